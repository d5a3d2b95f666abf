"""Synthetic problem generators for the workloads (coefficients and right-hand sides).

The reference fills its nodal tensors by hand with a = 1, V = 0
(tests/HIFDE3D/main.cpp:59-74) and uses v = 1 on the interior as its test
vector (:88-99).  The variable-coefficient workloads use this module's own
generators, all driven by SplitMix64 so that the same seed gives the same
numbers on every machine:

  * uniform(seed, n): SplitMix64 -> 53-bit uniform in [0, 1)
  * normal(seed, n): Box-Muller on pairs of those uniforms
  * pow10_coeff(NC, seed, lo, hi): a = 10^s, s ~ U[lo, hi] on the (NC+1)^3 nodes
  * lognormal_coeff(NC, seed, corr): a = exp(g), g a unit-variance Gaussian
    random field with Gaussian covariance exp(-|x-y|^2 / (2 corr^2)),
    synthesised spectrally on the (NC+1)^3 nodes (periodic FFT on a grid padded
    to twice the size so the wrap-around does not correlate opposite faces).

Tensors are returned as numpy arrays of shape (NC+1,)*3 indexed [i, j, k],
to be flattened in Fortran order (i fastest) as include/numtns.hpp:99 does.
"""
import numpy as np

_MASK = (1 << 64) - 1


def splitmix64(seed, n):
    """n successive SplitMix64 outputs (uint64) from the given seed."""
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _MASK) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n):
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal(seed, n):
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = rad * np.cos(2 * np.pi * u2)
    z[1::2] = rad * np.sin(2 * np.pi * u2)
    return z[:n]


def ones_coeff(NC):
    return np.ones((NC + 1,) * 3)


def zeros_coeff(NC):
    return np.zeros((NC + 1,) * 3)


def pow10_coeff(NC, seed, lo=-1.0, hi=1.0):
    m = NC + 1
    s = lo + (hi - lo) * uniform(seed, m ** 3)
    return (10.0 ** s).reshape((m, m, m), order="F")


def uniform_coeff(NC, seed, lo=0.0, hi=1.0):
    m = NC + 1
    return (lo + (hi - lo) * uniform(seed, m ** 3)).reshape((m, m, m), order="F")


def gaussian_field(NC, seed, corr=0.1):
    """Unit-variance Gaussian random field on the (NC+1)^3 nodes of the unit cube."""
    m = NC + 1
    h = 1.0 / NC
    P = 2 * m
    white = normal(seed, P ** 3).reshape((P, P, P), order="F")
    k = np.fft.fftfreq(P, d=h) * 2 * np.pi
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    # spectral density of exp(-r^2/(2 l^2)) in 3-D (up to a constant)
    spec = np.exp(-0.5 * corr ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    g = np.real(np.fft.ifftn(np.fft.fftn(white) * np.sqrt(spec)))[:m, :m, :m]
    g = g - g.mean()
    g = g / g.std()
    return g


def lognormal_coeff(NC, seed, corr=0.1):
    return np.exp(gaussian_field(NC, seed, corr))


def interior_mask(NC):
    idx = np.arange(NC ** 3)
    i, j, k = idx % NC, (idx // NC) % NC, idx // (NC * NC)
    return (i > 0) & (j > 0) & (k > 0)


def rhs_normal(NC, seed, nrhs=1):
    """iid N(0,1) on the interior DOFs (filled in linear-index order), 0 on the boundary."""
    mask = interior_mask(NC)
    nint = int(mask.sum())
    out = np.zeros((nrhs, NC ** 3))
    z = normal(seed, nint * nrhs).reshape(nrhs, nint)
    out[:, mask] = z
    return out[0] if nrhs == 1 else out


def reference_test_vector(NC):
    """v = 1 on the interior, 0 on the boundary (tests/HIFDE3D/main.cpp:88-99)."""
    return interior_mask(NC).astype(np.float64)


# Batched uneven instances: grid sizes drawn uniformly from C4_SIZES, each with
# the leaf width that makes NC/P a power of two, a = 10^s with s ~ U[-2, 2],
# 16 right-hand sides N(0, 1).
C4_SIZES = (24, 32, 40, 48, 64, 80, 96)
C4_LEAF = {24: 3, 48: 3, 96: 3, 32: 4, 64: 4, 40: 5, 80: 5}


def batch_instances(seed=3, count=48):
    """The independent problems of the batched workload: per problem NC, P and
    the seeds of its coefficient (s ~ U[-2, 2]) and right-hand sides."""
    u = uniform(seed, count)
    out = []
    for i in range(count):
        NC = C4_SIZES[min(int(u[i] * len(C4_SIZES)), len(C4_SIZES) - 1)]
        out.append({"NC": NC, "P": C4_LEAF[NC], "coeff_seed": seed * 1000 + 2 * i,
                    "rhs_seed": seed * 1000 + 2 * i + 1})
    return out
