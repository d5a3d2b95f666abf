"""Device problem assembly and generators (assemble.hip) against the host
assembly, the oracle's SetupStencil restatement and problems.py.

The stencil's CSR must be bitwise that of the host path (same rows, entry
order and values), so a factorization of a device-assembled operator is
bitwise that of the host-assembled one.  Uniforms are bitwise; values that go
through the device's log / sin / cos / pow / exp are within a few ulps of the
host's; the Gaussian field (direct separable filtering instead of an FFT) is
within 1e-10 absolute of problems.gaussian_field on unit-variance values.
"""
import json
import os

import numpy as np
import pytest

from hifde3d_amd import problems as pb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hifde3d_amd
    hifde3d_amd.lib()
    return hifde3d_amd


def _dev(x):
    import torch
    return torch.from_numpy(np.asfortranarray(x).ravel(order="F").copy()).cuda()


def _golden():
    here = os.path.dirname(os.path.abspath(__file__))
    return json.load(open(os.path.join(here, "golden", "generators.json")))


@pytest.mark.parametrize("sym", [False, True])
def test_device_stencil_equals_host_and_oracle_assembly(H, oracle_mod, sym):
    import torch
    N = 23
    NC = N + 1
    a, V = pb.pow10_coeff(NC, 5), pb.uniform_coeff(NC, 6)
    fh = H.HIFFactor(N, 3, 1e-4)
    fh.set_stencil(a, V, sym)
    fd = H.HIFFactor(N, 3, 1e-4)
    da, dV = _dev(a), _dev(V)
    torch.cuda.synchronize()
    fd.set_stencil_device(da.data_ptr(), dV.data_ptr(), sym)
    rp_h, ci_h, v_h = fh.operator_csr()
    rp_d, ci_d, v_d = fd.operator_csr()
    assert np.array_equal(rp_d, rp_h) and np.array_equal(ci_d, ci_h)
    assert np.array_equal(v_d.view(np.int64), v_h.view(np.int64))
    # the oracle's SetupStencil restatement (rows visited i, j, k with k
    # fastest, as the reference loops): the same entries, each row's in the
    # same order
    r, c, v = oracle_mod.stencil_coo(N, a, V, sym)
    o = np.argsort(r, kind="stable")
    rows = np.repeat(np.arange(NC ** 3), np.diff(rp_d))
    assert np.array_equal(rows, r[o]) and np.array_equal(ci_d, c[o])
    assert np.array_equal(v_d.view(np.int64), v[o].view(np.int64))
    # the reference-form SpMV (A^T x) of the device-assembled object
    x = np.random.default_rng(4).standard_normal(NC ** 3)
    assert np.array_equal(fd.spmv(x, transpose=True), fh.spmv(x, transpose=True))
    fh.close()
    fd.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_assembled_factor_is_bitwise_the_host_one(H, precision):
    import torch
    N, P = 31, 4
    NC = N + 1
    a, V = pb.lognormal_coeff(NC, 2), pb.zeros_coeff(NC)
    b = pb.rhs_normal(NC, 2)
    fh = H.HIFFactor(N, P, 1e-4, precision=precision)
    fh.set_stencil(a, V, True)
    fd = H.HIFFactor(N, P, 1e-4, precision=precision)
    da, dV = _dev(a), _dev(V)
    torch.cuda.synchronize()
    fd.set_stencil_device(da.data_ptr(), dV.data_ptr(), True)
    assert fh.Factor() == 0 and fd.Factor() == 0
    assert fh.level_counts() == fd.level_counts()
    assert np.array_equal(fh.Apply(b, True), fd.Apply(b, True))
    fh.close()
    fd.close()


def _ulps(x, y):
    xi = x.view(np.int64).astype(np.float64)
    yi = y.view(np.int64).astype(np.float64)
    return np.abs(xi - yi)


def test_device_generators_match_problems_and_pins(H):
    import torch
    g = _golden()
    n = 1 << 16
    u = torch.empty(n, dtype=torch.float64, device="cuda")
    H.gen_uniform(u.data_ptr(), n, 1)
    uh = u.cpu().numpy()
    assert np.array_equal(uh, pb.uniform(1, n))
    assert uh[:16].tolist() == g["uniform_seed1"]
    H.gen_uniform(u.data_ptr(), n, 9, -2.0, 2.0)
    assert np.array_equal(u.cpu().numpy(), -2.0 + 4.0 * pb.uniform(9, n))
    # 10^U: the device pow within a few ulps of the host's
    m = 9
    p10 = torch.empty(m ** 3, dtype=torch.float64, device="cuda")
    H.gen_uniform(p10.data_ptr(), m ** 3, 1, -1.0, 1.0, pow10=True)
    ph = p10.cpu().numpy()
    ref = pb.pow10_coeff(8, 1).ravel(order="F")
    assert _ulps(ph, ref).max() <= 4
    assert _ulps(ph[:16], np.array(g["pow10_NC8_seed1_first16"])).max() <= 4
    # Box-Muller normals
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    H.gen_normal(z.data_ptr(), n, 2)
    zh = z.cpu().numpy()
    assert np.allclose(zh, pb.normal(2, n), rtol=1e-14, atol=1e-14)
    assert np.allclose(zh[:16], g["normal_seed2"], rtol=1e-14, atol=1e-14)
    # right-hand sides: interior normals in linear order, zero boundary
    NC, R = 12, 3
    B = torch.full((R, NC ** 3), 7.0, dtype=torch.float64, device="cuda")
    H.gen_rhs_normal(B.data_ptr(), NC, 5, nrhs=R)
    Bh = B.cpu().numpy()
    Bp = pb.rhs_normal(NC, 5, R)
    assert np.allclose(Bh, Bp, rtol=1e-14, atol=1e-14)
    assert np.all(Bh[:, ~pb.interior_mask(NC)] == 0.0)
    r8 = torch.empty(8 ** 3, dtype=torch.float64, device="cuda")
    H.gen_rhs_normal(r8.data_ptr(), 8, 0)
    assert np.allclose(r8.cpu().numpy()[pb.interior_mask(8)][:16],
                       g["rhs_normal_NC8_seed0_interior_first16"], rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("NC,seed", [(16, 2), (32, 4)])
def test_device_gaussian_field_matches_spectral_synthesis(H, NC, seed):
    import torch
    m3 = (NC + 1) ** 3
    g = torch.empty(m3, dtype=torch.float64, device="cuda")
    H.gen_gaussian_field(g.data_ptr(), NC, seed, 0.1)
    gd = g.cpu().numpy()
    ref = pb.gaussian_field(NC, seed, 0.1).ravel(order="F")
    assert np.abs(gd - ref).max() < 1e-10
    assert abs(gd.mean()) < 1e-12 and abs(gd.std() - 1.0) < 1e-12
    if NC == 16 and seed == 2:
        assert np.abs(gd[:16] - _golden()["gaussian_field_NC16_seed2_corr0.1_first16"]).max() < 1e-10
    H.gen_gaussian_field(g.data_ptr(), NC, seed, 0.1, lognormal=True)
    assert np.allclose(g.cpu().numpy(), np.exp(ref), rtol=1e-10, atol=0)


def test_device_problem_end_to_end_c3_family_n64(H):
    """The n=128 headline problem family at n=64, generated and assembled on
    the GPU (lognormal field seed 2, symmetric stencil, P=4, eps 1e-4): the
    factorization and PCG agree with the host-generated problem (same level
    counts within 1%, PCG to 1e-10 with a true residual checked on the CPU)."""
    import scipy.sparse as sp
    import torch
    NC = 64
    a, V = H.device_problem(NC, "lognormal", 2)
    f = H.HIFFactor(NC - 1, 4, 1e-4, precision="mixed")
    f.set_stencil_device(a.data_ptr(), V.data_ptr(), True)
    assert f.Factor() == 0
    fh = H.HIFFactor(NC - 1, 4, 1e-4, precision="mixed")
    fh.set_stencil(pb.lognormal_coeff(NC, 2), pb.zeros_coeff(NC), True)
    assert fh.Factor() == 0
    for cd, ch in zip(f.level_counts(), fh.level_counts()):
        assert abs(cd[4] - ch[4]) <= 0.01 * max(ch[4], 1) + 2
    b = torch.empty(NC ** 3, dtype=torch.float64, device="cuda")
    H.gen_rhs_normal(b.data_ptr(), NC, 2)
    x = torch.zeros_like(b)
    it, rr = f.pcg_device(b.data_ptr(), x.data_ptr(), tol=1e-10, maxit=200)
    assert rr <= 1e-10 and it < 30
    rp, ci, v = f.operator_csr()
    A = sp.csr_matrix((v, ci, rp), shape=(NC ** 3,) * 2)
    bh, xh = b.cpu().numpy(), x.cpu().numpy()
    assert np.linalg.norm(A @ xh - bh) / np.linalg.norm(bh) < 2e-10
    f.close()
    fh.close()


def _batch_setup(H, NC, seeds, precision="mixed", eps=1e-4):
    import torch
    P = pb.C4_LEAF.get(NC, 4)
    coeffs = [H.device_problem(NC, "pow10", s_, lo=-2.0, hi=2.0) for s_ in seeds]
    bt = H.HIFBatch(NC - 1, P, eps, len(seeds), precision=precision)
    bt.set_stencil_batch_device([a.data_ptr() for a, _ in coeffs],
                                [V.data_ptr() for _, V in coeffs], True)
    torch.cuda.synchronize()
    return bt, coeffs


@pytest.mark.parametrize("nprob", [3, 8])
def test_packed_batch_solves_each_problem(H, nprob):
    """nprob problems of one grid (NC=32, P=4; the rest of the 8 octants
    padding) in one packed object: one factorization, one block of PCG
    solves with 4 right-hand sides per problem, each (problem, column) its
    own PCG.  Every solution's true residual against that problem's own
    operator (assembled by a standalone object) is below 2e-10, and agrees
    with the standalone object's solution to 1e-8; pack and unpack are
    inverse on the problems' unknowns."""
    import scipy.sparse as sp
    import torch
    NC, R = 32, 4
    seeds = [300 + s_ for s_ in range(nprob)]
    bt, coeffs = _batch_setup(H, NC, seeds)
    assert bt.Factor() == 0
    n3 = NC ** 3
    Bs = torch.empty((nprob * R, n3), dtype=torch.float64, device="cuda")
    for o in range(nprob):
        H.gen_rhs_normal(Bs[o * R].data_ptr(), NC, 400 + o, nrhs=R)
    Bp = torch.empty((R, bt.size), dtype=torch.float64, device="cuda")
    bt.pack(Bs.data_ptr(), Bp.data_ptr(), R)
    back = torch.full_like(Bs, 9.0)
    bt.unpack(Bp.data_ptr(), back.data_ptr(), R)
    assert torch.equal(back, Bs)
    Xp = torch.zeros_like(Bp)
    its, rr = bt.pcg_multi_device(Bp.data_ptr(), Xp.data_ptr(), R, tol=1e-10, maxit=200)
    assert its.shape == (nprob, R) and (rr <= 1e-10).all() and its.max() < 60
    Xs = torch.empty_like(Bs)
    bt.unpack(Xp.data_ptr(), Xs.data_ptr(), R)
    for o, (a, V) in enumerate(coeffs):
        f = H.HIFFactor(NC - 1, pb.C4_LEAF[NC], 1e-4, precision="mixed")
        f.set_stencil_device(a.data_ptr(), V.data_ptr(), True)
        assert f.Factor() == 0
        rp, ci, v = f.operator_csr()
        A = sp.csr_matrix((v, ci, rp), shape=(n3, n3))
        Bo = Bs[o * R:(o + 1) * R].contiguous()
        Xo = torch.zeros_like(Bo)
        f.pcg_multi_device(Bo.data_ptr(), Xo.data_ptr(), R, tol=1e-10, maxit=200)
        for c in range(R):
            b = Bo[c].cpu().numpy()
            x = Xs[o * R + c].cpu().numpy()
            assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 2e-10, (o, c)
            xr = Xo[c].cpu().numpy()
            assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-8, (o, c)
        f.close()
    bt.close()


def test_packed_batch_exact_factor_applies_each_operator(H):
    """FP64 with eps = 1e-14 (no truncation): the packed factorization's
    forward apply equals each problem's operator on its octant, and identity
    on the separating planes."""
    import scipy.sparse as sp
    import torch
    NC = 16
    bt, coeffs = _batch_setup(H, NC, [5, 6], precision="fp64", eps=1e-14)
    assert bt.Factor() == 0
    n3 = NC ** 3
    U = torch.empty((2, n3), dtype=torch.float64, device="cuda")
    H.gen_rhs_normal(U.data_ptr(), NC, 77, nrhs=2)
    Up = torch.empty(bt.size, dtype=torch.float64, device="cuda")
    # two problems, one vector each: vector c of problem o at U[o]
    bt.pack(U.data_ptr(), Up.data_ptr(), 1)
    Fu = torch.from_numpy(bt.Apply(Up.cpu().numpy(), False)).cuda()
    back = torch.empty_like(U)
    bt.unpack(Fu.data_ptr(), back.data_ptr(), 1)
    for o, (a, V) in enumerate(coeffs):
        f = H.HIFFactor(NC - 1, 4, 1e-4)
        f.set_stencil_device(a.data_ptr(), V.data_ptr(), True)
        rp, ci, v = f.operator_csr()
        A = sp.csr_matrix((v, ci, rp), shape=(n3, n3))
        ref = A @ U[o].cpu().numpy()
        got = back[o].cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-11, o
        f.close()
    bt.close()
