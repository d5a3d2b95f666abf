"""Writes tests/golden/c3_n<n>.json from the CPU oracle: the headline
workload family (lognormal coefficient a = exp(g), g the unit-variance
Gaussian random field of correlation length 0.1 from seed 2; symmetric
7-point stencil; P = 4; ID tolerance 1e-4) at grid n.

Records the per-level counts, the preconditioner residual
||A F^{-1} b - b|| / ||b|| for b = rhs_normal(n, 2), the reference driver's
e_a / e_s, and the oracle-preconditioned CG to 1e-10 (iterations, final
residual, the solution's 2-norm and its values at 64 fixed DOFs).

Usage: python tests/golden/make_c3.py [n] [threads]   (default n = 64, 8 threads)
"""
import json
import os
import sys
import time

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from hifde3d_amd import problems as pb  # noqa: E402
from oracle import hifo  # noqa: E402


def sample_dofs(n):
    """64 interior DOFs spread over the grid (fixed by n)."""
    rng = np.random.default_rng(12345)
    idx = rng.integers(1, n, size=(64, 3))
    return [int(i + n * j + n * n * k) for i, j, k in idx]


def pcg(A, b, prec, tol=1e-10, maxit=300):
    x = np.zeros_like(b)
    r = b.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    it = 0
    while np.linalg.norm(r) / bn > tol and it < maxit:
        it += 1
        q = A @ p
        al = rz / (p @ q)
        x += al * p
        r -= al * q
        if np.linalg.norm(r) / bn <= tol:
            break
        z = prec(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, it, float(np.linalg.norm(b - A @ x) / bn)


def case(n, threads):
    N, P, eps = n - 1, 4, 1e-4
    a = pb.lognormal_coeff(n, 2, 0.1)
    V = pb.zeros_coeff(n)
    r, c, v = hifo.stencil_coo(N, a, V, True)
    A = sp.csr_matrix((v, (r, c)), shape=(n ** 3, n ** 3))
    o = hifo.Oracle(N, P, eps)
    o.set_matrix_coo(r, c, v)
    del r, c, v
    t0 = time.time()
    info = o.factor(threads)
    tf = time.time() - t0
    vv = pb.reference_test_vector(n)
    w = A.T @ vv
    ea = float(np.linalg.norm(w - o.apply(vv, False)) / np.linalg.norm(w))
    es = float(np.linalg.norm(vv - o.apply(w, True)) / np.linalg.norm(vv))
    b = pb.rhs_normal(n, 2)
    res = float(np.linalg.norm(A @ o.apply(b, True) - b) / np.linalg.norm(b))
    t0 = time.time()
    x, it, rr = pcg(A, b, lambda y: o.apply(y, True))
    ts = time.time() - t0
    idx = sample_dofs(n)
    return {"n": n, "P": P, "eps": eps, "coeff": "lognormal GRF corr 0.1", "coeff_seed": 2,
            "rhs_seed": 2, "symmetric": True, "info": int(info),
            "level_counts": [list(map(int, c_)) for c_ in o.level_counts()],
            "e_a": ea, "e_s": es, "precond_residual": res,
            "pcg_iters": int(it), "pcg_relres": rr, "x_norm": float(np.linalg.norm(x)),
            "x_sample_dofs": idx, "x_sample": [float(x[i]) for i in idx],
            "oracle_factor_s": tf, "oracle_pcg_s": ts, "oracle_threads": threads}


if __name__ == "__main__":
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 64
    th = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    out = case(n, th)
    with open(os.path.join(HERE, f"c3_n{n}.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k not in ("x_sample", "x_sample_dofs")}))
