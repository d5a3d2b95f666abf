"""Writes golden fixtures from the CPU oracle (oracle/hif_oracle.cpp).

Usage: python tests/golden/make_golden.py   (minutes: the n=64 factor is large)

c2_n64.json: the n = 64 workload (N = 63 interior points per side, P = 4,
ID tolerance 1e-6, a = 10^s with s ~ U[-1, 1] from seed 1, symmetric
7-point stencil).  Records the per-level box and elimination counts, the
reference driver's e_a / e_s (v = 1 on the interior, w = A^T v), and the
preconditioner residual ||A F^{-1} b - b|| / ||b|| for b ~ N(0, 1) (seed 1).
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


def case(n, P, eps, coeff_seed, rhs_seed, threads=8):
    N = n - 1
    a = pb.pow10_coeff(n, coeff_seed)
    V = pb.zeros_coeff(n)
    r, c, v = hifo.stencil_coo(N, a, V, True)
    A = sp.csr_matrix((v, (r, c)), shape=(n ** 3, n ** 3))
    o = hifo.Oracle(N, P, eps)
    o.set_matrix_coo(r, c, v)
    t0 = time.time()
    info = o.factor(threads)
    tf = time.time() - t0
    vv = pb.reference_test_vector(n)
    w = A.T @ vv
    ea = float(np.linalg.norm(w - o.apply(vv, False)) / np.linalg.norm(w))
    es = float(np.linalg.norm(vv - o.apply(w, True)) / np.linalg.norm(vv))
    b = pb.rhs_normal(n, rhs_seed)
    res = float(np.linalg.norm(A @ o.apply(b, True) - b) / np.linalg.norm(b))
    return {"n": n, "P": P, "eps": eps, "coeff": "pow10 U[-1,1]", "coeff_seed": coeff_seed,
            "rhs_seed": rhs_seed, "symmetric": True, "info": int(info),
            "level_counts": [list(map(int, x)) for x in o.level_counts()],
            "e_a": ea, "e_s": es, "precond_residual": res, "oracle_factor_s": tf,
            "oracle_threads": threads}


if __name__ == "__main__":
    out = case(64, 4, 1e-6, 1, 1)
    with open(os.path.join(HERE, "c2_n64.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))
