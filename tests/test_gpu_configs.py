"""The headline configs run at their own sizes in the GPU suite.

* n=128 lognormal (the measured workload): the C3-family oracle fixture at
  n=64 (tests/golden/c3_n64.json, written by make_c3.py) pins counts,
  preconditioner quality and the solution; at n=128 the workload itself is
  checked through size-independent properties: PCG to 1e-10 with the true
  residual recomputed on the CPU from the assembled operator, the mixed and
  FP64 factorizations giving the same level counts (within 1%) and solutions
  within 1e-8 of each other, and -- when the full-size oracle fixture
  c3_n128.json is present -- counts within 5% of the oracle's.
* NC=256, P=8, eps 1e-3 with a reaction term and 64 right-hand sides (the
  large-leaf config), at full size: exact level-0 sizes, every column of the
  block PCG to a true residual of 2e-10.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from hifde3d_amd import problems as pb

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def H():
    import hifde3d_amd
    hifde3d_amd.lib()
    return hifde3d_amd


def _golden(name):
    path = os.path.join(HERE, "golden", name)
    return json.load(open(path)) if os.path.exists(path) else None


def _csr(f):
    rp, ci, v = f.operator_csr()
    return sp.csr_matrix((v, ci, rp), shape=(f.size, f.size))


def _counts_close(cg, co, rel):
    for lv, c in enumerate(co):
        assert cg[lv][0] == c[0] and cg[lv][2] == c[2], lv
        for idx in (1, 3):
            assert abs(cg[lv][idx] - c[idx]) <= rel * max(c[idx], 20), (lv, idx, cg[lv], c)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c3_family_n64_against_oracle_golden(H, precision):
    """Lognormal GRF (seed 2), n=64, P=4, eps 1e-4, symmetric stencil: level
    counts within 1% of the oracle (measured equal, Gram and Householder
    paths alike; the bound leaves room for a pivot decided by rounding at the
    tolerance); FP64 preconditioner residual
    ||A F^{-1} b - b|| / ||b|| within 2x (mixed 4x); PCG to 1e-10 in at most
    the oracle-preconditioned CG's iterations + 3, solution within 1e-5 of
    the oracle's largest sampled value at 64 fixed DOFs."""
    import torch
    g = _golden("c3_n64.json")
    n, P, eps = g["n"], g["P"], g["eps"]
    f = H.HIFFactor(n - 1, P, eps, precision=precision)
    f.set_stencil(pb.lognormal_coeff(n, g["coeff_seed"]), pb.zeros_coeff(n), True)
    assert f.Factor() == 0
    _counts_close(f.level_counts(), g["level_counts"], 0.01)
    b = pb.rhs_normal(n, g["rhs_seed"])
    res = np.linalg.norm(f.spmv(f.Apply(b, True)) - b) / np.linalg.norm(b)
    assert res <= (2.0 if precision == "fp64" else 4.0) * g["precond_residual"]
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    it, rr = f.pcg_device(db.data_ptr(), dx.data_ptr(), tol=1e-10, maxit=200)
    assert rr <= 1e-10 and it <= g["pcg_iters"] + 3
    x = dx.cpu().numpy()
    A = _csr(f)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 2e-10
    xs = np.array(g["x_sample"])
    assert np.abs(x[g["x_sample_dofs"]] - xs).max() <= 1e-5 * np.abs(xs).max()
    f.close()


def test_fp32_elements_match_fp64_elements(H):
    """With every level in FP32 (eps >= 1e-5) the level-matrix elements are
    FP32 too; option elements_fp32 = 0 keeps them FP64.  Lognormal GRF,
    n=64, eps 1e-4: the same level counts within 1%, preconditioner
    residuals within 2x of each other (they differ through skeleton choices
    at the tolerance: 3.6e-3 and 5.5e-3 here), PCG iterations within one,
    and solutions within 1e-8."""
    import torch
    n = 64
    a, V, b = pb.lognormal_coeff(n, 2), pb.zeros_coeff(n), pb.rhs_normal(n, 2)
    db = torch.from_numpy(b).cuda()
    out = {}
    for opt in (1, 0):
        f = H.HIFFactor(n - 1, 4, 1e-4, precision="mixed")
        f.set_option("elements_fp32", opt)
        f.set_stencil(a, V, True)
        assert f.Factor() == 0
        res = np.linalg.norm(f.spmv(f.Apply(b, True)) - b) / np.linalg.norm(b)
        dx = torch.zeros_like(db)
        it, rr = f.pcg_device(db.data_ptr(), dx.data_ptr(), tol=1e-10, maxit=200)
        assert rr <= 1e-10
        out[opt] = (f.level_counts(), res, it, dx.cpu().numpy())
        f.close()
    (c1, r1, i1, x1), (c0, r0, i0, x0) = out[1], out[0]
    for lv in range(len(c0)):
        assert c1[lv][:3:2] == c0[lv][:3:2]
        for idx in (1, 3, 4):
            assert abs(c1[lv][idx] - c0[lv][idx]) <= 0.01 * max(c0[lv][idx], 100)
    assert r1 <= 2 * r0 and r0 <= 2 * r1
    assert abs(i1 - i0) <= 1
    assert np.linalg.norm(x1 - x0) / np.linalg.norm(x0) < 1e-8


def test_c3_full_size_workload_properties(H):
    """The measured workload itself: n=128 (127^3 unknowns), lognormal GRF
    seed 2, P=4, eps 1e-4; mixed factor + FP64 PCG to 1e-10, and the FP64
    factor + PCG of the same problem."""
    import torch
    n = 128
    a = pb.lognormal_coeff(n, 2)
    V = pb.zeros_coeff(n)
    b = pb.rhs_normal(n, 2)
    db = torch.from_numpy(b).cuda()
    out = {}
    for precision in ("mixed", "fp64"):
        f = H.HIFFactor(n - 1, 4, 1e-4, precision=precision)
        f.set_stencil(a, V, True)
        assert f.Factor() == 0
        dx = torch.zeros_like(db)
        it, rr = f.pcg_device(db.data_ptr(), dx.data_ptr(), tol=1e-10, maxit=200)
        assert rr <= 1e-10 and it <= 15, (precision, it)
        out[precision] = (f.level_counts(), dx.cpu().numpy())
        if precision == "mixed":
            A = _csr(f)
        f.close()
    for precision, (_, x) in out.items():
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 2e-10, precision
    cm, cf = out["mixed"][0], out["fp64"][0]
    for lv in range(len(cf)):
        assert cm[lv][:3:2] == cf[lv][:3:2]
        for idx in (1, 3, 4):
            assert abs(cm[lv][idx] - cf[lv][idx]) <= 0.01 * max(cf[lv][idx], 100)
    xm, x64 = out["mixed"][1], out["fp64"][1]
    assert np.linalg.norm(xm - x64) / np.linalg.norm(x64) < 1e-8
    g = _golden("c3_n128.json")
    if g is not None:
        _counts_close(cf, g["level_counts"], 0.05)
        _counts_close(cm, g["level_counts"], 0.05)


def test_large_leaf_config_full_size_block_pcg(H):
    """NC=256 (255^3 = 16.6M unknowns), P=8, eps 1e-3, lognormal a (seed 4),
    V ~ U[0,1] (seed 40), generated and assembled on the GPU; mixed factor
    with the workspaces released, then 64 right-hand sides by block PCG to
    1e-10.  Level 0 has 32^3 cells of 343 interior DOFs; every column's true
    residual (the assembled operator applied on the GPU) is below 2e-10."""
    import torch
    NC, P, eps, R = 256, 8, 1e-3, 64
    a, V = H.device_problem(NC, "lognormal", 4, V="uniform", v_seed=40)
    f = H.HIFFactor(NC - 1, P, eps, precision="mixed")
    f.set_stencil_device(a.data_ptr(), V.data_ptr(), True)
    del a, V
    f.set_option("keep_workspaces", 0)
    assert f.Factor() == 0
    c0 = f.level_counts()[0]
    assert c0[0] == 32 ** 3 and c0[1] == 32 ** 3 * 343
    B = torch.empty((R, NC ** 3), dtype=torch.float64, device="cuda")
    H.gen_rhs_normal(B.data_ptr(), NC, 41, nrhs=R)
    X = torch.zeros_like(B)
    its, rr = f.pcg_multi_device(B.data_ptr(), X.data_ptr(), R, tol=1e-10, maxit=300)
    assert (rr <= 1e-10).all() and its.max() < 80
    rp, ci, v = f.operator_csr()
    A = torch.sparse_csr_tensor(torch.from_numpy(rp.astype(np.int64)),
                                torch.from_numpy(ci.astype(np.int64)), torch.from_numpy(v),
                                size=(NC ** 3, NC ** 3)).cuda()
    for c in range(R):
        r = B[c] - A @ X[c]
        assert float(torch.linalg.norm(r) / torch.linalg.norm(B[c])) < 2e-10, c
    f.close()


def test_batched_config_packed_shared_launches(H):
    """The batched config through packed factor objects (up to eight problems
    of one grid size per object, every level's boxes in shared launches): all
    48 problems, 16 right-hand sides each, every (problem, column) PCG to
    1e-10 and its true residual -- against the problem's own operator,
    assembled by a standalone object -- below 2e-10."""
    import torch
    inst = pb.batch_instances()
    R = 16
    by_nc = {}
    for p in inst:
        by_nc.setdefault(p["NC"], []).append(p)
    done = 0
    for NC, ps in sorted(by_nc.items()):
        for c0 in range(0, len(ps), 8):
            chunk = ps[c0:c0 + 8]
            coeffs = [H.device_problem(NC, "pow10", p["coeff_seed"], lo=-2.0, hi=2.0)
                      for p in chunk]
            f = H.HIFBatch(NC - 1, chunk[0]["P"], 1e-4, len(chunk), precision="mixed")
            f.set_stencil_batch_device([a.data_ptr() for a, _ in coeffs],
                                       [V.data_ptr() for _, V in coeffs], True)
            assert f.Factor() == 0
            Bs = torch.empty((len(chunk) * R, NC ** 3), dtype=torch.float64, device="cuda")
            for o, p in enumerate(chunk):
                H.gen_rhs_normal(Bs[o * R].data_ptr(), NC, p["rhs_seed"], nrhs=R)
            Bp = torch.empty((R, f.size), dtype=torch.float64, device="cuda")
            f.pack(Bs.data_ptr(), Bp.data_ptr(), R)
            Xp = torch.zeros_like(Bp)
            its, rr = f.pcg_multi_device(Bp.data_ptr(), Xp.data_ptr(), R, tol=1e-10, maxit=300)
            assert (rr <= 1e-10).all() and its.max() < 100, (NC, rr.max(), its.max())
            Xs = torch.empty_like(Bs)
            f.unpack(Xp.data_ptr(), Xs.data_ptr(), R)
            f.close()
            for o, (a, V) in enumerate(coeffs):
                g = H.HIFFactor(NC - 1, chunk[o]["P"], 1e-4)
                g.set_stencil_device(a.data_ptr(), V.data_ptr(), True)
                A = _csr(g)
                g.close()
                Bo = Bs[o * R:(o + 1) * R].cpu().numpy()
                Xo = Xs[o * R:(o + 1) * R].cpu().numpy()
                for c in range(R):
                    res = np.linalg.norm(A @ Xo[c] - Bo[c]) / np.linalg.norm(Bo[c])
                    assert res < 2e-10, (NC, o, c, res)
                done += 1
    assert done == 48
