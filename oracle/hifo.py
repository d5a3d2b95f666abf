"""ctypes wrapper of the CPU restatement in oracle/hif_oracle.cpp.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or as the timed CPU
baseline -- never by the product package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhif_oracle.so")
_lib = None

_i = ctypes.c_int
_l = ctypes.c_long
_d = ctypes.c_double
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.hifo_create.restype = _vp
        L.hifo_create.argtypes = [_i, _i, _d, _i]
        L.hifo_destroy.argtypes = [_vp]
        L.hifo_num_levels.argtypes = [_vp]
        L.hifo_set_matrix_coo.argtypes = [_vp, _l, _ip, _ip, _dp]
        L.hifo_factor.argtypes = [_vp, _i]
        L.hifo_apply.argtypes = [_vp, _dp, _i]
        L.hifo_num_boxes.argtypes = [_vp, _i, _i]
        L.hifo_box_sizes.argtypes = [_vp, _i, _i, _i, _ip]
        L.hifo_box_data.argtypes = [_vp, _i, _i, _i, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
        L.hifo_level_times.argtypes = [_vp, _i, _dp]
        L.hifo_face_rows.argtypes = [_vp, _i, _i, _ip]
        L.hifo_stencil_coo.restype = _l
        L.hifo_stencil_coo.argtypes = [_i, _d, _dp, _dp, _i, _ip, _ip, _dp]
        L.hifo_qrcp.argtypes = [_i, _i, _dp, _i, _ip, _dp]
        L.hifo_lu_inverse.argtypes = [_i, _dp]
        L.hifo_interp_decomp.argtypes = [_i, _i, _dp, _d, _ip, _ip, _dp, _ip]
        L.hifo_schur.argtypes = [_dp, _i, _ip, _i, _ip, _i, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def stencil_coo(N, a, V, symmetric=False):
    """Reference SetupStencil on nodal tensors a, V of shape (NC+1)^3 stored
    i-fastest (flattened in Fortran order).  Returns (rows, cols, vals)."""
    NC = N + 1
    a = np.asfortranarray(a, dtype=np.float64).ravel(order="F")
    V = np.asfortranarray(V, dtype=np.float64).ravel(order="F")
    assert a.size == (NC + 1) ** 3 and V.size == (NC + 1) ** 3
    cap = 7 * N ** 3
    rows = np.empty(cap, np.int32)
    cols = np.empty(cap, np.int32)
    vals = np.empty(cap, np.float64)
    nnz = lib().hifo_stencil_coo(N, 1.0 / NC, _p(a, ctypes.c_double), _p(V, ctypes.c_double),
                                 int(symmetric), _p(rows, ctypes.c_int), _p(cols, ctypes.c_int),
                                 _p(vals, ctypes.c_double))
    return rows[:nnz].copy(), cols[:nnz].copy(), vals[:nnz].copy()


class Oracle:
    """HIFFactor<double> restatement: Oracle(N, P, eps) with the reference's
    N = interior points per side (storage NC = N + 1 per side)."""

    def __init__(self, N, P, eps, front_quirk=True):
        self.N, self.P, self.eps = N, P, eps
        self.NC = N + 1
        self._h = lib().hifo_create(N, P, eps, int(front_quirk))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().hifo_destroy(self._h)
            self._h = None

    @property
    def num_levels(self):
        return lib().hifo_num_levels(self._h)

    def set_matrix_coo(self, rows, cols, vals):
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        vals = np.ascontiguousarray(vals, np.float64)
        lib().hifo_set_matrix_coo(self._h, rows.size, _p(rows, ctypes.c_int),
                                  _p(cols, ctypes.c_int), _p(vals, ctypes.c_double))

    def factor(self, nthreads=1):
        return lib().hifo_factor(self._h, nthreads)

    def apply(self, u, inverse=False):
        u = np.array(u, dtype=np.float64, copy=True)
        lib().hifo_apply(self._h, _p(u, ctypes.c_double), int(inverse))
        return u

    def num_boxes(self, level, kind):
        return lib().hifo_num_boxes(self._h, level, kind)

    def box(self, level, kind, b):
        sz = np.zeros(4, np.int32)
        assert lib().hifo_box_sizes(self._h, level, kind, b, _p(sz, ctypes.c_int)) == 0
        n, r, s, face = (int(x) for x in sz)
        g = np.zeros(n, np.int32)
        red = np.zeros(r, np.int32)
        skel = np.zeros(s, np.int32)
        A22 = np.zeros(r * r)
        A22inv = np.zeros(r * r)
        X = np.zeros(r * s)
        W = np.zeros(r * s)
        lib().hifo_box_data(self._h, level, kind, b, _p(g, ctypes.c_int), _p(red, ctypes.c_int),
                            _p(skel, ctypes.c_int), _p(A22, ctypes.c_double),
                            _p(A22inv, ctypes.c_double), _p(X, ctypes.c_double),
                            _p(W, ctypes.c_double))
        out = dict(g=g, red=red, skel=skel, face=face,
                   A22=A22.reshape(r, r, order="F"), A22inv=A22inv.reshape(r, r, order="F"),
                   X=X.reshape(r, s, order="F"))
        if kind == 1:
            out["W"] = W.reshape(s, r, order="F")
        return out

    def face_rows(self, level, b):
        """Global ids of the rows of face b's ID matrix at `level`, in the
        order SkelInteractionIndexData lists them (duplicates kept)."""
        n = lib().hifo_face_rows(self._h, level, b, None)
        assert n >= 0
        out = np.zeros(max(n, 1), np.int32)
        lib().hifo_face_rows(self._h, level, b, _p(out, ctypes.c_int))
        return out[:n]

    def level_counts(self):
        """Per level: (cells, cell DOFs eliminated, faces, face DOFs eliminated)."""
        out = []
        sz = np.zeros(4, np.int32)
        for lv in range(self.num_levels):
            row = []
            for kind in (0, 1):
                nb = self.num_boxes(lv, kind)
                red = 0
                for b in range(nb):
                    lib().hifo_box_sizes(self._h, lv, kind, b, _p(sz, ctypes.c_int))
                    red += int(sz[1])
                row += [nb, red]
            out.append(tuple(row))
        return out

    def level_times(self):
        t = np.zeros(3)
        out = []
        for lv in range(self.num_levels):
            lib().hifo_level_times(self._h, lv, _p(t, ctypes.c_double))
            out.append(tuple(t))
        return out


def qrcp(A):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    jpvt = np.zeros(n, np.int32)
    tau = np.zeros(n)
    lib().hifo_qrcp(m, n, _p(A, ctypes.c_double), max(m, 1), _p(jpvt, ctypes.c_int),
                    _p(tau, ctypes.c_double))
    return A, jpvt, tau


def lu_inverse(A):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    info = lib().hifo_lu_inverse(A.shape[0], _p(A, ctypes.c_double))
    return A, info


def interp_decomp(M, eps):
    M = np.array(M, dtype=np.float64, order="F", copy=True)
    m, n = M.shape
    skel = np.zeros(n, np.int32)
    red = np.zeros(n, np.int32)
    W = np.zeros(n * n)
    cnt = np.zeros(2, np.int32)
    lib().hifo_interp_decomp(m, n, _p(M, ctypes.c_double), eps, _p(skel, ctypes.c_int),
                             _p(red, ctypes.c_int), _p(W, ctypes.c_double),
                             _p(cnt, ctypes.c_int))
    k, q = int(cnt[0]), int(cnt[1])
    return skel[:k].copy(), red[:q].copy(), W[:k * q].reshape(k, q, order="F").copy(), M


def schur(M, red, skel):
    M = np.array(M, dtype=np.float64, order="F", copy=True)
    n = M.shape[0]
    red = np.ascontiguousarray(red, np.int32)
    skel = np.ascontiguousarray(skel, np.int32)
    r, s = red.size, skel.size
    A22inv = np.zeros(r * r)
    X = np.zeros(r * s)
    S = np.zeros(s * s)
    info = lib().hifo_schur(_p(M, ctypes.c_double), n, _p(red, ctypes.c_int), r,
                            _p(skel, ctypes.c_int), s, _p(A22inv, ctypes.c_double),
                            _p(X, ctypes.c_double), _p(S, ctypes.c_double))
    return (A22inv.reshape(r, r, order="F"), X.reshape(r, s, order="F"),
            S.reshape(s, s, order="F"), info)
