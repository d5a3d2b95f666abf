"""MI355X (gfx950) HIF-DE factorization: Python binding of libhifde_amd.so.

The class mirrors the reference's HIFFactor<Scalar> interface
(include/Factor.hpp:8-57): construct with (N, P, epsilon), fill the matrix
(sp_matrix() / SetupStencil), Factor(), then Apply(u, inverse).  Every call
goes through the C ABI declared in include/hifde_amd.h; there is no CPU
fallback -- importing on a machine without the built library raises.
"""
import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhifde_amd.so")

FP64 = 0
MIXED = 1
FLAG_FRONT_QUIRK = 1

_lib = None

_i = ctypes.c_int
_ll = ctypes.c_longlong
_d = ctypes.c_double
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)

EXPORTS = [
    "hifde_create", "hifde_destroy", "hifde_last_error", "hifde_set_matrix_coo",
    "hifde_set_stencil", "hifde_factor_run", "hifde_apply", "hifde_apply_device", "hifde_spmv",
    "hifde_pcg", "hifde_pcg_device", "hifde_pcg_multi_device", "hifde_num_levels",
    "hifde_level_counts",
    "hifde_box_count", "hifde_box_sizes", "hifde_box_data", "hifde_stage_times",
    "hifde_level_stage_times",
    "hifde_stage_name", "hifde_factor_bytes", "hifde_interp_decomp_batch", "hifde_schur_batch",
    "hifde_set_profiling", "hifde_set_profile_kernel", "hifde_reset_kernel_stats",
    "hifde_kernel_stats", "hifde_set_option",
    "hifde_gemm_batch", "hifde_factor_save", "hifde_factor_load", "hifde_set_stream",
    "hifde_apply_device_noalloc", "hifde_pivoted_qr_batch", "hifde_triangular_solve_batch",
    "hifde_set_stencil_device", "hifde_gen_uniform", "hifde_gen_normal", "hifde_gen_rhs_normal",
    "hifde_gen_gaussian_field", "hifde_operator_csr", "hifde_gemm_table_load",
    "hifde_gemm_table_clear", "hifde_gemm_log", "hifde_gemm_time", "hifde_create_batch",
    "hifde_set_stencil_batch_device", "hifde_batch_pack",
]


class HifdeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hifde error {code}: {msg}")
        self.code = code


def lib():
    """Load libhifde_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built; run `python -m hifde3d_amd.build`")
    # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7
    # under the same soname, whichever is loaded first serves both, and
    # PyTorch finds no GPU on the system copy.  Device buffers are exchanged
    # with PyTorch, so its runtime is loaded first when it is installed.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    L.hifde_create.argtypes = [_i, _i, _d, _i, _i, ctypes.POINTER(_vp)]
    L.hifde_destroy.argtypes = [_vp]
    L.hifde_destroy.restype = None
    L.hifde_last_error.restype = ctypes.c_char_p
    L.hifde_set_matrix_coo.argtypes = [_vp, _ll, _ip, _ip, _dp]
    L.hifde_set_stencil.argtypes = [_vp, _dp, _dp, _i]
    L.hifde_factor_run.argtypes = [_vp]
    L.hifde_apply.argtypes = [_vp, _dp, _i]
    L.hifde_apply_device.argtypes = [_vp, _vp, _i, _ll, _i]
    L.hifde_set_stream.argtypes = [_vp, _vp]
    L.hifde_set_stencil_device.argtypes = [_vp, _vp, _vp, _i]
    L.hifde_operator_csr.argtypes = [_vp, _ip, _ip, _dp]
    L.hifde_operator_csr.restype = _ll
    L.hifde_gemm_table_load.argtypes = [ctypes.c_char_p]
    L.hifde_create_batch.argtypes = [_i, _i, _d, _i, _i, _i, ctypes.POINTER(_vp)]
    L.hifde_set_stencil_batch_device.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i]
    L.hifde_batch_pack.argtypes = [_vp, _vp, _ll, _vp, _ll, _i, _i]
    L.hifde_gemm_table_clear.argtypes = []
    L.hifde_gemm_log.argtypes = [_vp, _ip, _ll]
    L.hifde_gemm_log.restype = _ll
    L.hifde_gemm_time.argtypes = [_i, _i, _i, _i, _ip, _ip, _ip, _i, _i, _dp]
    _u64 = ctypes.c_ulonglong
    L.hifde_gen_uniform.argtypes = [_vp, _ll, _u64, _d, _d, _i, _vp]
    L.hifde_gen_normal.argtypes = [_vp, _ll, _u64, _vp]
    L.hifde_gen_rhs_normal.argtypes = [_vp, _i, _i, _ll, _u64, _vp]
    L.hifde_gen_gaussian_field.argtypes = [_vp, _i, _u64, _d, _i, _vp]
    L.hifde_apply_device_noalloc.argtypes = [_vp, _vp, _i]
    L.hifde_spmv.argtypes = [_vp, _dp, _dp, _i]
    L.hifde_pcg.argtypes = [_vp, _dp, _dp, _d, _i, _ip, _dp]
    L.hifde_pcg_device.argtypes = [_vp, _vp, _vp, _d, _i, _ip, _dp]
    L.hifde_pcg_multi_device.argtypes = [_vp, _vp, _ll, _vp, _ll, _i, _d, _i, _ip, _dp]
    L.hifde_num_levels.argtypes = [_vp]
    L.hifde_level_counts.argtypes = [_vp, _i, _ip]
    L.hifde_box_count.argtypes = [_vp, _i, _i]
    L.hifde_box_sizes.argtypes = [_vp, _i, _i, _i, _ip]
    L.hifde_box_data.argtypes = [_vp, _i, _i, _i, _ip, _ip, _dp, _dp, _dp, _dp]
    L.hifde_stage_times.argtypes = [_vp, _dp, _i]
    L.hifde_level_stage_times.argtypes = [_vp, _i, _dp, _i]
    L.hifde_stage_name.argtypes = [_i]
    L.hifde_stage_name.restype = ctypes.c_char_p
    L.hifde_factor_bytes.argtypes = [_vp]
    L.hifde_factor_bytes.restype = _ll
    L.hifde_set_profiling.argtypes = [_vp, _i]
    L.hifde_set_profile_kernel.argtypes = [_vp, ctypes.c_char_p]
    L.hifde_factor_save.argtypes = [_vp, ctypes.c_char_p]
    L.hifde_factor_load.argtypes = [_vp, ctypes.c_char_p]
    L.hifde_set_option.argtypes = [_vp, ctypes.c_char_p, _d]
    L.hifde_reset_kernel_stats.argtypes = [_vp]
    L.hifde_kernel_stats.argtypes = [_vp, _i, ctypes.c_char_p, _i, _dp]
    L.hifde_interp_decomp_batch.argtypes = [_i, _ip, _ip, _dp, _d, _ip, _ip, _dp]
    L.hifde_pivoted_qr_batch.argtypes = [_i, _ip, _ip, _dp, _ip, _dp]
    L.hifde_triangular_solve_batch.argtypes = [_i, _ip, _ip, _dp, _dp]
    L.hifde_gemm_batch.argtypes = [_i, _i, _i, _ip, _ip, _ip, _dp, _dp, _dp]
    L.hifde_schur_batch.argtypes = [_i, _i, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
    _lib = L
    return L


def _ptr_array(ptrs):
    arr = (_vp * len(ptrs))(*[ctypes.c_void_p(int(p)) for p in ptrs])
    return ctypes.cast(arr, ctypes.POINTER(_vp)), arr


def _check(rc):
    if rc < 0:
        raise HifdeError(rc, lib().hifde_last_error().decode())
    return rc


def _ptr(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _current_stream():
    """PyTorch's current stream on this thread (a hipStream_t handle), or 0
    (the legacy default stream) when PyTorch is not in use."""
    torch = sys.modules.get("torch")
    if torch is None or not torch.cuda.is_initialized():
        return 0
    return torch.cuda.current_stream().cuda_stream


class HIFFactor:
    """HIFFactor(N, P, epsilon): N interior points per side, leaf width P,
    ID relative tolerance epsilon (include/Factor.hpp:12-13).

    precision: "fp64" (HIFFactor<double>) or "mixed" (FP32 Schur/LU blocks and
    stored factors; FP64 pivots, rank decisions and level matrix).
    front_quirk: reproduce the FRONT-face guard of src/Factor.cpp:486.
    """

    def __init__(self, N, P, epsilon, precision="fp64", front_quirk=True):
        self.N, self.P, self.epsilon = int(N), int(P), float(epsilon)
        self.NC = self.N + 1
        self.precision = precision
        prec = {"fp64": FP64, "mixed": MIXED}[precision]
        h = _vp()
        _check(lib().hifde_create(self.N, self.P, self.epsilon, prec,
                                  FLAG_FRONT_QUIRK if front_quirk else 0, ctypes.byref(h)))
        self._h = h
        self.lu_info = 0

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hifde_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def size(self):
        return self.NC ** 3

    @property
    def num_levels(self):
        return _check(lib().hifde_num_levels(self._h))

    # --- matrix -------------------------------------------------------------
    def set_matrix_coo(self, rows, cols, vals):
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        vals = np.ascontiguousarray(vals, np.float64)
        _check(lib().hifde_set_matrix_coo(self._h, rows.size, _ptr(rows, ctypes.c_int),
                                          _ptr(cols, ctypes.c_int), _ptr(vals)))

    def set_stencil(self, a, V, symmetric=False):
        """SetupStencil on nodal tensors of shape (N+2)^3, i fastest."""
        m = self.NC + 1
        a = np.asfortranarray(a, np.float64).ravel(order="F")
        V = np.asfortranarray(V, np.float64).ravel(order="F")
        if a.size != m ** 3 or V.size != m ** 3:
            raise ValueError("nodal tensors must have (N+2)^3 entries")
        _check(lib().hifde_set_stencil(self._h, _ptr(a), _ptr(V), int(symmetric)))

    def set_stencil_device(self, a_ptr, V_ptr, symmetric=False):
        """SetupStencil from device nodal tensors ((N+2)^3 doubles each, i
        fastest): the operator is assembled in HBM."""
        self._follow_caller_stream()
        _check(lib().hifde_set_stencil_device(self._h, ctypes.c_void_p(a_ptr),
                                              ctypes.c_void_p(V_ptr), int(symmetric)))

    def operator_csr(self):
        """The assembled operator as (rowptr, colind, vals)."""
        nnz = lib().hifde_operator_csr(self._h, None, None, None)
        _check(int(nnz))
        rp = np.zeros(self.size + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1))
        _check(int(lib().hifde_operator_csr(self._h, _ptr(rp, ctypes.c_int), _ptr(ci, ctypes.c_int),
                                            _ptr(v))))
        return rp, ci[:nnz], v[:nnz]

    # --- factor / apply -----------------------------------------------------
    def Factor(self):
        self.lu_info = _check(lib().hifde_factor_run(self._h))
        return self.lu_info

    factor = Factor

    def Apply(self, u, apply_inverse=False):
        """Return F u (or F^{-1} u) for a host vector of length (N+1)^3."""
        u = np.array(u, dtype=np.float64, copy=True)
        if u.size != self.size:
            raise ValueError("vector length must be (N+1)^3")
        _check(lib().hifde_apply(self._h, _ptr(u), int(apply_inverse)))
        return u

    apply = Apply

    def _follow_caller_stream(self):
        # device entry points order their work after the caller's stream
        _check(lib().hifde_set_stream(self._h, ctypes.c_void_p(_current_stream())))

    def apply_device(self, ptr, nrhs=1, ldu=None, inverse=True):
        self._follow_caller_stream()
        _check(lib().hifde_apply_device(self._h, ctypes.c_void_p(ptr), nrhs,
                                        ldu if ldu is not None else self.size, int(inverse)))

    def apply_device_noalloc(self, ptr, inverse=True):
        """One device vector through the allocation-free apply (grids up to
        128^3; calls overlapping on one GPU race on its shared scratch)."""
        self._follow_caller_stream()
        _check(lib().hifde_apply_device_noalloc(self._h, ctypes.c_void_p(ptr), int(inverse)))

    def spmv(self, x, transpose=False):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.size)
        _check(lib().hifde_spmv(self._h, _ptr(x), _ptr(y), int(transpose)))
        return y

    def pcg(self, b, x0=None, tol=1e-10, maxit=200):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.size) if x0 is None else np.array(x0, np.float64, copy=True)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0)
        _check(lib().hifde_pcg(self._h, _ptr(b), _ptr(x), tol, maxit, ctypes.byref(it),
                               ctypes.byref(rr)))
        return x, it.value, rr.value

    def pcg_device(self, b_ptr, x_ptr, tol=1e-10, maxit=200):
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0)
        self._follow_caller_stream()
        _check(lib().hifde_pcg_device(self._h, ctypes.c_void_p(b_ptr), ctypes.c_void_p(x_ptr),
                                      tol, maxit, ctypes.byref(it), ctypes.byref(rr)))
        return it.value, rr.value

    def pcg_multi_device(self, b_ptr, x_ptr, nrhs, ldb=None, ldx=None, tol=1e-10, maxit=200):
        """nrhs independent PCG solves (columns of B and X in device memory,
        column-major).  Returns per-column (iterations, relative residuals)."""
        it = np.zeros(max(nrhs, 1), np.int32)
        rr = np.zeros(max(nrhs, 1))
        self._follow_caller_stream()
        _check(lib().hifde_pcg_multi_device(
            self._h, ctypes.c_void_p(b_ptr), ldb if ldb is not None else self.size,
            ctypes.c_void_p(x_ptr), ldx if ldx is not None else self.size, nrhs, tol, maxit,
            it.ctypes.data_as(_ip), rr.ctypes.data_as(_dp)))
        return it[:nrhs].copy(), rr[:nrhs].copy()

    def save(self, path):
        """Write the factorization (operator, box index, factor blocks) to path."""
        _check(lib().hifde_factor_save(self._h, os.fsencode(path)))

    def load(self, path):
        """Read a factorization saved by an object of the same N, P, epsilon and
        precision; apply and solve then work without factoring."""
        self.lu_info = _check(lib().hifde_factor_load(self._h, os.fsencode(path)))
        return self.lu_info

    # --- introspection ------------------------------------------------------
    def level_counts(self):
        out = []
        buf = np.zeros(5, np.int32)
        for lv in range(self.num_levels):
            _check(lib().hifde_level_counts(self._h, lv, _ptr(buf, ctypes.c_int)))
            out.append(tuple(int(x) for x in buf))
        return out

    def num_boxes(self, level, kind):
        return _check(lib().hifde_box_count(self._h, level, kind))

    def box(self, level, kind, b):
        sz = np.zeros(2, np.int32)
        _check(lib().hifde_box_sizes(self._h, level, kind, b, _ptr(sz, ctypes.c_int)))
        r, s = int(sz[0]), int(sz[1])
        red = np.zeros(r, np.int32)
        skel = np.zeros(s, np.int32)
        A22 = np.zeros(r * r)
        A22inv = np.zeros(r * r)
        X = np.zeros(r * s)
        W = np.zeros(r * s)
        _check(lib().hifde_box_data(self._h, level, kind, b, _ptr(red, ctypes.c_int),
                                    _ptr(skel, ctypes.c_int), _ptr(A22), _ptr(A22inv), _ptr(X),
                                    _ptr(W)))
        out = dict(red_gid=red, skel_gid=skel, A22=A22.reshape(r, r, order="F"),
                   A22inv=A22inv.reshape(r, r, order="F"), X=X.reshape(r, s, order="F"))
        if kind == 1:
            out["W"] = W.reshape(s, r, order="F")
        return out

    def stage_times(self):
        buf = np.zeros(16)
        n = _check(lib().hifde_stage_times(self._h, _ptr(buf), 16))
        return {lib().hifde_stage_name(i).decode(): float(buf[i]) for i in range(n)}

    def level_stage_times(self):
        """Per level of the last factor: {stage name: seconds}."""
        out = []
        buf = np.zeros(16)
        for lv in range(self.num_levels):
            n = _check(lib().hifde_level_stage_times(self._h, lv, _ptr(buf), 16))
            out.append({lib().hifde_stage_name(i).decode(): float(buf[i]) for i in range(n)})
        return out

    def set_option(self, name, value):
        _check(lib().hifde_set_option(self._h, name.encode(), float(value)))

    def set_profiling(self, on=True, kernel=None):
        """Per-family device timing; with `kernel`, only that kernel's launches
        are timed (low overhead)."""
        _check(lib().hifde_set_profiling(self._h, int(on)))
        _check(lib().hifde_set_profile_kernel(self._h, kernel.encode() if kernel else None))

    def reset_kernel_stats(self):
        _check(lib().hifde_reset_kernel_stats(self._h))

    def kernel_stats(self):
        """{family: dict(launches, sec, flops, bytes)} of device time per kernel family."""
        L = lib()
        n = _check(L.hifde_kernel_stats(self._h, -1, None, 0, None))
        out = {}
        name = ctypes.create_string_buffer(64)
        vals = np.zeros(4)
        for i in range(n):
            L.hifde_kernel_stats(self._h, i, name, 64, _ptr(vals))
            out[name.value.decode()] = dict(launches=int(vals[0]), sec=float(vals[1]),
                                            flops=float(vals[2]), bytes=float(vals[3]))
        return out

    def factor_bytes(self):
        return int(lib().hifde_factor_bytes(self._h))

    def gemm_log(self):
        """Descriptors of the last factorization (option "gemm_record" = 1):
        int array of rows (call, esize, transA, lower, m, n, k, ncall, kind)."""
        n = int(lib().hifde_gemm_log(self._h, None, 0))
        out = np.zeros(max(n, 9), np.int32)
        lib().hifde_gemm_log(self._h, _ptr(out, ctypes.c_int), n)
        return out[:n].reshape(-1, 9)


# --- device generators (the definitions of problems.py, computed on the GPU)
def _stream_arg():
    return ctypes.c_void_p(_current_stream())


def gen_uniform(ptr, n, seed, lo=0.0, hi=1.0, pow10=False):
    """n doubles at device address ptr: lo + (hi - lo) U (or 10^ of it)."""
    _check(lib().hifde_gen_uniform(ctypes.c_void_p(ptr), int(n), int(seed) & ((1 << 64) - 1),
                                   float(lo), float(hi), int(pow10), _stream_arg()))


def gen_normal(ptr, n, seed):
    _check(lib().hifde_gen_normal(ctypes.c_void_p(ptr), int(n), int(seed) & ((1 << 64) - 1),
                                  _stream_arg()))


def gen_rhs_normal(ptr, NC, seed, nrhs=1, ld=None):
    _check(lib().hifde_gen_rhs_normal(ctypes.c_void_p(ptr), int(NC), int(nrhs),
                                      int(ld if ld is not None else NC ** 3),
                                      int(seed) & ((1 << 64) - 1), _stream_arg()))


def gen_gaussian_field(ptr, NC, seed, corr=0.1, lognormal=False):
    _check(lib().hifde_gen_gaussian_field(ctypes.c_void_p(ptr), int(NC),
                                          int(seed) & ((1 << 64) - 1), float(corr),
                                          int(lognormal), _stream_arg()))


def device_problem(NC, coeff, seed, V="zeros", v_seed=None, corr=0.1, lo=-1.0, hi=1.0):
    """Nodal tensors a and V ((NC+1)^3 float64 torch tensors on the current
    device, i fastest) generated on the GPU: coeff "pow10" (a = 10^U[lo, hi]),
    "lognormal" (exp of the Gaussian field) or "ones"; V "zeros" or "uniform"
    (U[0, 1], seed v_seed)."""
    import torch
    m3 = (NC + 1) ** 3
    a = torch.empty(m3, dtype=torch.float64, device="cuda")
    if coeff == "pow10":
        gen_uniform(a.data_ptr(), m3, seed, lo, hi, pow10=True)
    elif coeff == "lognormal":
        gen_gaussian_field(a.data_ptr(), NC, seed, corr, lognormal=True)
    elif coeff == "ones":
        a.fill_(1.0)
    else:
        raise ValueError(coeff)
    Vt = torch.zeros(m3, dtype=torch.float64, device="cuda")
    if V == "uniform":
        gen_uniform(Vt.data_ptr(), m3, v_seed, 0.0, 1.0)
    elif V != "zeros":
        raise ValueError(V)
    return a, Vt


class HIFBatch(HIFFactor):
    """nprob (1 to 8) independent problems of one grid -- N interior points
    per side, leaf P, ID tolerance epsilon -- factored by one object: packed
    as the octants of a grid twice as wide (identity rows on the separating
    planes and unused octants), so that every level's boxes of all of them go
    through the same launches.  Vectors move between the problems' own layout
    ((N+1)^3 each) and the packed one with pack / unpack; pcg_multi_device
    runs one PCG per (problem, column) and returns per-pair iteration counts
    and residuals shaped (nprob, nrhs)."""

    def __init__(self, N, P, epsilon, nprob, precision="mixed", front_quirk=True):
        self.N, self.P, self.epsilon = 2 * (int(N) + 1) - 1, int(P), float(epsilon)
        self.NC = self.N + 1
        self.N_sub, self.NC_sub, self.nprob = int(N), int(N) + 1, int(nprob)
        self.precision = precision
        prec = {"fp64": FP64, "mixed": MIXED}[precision]
        h = _vp()
        _check(lib().hifde_create_batch(int(N), self.P, self.epsilon, prec,
                                        FLAG_FRONT_QUIRK if front_quirk else 0, self.nprob,
                                        ctypes.byref(h)))
        self._h = h
        self.lu_info = 0

    def set_stencil_batch_device(self, a_ptrs, V_ptrs, symmetric=True):
        """Problem o's stencil from device nodal tensors a_ptrs[o], V_ptrs[o]
        ((N+2)^3 doubles each, i fastest)."""
        if len(a_ptrs) != self.nprob or len(V_ptrs) != self.nprob:
            raise ValueError("one coefficient pair per problem")
        self._follow_caller_stream()
        pa, keep_a = _ptr_array(a_ptrs)
        pv, keep_v = _ptr_array(V_ptrs)
        _check(lib().hifde_set_stencil_batch_device(self._h, pa, pv, int(symmetric)))

    def pack(self, sub_ptr, packed_ptr, nrhs=1, ld_sub=None, ld_packed=None):
        """Problems' vectors (vector c of problem o at sub + (o*nrhs + c)*ld_sub)
        into packed columns."""
        self._follow_caller_stream()
        _check(lib().hifde_batch_pack(self._h, ctypes.c_void_p(sub_ptr),
                                      ld_sub or self.NC_sub ** 3, ctypes.c_void_p(packed_ptr),
                                      ld_packed or self.size, int(nrhs), 0))

    def unpack(self, packed_ptr, sub_ptr, nrhs=1, ld_sub=None, ld_packed=None):
        self._follow_caller_stream()
        _check(lib().hifde_batch_pack(self._h, ctypes.c_void_p(sub_ptr),
                                      ld_sub or self.NC_sub ** 3, ctypes.c_void_p(packed_ptr),
                                      ld_packed or self.size, int(nrhs), 1))

    def pcg_multi_device(self, b_ptr, x_ptr, nrhs, ldb=None, ldx=None, tol=1e-10, maxit=200):
        it = np.zeros(max(nrhs * self.nprob, 1), np.int32)
        rr = np.zeros(max(nrhs * self.nprob, 1))
        self._follow_caller_stream()
        _check(lib().hifde_pcg_multi_device(
            self._h, ctypes.c_void_p(b_ptr), ldb if ldb is not None else self.size,
            ctypes.c_void_p(x_ptr), ldx if ldx is not None else self.size, nrhs, tol, maxit,
            it.ctypes.data_as(_ip), rr.ctypes.data_as(_dp)))
        n = nrhs * self.nprob
        return it[:n].reshape(self.nprob, nrhs).copy(), rr[:n].reshape(self.nprob, nrhs).copy()


def interp_decomp_batch(mats, eps):
    """InterpDecomp on a list of 2-D arrays.  Returns [(skel, red, W)]."""
    m = np.array([a.shape[0] for a in mats], np.int32)
    n = np.array([a.shape[1] for a in mats], np.int32)
    packed = np.concatenate([np.asfortranarray(a, np.float64).ravel(order="F") for a in mats]
                            + [np.zeros(1)])
    k = np.zeros(len(mats), np.int32)
    perm = np.zeros(max(int(n.sum()), 1), np.int32)
    W = np.zeros(max(int((n.astype(np.int64) ** 2).sum()), 1))
    _check(lib().hifde_interp_decomp_batch(len(mats), _ptr(m, ctypes.c_int), _ptr(n, ctypes.c_int),
                                           _ptr(packed), eps, _ptr(k, ctypes.c_int),
                                           _ptr(perm, ctypes.c_int), _ptr(W)))
    out, po, wo = [], 0, 0
    for b in range(len(mats)):
        nb, kb = int(n[b]), int(k[b])
        p = perm[po:po + nb]
        Wb = W[wo:wo + kb * (nb - kb)].reshape(kb, nb - kb, order="F")
        out.append((p[:kb].copy(), p[kb:].copy(), Wb.copy()))
        po += nb
        wo += nb * nb
    return out


def pivoted_qr_batch(mats):
    """PivotedQRWrapper on a list of 2-D arrays.  Returns [(QR, jpvt, tau)]:
    QR holds R on and above the diagonal and the Householder vectors below
    it (dgeqp3's layout), jpvt the 0-based column pivots."""
    m = np.array([a.shape[0] for a in mats], np.int32)
    n = np.array([a.shape[1] for a in mats], np.int32)
    packed = np.concatenate([np.asfortranarray(a, np.float64).ravel(order="F") for a in mats]
                            + [np.zeros(1)])
    jpvt = np.zeros(max(int(n.sum()), 1), np.int32)
    mn = np.minimum(m, n)
    tau = np.zeros(max(int(mn.sum()), 1))
    _check(lib().hifde_pivoted_qr_batch(len(mats), _ptr(m, ctypes.c_int), _ptr(n, ctypes.c_int),
                                        _ptr(packed), _ptr(jpvt, ctypes.c_int), _ptr(tau)))
    out, o, po, to = [], 0, 0, 0
    for b in range(len(mats)):
        mb, nb = int(m[b]), int(n[b])
        out.append((packed[o:o + mb * nb].reshape(mb, nb, order="F").copy(),
                    jpvt[po:po + nb].copy(), tau[to:to + int(mn[b])].copy()))
        o += mb * nb
        po += nb
        to += int(mn[b])
    return out


def triangular_solve_batch(pairs):
    """TriangularSolveWrapper on a list of (R, B): returns [R^{-1} B] with R
    upper triangular (only its upper triangle is read)."""
    m = np.array([r.shape[0] for r, _ in pairs], np.int32)
    n = np.array([b.shape[1] for _, b in pairs], np.int32)
    R = np.concatenate([np.asfortranarray(r, np.float64).ravel(order="F") for r, _ in pairs]
                       + [np.zeros(1)])
    B = np.concatenate([np.asfortranarray(b, np.float64).ravel(order="F") for _, b in pairs]
                       + [np.zeros(1)])
    _check(lib().hifde_triangular_solve_batch(len(pairs), _ptr(m, ctypes.c_int),
                                              _ptr(n, ctypes.c_int), _ptr(R), _ptr(B)))
    out, o = [], 0
    for b in range(len(pairs)):
        cnt = int(m[b]) * int(n[b])
        out.append(B[o:o + cnt].reshape(int(m[b]), int(n[b]), order="F").copy())
        o += cnt
    return out


def gemm_time(precision, transA, lower, shapes, kind, reps=5):
    """Average device seconds of one batched launch of products (m, n, k)
    forced onto kernel kind (0 FMA, 1 matrix-core 128, 2 matrix-core 64)."""
    m = np.array([x[0] for x in shapes], np.int32)
    n = np.array([x[1] for x in shapes], np.int32)
    k = np.array([x[2] for x in shapes], np.int32)
    t = ctypes.c_double(0)
    _check(lib().hifde_gemm_time({"fp64": FP64, "mixed": MIXED}[precision], int(transA),
                                 int(lower), len(shapes), _ptr(m, ctypes.c_int),
                                 _ptr(n, ctypes.c_int), _ptr(k, ctypes.c_int), int(kind),
                                 int(reps), ctypes.byref(t)))
    return t.value


def gemm_table_load(path=None):
    """Load a GEMM routing table (None: the one shipped beside the library)."""
    return _check(lib().hifde_gemm_table_load(os.fsencode(path) if path else None))


def gemm_table_clear():
    _check(lib().hifde_gemm_table_clear())


BF16X3 = 0x200


GEMM_TILE64 = 0x400


GEMM_FORCE = {0: 0x800, 1: 0x1000, 2: 0x2000}


GEMM_RULE = 0x4000


GEMM_LOWER = 0x8000


def gemm_batch(pairs, precision="fp64", transA=False, bf16x3=False, tile64=False, kind=None,
               rule=False, lower=False):
    """[op(A) @ B for (A, B) in pairs] through the library's batched GEMM.
    bf16x3 (precision "mixed" only): matrix-core products as bf16x3 split
    products.  tile64: products with both output sides of at least 32 on
    64 x 64 matrix-core tiles.  kind: force one kernel (0 FMA, 1 matrix-core
    128, 2 matrix-core 64).  rule: route by the fixed rule even where a loaded
    routing table has an entry.  lower: square products compute only their
    output tiles on or below the diagonal (only the lower triangle of such a
    result is defined)."""
    prec = ({"fp64": FP64, "mixed": MIXED}[precision] | (BF16X3 if bf16x3 else 0)
            | (GEMM_TILE64 if tile64 else 0) | (GEMM_FORCE[kind] if kind is not None else 0)
            | (GEMM_RULE if rule else 0) | (GEMM_LOWER if lower else 0))
    m = np.array([(a.shape[1] if transA else a.shape[0]) for a, _ in pairs], np.int32)
    k = np.array([(a.shape[0] if transA else a.shape[1]) for a, _ in pairs], np.int32)
    n = np.array([b.shape[1] for _, b in pairs], np.int32)
    A = np.concatenate([np.asfortranarray(a, np.float64).ravel(order="F") for a, _ in pairs]
                       + [np.zeros(1)])
    B = np.concatenate([np.asfortranarray(b, np.float64).ravel(order="F") for _, b in pairs]
                       + [np.zeros(1)])
    C = np.zeros(max(int((m.astype(np.int64) * n).sum()), 1))
    _check(lib().hifde_gemm_batch(prec, int(transA), len(pairs), _ptr(m, ctypes.c_int),
                                  _ptr(n, ctypes.c_int), _ptr(k, ctypes.c_int), _ptr(A), _ptr(B),
                                  _ptr(C)))
    out, o = [], 0
    for b in range(len(pairs)):
        cnt = int(m[b]) * int(n[b])
        out.append(C[o:o + cnt].reshape(int(m[b]), int(n[b]), order="F").copy())
        o += cnt
    return out


SCHUR_SYMMETRIC = 0x100


def schur_batch(blocks, precision="fp64", symmetric=False, bf16x3=False):
    """Schur on a list of (A22, A21, A12).  Returns [(A22inv, X, S, info)].
    symmetric: every A22 is symmetric (only its lower triangle is read).
    bf16x3 (precision "mixed" only): X and S as bf16x3 split products."""
    prec = ({"fp64": FP64, "mixed": MIXED}[precision] | (SCHUR_SYMMETRIC if symmetric else 0)
            | (BF16X3 if bf16x3 else 0))
    r = np.array([b[0].shape[0] for b in blocks], np.int32)
    s = np.array([b[1].shape[1] for b in blocks], np.int32)

    def pack(i):
        return np.concatenate([np.asfortranarray(b[i], np.float64).ravel(order="F")
                               for b in blocks] + [np.zeros(1)])
    A22, A21, A12 = pack(0), pack(1), pack(2)
    rr = int((r.astype(np.int64) ** 2).sum())
    rs = int((r.astype(np.int64) * s).sum())
    ss = int((s.astype(np.int64) ** 2).sum())
    Ainv = np.zeros(max(rr, 1))
    X = np.zeros(max(rs, 1))
    S = np.zeros(max(ss, 1))
    info = np.zeros(len(blocks), np.int32)
    _check(lib().hifde_schur_batch(prec, len(blocks), _ptr(r, ctypes.c_int), _ptr(s, ctypes.c_int),
                                   _ptr(A22), _ptr(A21), _ptr(A12), _ptr(Ainv), _ptr(X), _ptr(S),
                                   _ptr(info, ctypes.c_int)))
    out = []
    o1 = o2 = o3 = 0
    for b in range(len(blocks)):
        rb, sb = int(r[b]), int(s[b])
        out.append((Ainv[o1:o1 + rb * rb].reshape(rb, rb, order="F").copy(),
                    X[o2:o2 + rb * sb].reshape(rb, sb, order="F").copy(),
                    S[o3:o3 + sb * sb].reshape(sb, sb, order="F").copy(), int(info[b])))
        o1 += rb * rb
        o2 += rb * sb
        o3 += sb * sb
    return out
