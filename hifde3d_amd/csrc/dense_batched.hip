// Batched dense kernels for the per-level box work of the factorization.
//
//  * k_gj_inverse   explicit inverse with partial pivoting, one workgroup per
//                   box (the LU + explicit inverse of hmat_tools::Invert,
//                   src/hmat_tools/Invert.cpp:35-42, done as in-place
//                   Gauss-Jordan: same pivot rule, same result up to rounding).
//  * k_gemm_nn      C = alpha * A * B over a flattened list of 64x64 output
//                   tiles of every box (Schur.hpp:97,99 products).
//  * k_interp_decomp  column-pivoted Householder QR (LAPACK dgeqp3 pivot rule
//                   and partial-norm downdating, src/InterpDecomp.cpp:94,
//                   :156-166) followed by the reference's relative-tolerance
//                   rank rule and triangular solve (src/InterpDecomp.cpp:
//                   100-137).  Always FP64: pivots and rank decisions.
//  * k_face_transform  closed form of SchurAfterID's Rot^T A(F,F) Rot
//                   (src/Factor.cpp:113-153) restricted to the three blocks
//                   Schur needs.

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <type_traits>

#include "hifde_internal.hpp"

namespace hifde {

namespace {

constexpr int kWave = 64;

template <class O>
__device__ __forceinline__ int find_box(const O* __restrict__ off, int nbox, long long e) {
    int lo = 0, hi = nbox;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if ((long long)off[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Block-wide sum; every thread receives the result.  red must hold >= 16 doubles.
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    int nw = (blockDim.x + kWave - 1) / kWave;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += red[w];
    __syncthreads();
    return t;
}

// Block-wide arg-max of |value| with ties to the smaller index.
template <class T>
__device__ void block_argmax(T v, int idx, T* redv, int* redi, T* outv, int* outi) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        T ov = __shfl_xor(v, o, kWave);
        int oi = __shfl_xor(idx, o, kWave);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    int nw = (blockDim.x + kWave - 1) / kWave;
    __syncthreads();
    if (lane == 0) { redv[wid] = v; redi[wid] = idx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        T bv = redv[0];
        int bi = redi[0];
        for (int w = 1; w < nw; ++w)
            if (redv[w] > bv || (redv[w] == bv && redi[w] < bi)) { bv = redv[w]; bi = redi[w]; }
        *outv = bv;
        *outi = bi;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Explicit inverse by Gauss-Jordan elimination.
//
// Boxes with n <= kSmallGJ: in-place GJ with row partial pivoting (the pivot
// of column c is the first row >= c of largest magnitude, as in getrf) in
// LDS, one workgroup per box.  In-place GJ keeps [A | I] in one array; the
// final column interchanges (reverse pivot order) turn it into A^{-1}.
//
// Larger boxes: two-level block Gauss-Jordan.  The outer level walks
// kGJPanel-wide diagonal windows K of the box:
//   A_KK <- D = A_KK^{-1}            inner level below, in place
//   R    <- D A_KJ                   (J != K)      GEMM into a workspace
//   A_IJ <- A_IJ - A_IK R            (I, J != K)   GEMM, rank kGJPanel
//   C    <- -A_IK D                  (I != K)      GEMM into a workspace
//   A_IK <- C,  A_KJ <- R                          strided copies
// so the bulk of the 2n^3 flops is a rank-256 update, which is compute-bound
// on the matrix cores.  The inner level inverts a window with kGJB-wide
// diagonal blocks and the same phases, in four launches:
//   A_kk <- D = A_kk^{-1}                        k_bgj_diag   (1 wave per box)
//   R    <- D A_kj              (j != k)         k_bgj_row    (column chunks)
//   A_ij <- A_ij - A_ik R       (i, j != k)      GEMM
//   A_kj <- R;  A_ik <- -A_ik D, A_kk <- D       k_bgj_col    (row chunks)
// Pivoting is partial pivoting inside each kGJB diagonal block.  The blocks
// of this path are the interior blocks of symmetric positive definite or
// M-matrix operators, whose diagonal blocks are nonsingular; a singular
// diagonal block is reported through info.
// ---------------------------------------------------------------------------
constexpr int kSmallGJ = 64;
constexpr int kGJB = 64;
constexpr int kGJPanel = 512;
// GemmPlan::add with one_launch: calls of at most this many descriptors
// are issued as one launch
constexpr int kUnifyMaxDesc = 16;

// Wave-wide arg-max of non-negative values, ties to the lower lane (values
// equal in all but their 6 lowest mantissa bits count as tied).  The key is
// the value's bit pattern with those 6 bits replaced by 63 - lane; for
// non-negative floating-point values the bit pattern orders like the value,
// so the largest key is found with floating-point max.  Four DPP row shifts
// reduce each row of 16 lanes (lanes shifted in from outside the row read 0),
// and the four row maxima are read back with readlane.  Returns the winning
// lane (the same in every lane); nonzero is false when every value is zero.
template <int N>
__device__ __forceinline__ double row_shr(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x110 + N, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x110 + N, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
template <int N>
__device__ __forceinline__ float row_shr(float x) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x110 + N, 0xF, 0xF, true));
}
__device__ __forceinline__ double read_lane(double x, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                            __builtin_amdgcn_readlane(__double2loint(x), l));
}
__device__ __forceinline__ float read_lane(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double with_lane(double v, int lane) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return __longlong_as_double((long long)((b & ~0x3Full) | (unsigned long long)(63 - lane)));
}
__device__ __forceinline__ float with_lane(float v, int lane) {
    const unsigned b = (unsigned)__float_as_int(v);
    return __int_as_float((int)((b & ~0x3Fu) | (unsigned)(63 - lane)));
}
__device__ __forceinline__ int key_lane(double k) {
    return 63 - (int)((unsigned long long)__double_as_longlong(k) & 0x3Full);
}
__device__ __forceinline__ int key_lane(float k) {
    return 63 - (int)((unsigned)__float_as_int(k) & 0x3Fu);
}
__device__ __forceinline__ bool key_nonzero(double k) {
    return ((unsigned long long)__double_as_longlong(k) & ~0x3Full) != 0;
}
__device__ __forceinline__ bool key_nonzero(float k) {
    return ((unsigned)__float_as_int(k) & ~0x3Fu) != 0;
}
template <class T>
__device__ __forceinline__ int wave_argmax_nonneg(T v, int lane, bool& nonzero) {
    T k = with_lane(v, lane);
    k = fmax(k, row_shr<1>(k));
    k = fmax(k, row_shr<2>(k));
    k = fmax(k, row_shr<4>(k));
    k = fmax(k, row_shr<8>(k));
    T m = fmax(fmax(read_lane(k, 15), read_lane(k, 31)), fmax(read_lane(k, 47), read_lane(k, 63)));
    nonzero = key_nonzero(m);
    return key_lane(m);
}

// Pivoted in-place GJ of an n x n block (n <= 64) held in LDS as S[col][row],
// padded with the identity to 64 x 64 (the inverse of diag(A, I) is
// diag(A^{-1}, I)).  256 threads: thread t owns column t >> 2 and rows
// (t & 3) + 4q.  Pivot search is one wave's shuffle reduction; the column
// interchanges are applied at the end as one gather through the composed
// permutation.  Returns false on an exactly singular pivot column.
template <class T>
__device__ bool lds_gj_inverse64(T (*S)[kSmallGJ + 1], int n, T* colk, T* rowk, int* perm,
                                 int* s_piv) {
    const int tid = threadIdx.x;
    const int c = tid >> 2, r0 = tid & 3;
    for (int k = 0; k < n; ++k) {
        if (tid < 64) {  // wave 0, all lanes
            bool nz;
            const int idx =
                wave_argmax_nonneg<T>((tid >= k && tid < n) ? fabs(S[k][tid]) : (T)0, tid, nz);
            if (tid == 0) s_piv[0] = nz ? idx : -1;
        }
        __syncthreads();
        const int p = s_piv[0];
        if (p < 0) return false;
        if (tid < 64) {
            if (p != k) {
                T t = S[tid][k];
                S[tid][k] = S[tid][p];
                S[tid][p] = t;
            }
            if (tid == 0) perm[k] = p;
        }
        __syncthreads();
        if (tid < 64) {
            const T d = (T)1 / S[k][k];
            colk[tid] = S[k][tid];
            rowk[tid] = (tid == k) ? d : S[tid][k] * d;
        }
        __syncthreads();
        const T rc = rowk[c];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int i = r0 + 4 * q;
            if (i == k) {
                S[c][i] = rc;
            } else {
                const T a = (c == k) ? (T)0 : S[c][i];
                S[c][i] = a - colk[i] * rc;
            }
        }
        __syncthreads();
    }
    // columns of the result: res(:, j) = S(:, pi(j)), pi = swaps applied in reverse
    if (tid == 0) {
        for (int j = 0; j < 64; ++j) s_piv[1 + j] = j;
        for (int k = n - 1; k >= 0; --k) {
            const int p = perm[k];
            if (p != k) {
                int t = s_piv[1 + k];
                s_piv[1 + k] = s_piv[1 + p];
                s_piv[1 + p] = t;
            }
        }
    }
    __syncthreads();
    T tmp[16];
    const int src = s_piv[1 + c];
#pragma unroll
    for (int q = 0; q < 16; ++q) tmp[q] = S[src][r0 + 4 * q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) S[c][r0 + 4 * q] = tmp[q];
    __syncthreads();
    return true;
}

// Boxes of at most kWaveGJ rows: one wave per box, no barriers and no LDS.
// Lane i holds row i in registers.  Row interchanges exchange positions, not
// data: the pivot of step k is the largest |a| among the rows at positions
// >= k, ties to the lowest position, as with physical interchanges, and the
// arithmetic per entry is that of lds_gj_inverse64, so both give the same
// result.  The columns are permuted on output, as there.
constexpr int kWaveGJ = 32;
template <class T>
__global__ void __launch_bounds__(256) k_gj_wave(int nbox, const int* __restrict__ dims,
                                                 const long long* __restrict__ off, T* mats,
                                                 int* info) {
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= nbox) return;
    const int n = dims[b];
    if (n == 0 || n > kWaveGJ) return;
    T* A = mats + off[b];
    const bool row = lane < n;
    T a[kWaveGJ];
#pragma unroll
    for (int c = 0; c < kWaveGJ; ++c) a[c] = (row && c < n) ? A[lane + (size_t)c * n] : (T)0;
    int pos = lane;
    int prm[kWaveGJ];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < kWaveGJ; ++k) {
        prm[k] = k;
        if (k >= n || !ok) continue;  // (no break: the loop stays unrolled)
        bool nz;
        const int pp =
            wave_argmax_nonneg<T>((row && pos >= k) ? fabs(a[k]) : (T)0, pos, nz);
        if (!nz) {
            ok = false;
            continue;
        }
        prm[k] = pp;
        const int pl = __ffsll((long long)__ballot(row && pos == pp)) - 1;
        const int old = pos;
        if (lane == pl) pos = k;
        else if (old == k) pos = pp;
        const T d = (T)1 / read_lane(a[k], pl);
        const T f = a[k];
        const bool me = lane == pl;
#pragma unroll
        for (int c = 0; c < kWaveGJ; ++c) {
            const T rc = (c == k) ? d : read_lane(a[c], pl) * d;
            if (me) {
                a[c] = rc;
            } else {
                const T av = (c == k) ? (T)0 : a[c];
                a[c] = av - f * rc;
            }
        }
    }
    if (!ok) {
        if (lane == 0) info[b] = 1;
        return;
    }
    // column c of the in-place result is column j of the inverse, where j is
    // c taken through the interchanges in reverse order
#pragma unroll
    for (int c = 0; c < kWaveGJ; ++c) {
        if (c >= n) continue;
        int j = c;
#pragma unroll
        for (int k = kWaveGJ - 1; k >= 0; --k) {
            if (j == k) j = prm[k];
            else if (j == prm[k]) j = k;
        }
        if (row) A[pos + (size_t)j * n] = a[c];
    }
    if (lane == 0) info[b] = 0;
}

template <class T>
__global__ void __launch_bounds__(256) k_gj_small(const int* __restrict__ dims,
                                                  const long long* __restrict__ off, T* mats,
                                                  int* info) {
    const int b = blockIdx.x;
    const int n = dims[b];
    if (n <= kWaveGJ || n > kSmallGJ) return;  // smaller boxes: k_gj_wave
    T* A = mats + off[b];
    __shared__ T S[kSmallGJ][kSmallGJ + 1];  // S[col][row]
    __shared__ T colk[kSmallGJ], rowk[kSmallGJ];
    __shared__ int perm[kSmallGJ];
    __shared__ int s_piv[kSmallGJ + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < kSmallGJ * kSmallGJ; e += 256) {
        const int c = e >> 6, i = e & 63;
        S[c][i] = (c < n && i < n) ? A[i + (size_t)c * n] : (T)(c == i ? 1 : 0);
    }
    __syncthreads();
    if (!lds_gj_inverse64<T>(S, n, colk, rowk, perm, s_piv)) {
        if (tid == 0) info[b] = 1;
        return;
    }
    for (int e = tid; e < n * n; e += 256) A[e] = S[e / n][e % n];
    if (tid == 0) info[b] = 0;
}

// Inverse D of the kb x kb diagonal block K of a window (kb <= 64), one
// workgroup of kDiagWaves waves per box.  Lane i of every wave holds row i of
// diag(A_KK, I_{64-kb}); wave w holds the columns w + kDiagWaves q in
// registers, so the pivot row of a step is lane p of each wave and is read
// with readlane, without LDS traffic.  In-place Gauss-Jordan with partial
// pivoting: the pivot of step k is the not yet used row of largest |a_ik|
// (compared on a 32-bit key, ties to the lower row).  The wave holding column
// k publishes it in LDS
// (double-buffered, one barrier per step); every wave reads it, finds the same
// pivot and updates its own columns.  With p_k the pivot row of step k, the
// in-place array Z satisfies A^{-1}(k, p_c) = Z(p_k, c), which is how each
// lane writes its row out.
constexpr int kDiagWaves = 8;
constexpr int kDiagCols = kGJB / kDiagWaves;
// 32-bit arg-max key of a non-negative value: the high word of a double (sign,
// exponent, 20 mantissa bits) or a float's bits, orders like the value; its
// low 6 bits are replaced by 63 - lane, so values equal in the remaining bits
// (relative 2^-14 for double, 2^-17 for float) tie to the lower lane.
__device__ __forceinline__ unsigned argmax_key(double v, int lane) {
    return ((unsigned)__double2hiint(v) & ~0x3Fu) | (unsigned)(63 - lane);
}
__device__ __forceinline__ unsigned argmax_key(float v, int lane) {
    return ((unsigned)__float_as_int(v) & ~0x3Fu) | (unsigned)(63 - lane);
}
template <int N>
__device__ __forceinline__ unsigned row_shr_u(unsigned x) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x110 + N, 0xF, 0xF, true);
}
// Arg-max of the keys over the wave (uniform); nonzero is false when every
// value is zero.
__device__ __forceinline__ int wave_argmax_key(unsigned k, bool& nonzero) {
    k = max(k, row_shr_u<1>(k));
    k = max(k, row_shr_u<2>(k));
    k = max(k, row_shr_u<4>(k));
    k = max(k, row_shr_u<8>(k));
    const unsigned m0 = (unsigned)__builtin_amdgcn_readlane((int)k, 15);
    const unsigned m1 = (unsigned)__builtin_amdgcn_readlane((int)k, 31);
    const unsigned m2 = (unsigned)__builtin_amdgcn_readlane((int)k, 47);
    const unsigned m3 = (unsigned)__builtin_amdgcn_readlane((int)k, 63);
    const unsigned m = max(max(m0, m1), max(m2, m3));
    nonzero = (m & ~0x3Fu) != 0;
    return 63 - (int)(m & 0x3Fu);
}
// 1/x from the hardware reciprocal and Newton steps r <- r + r (1 - x r).
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}
__device__ __forceinline__ float recip(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
}
template <class T, int W>
__device__ __forceinline__ void bgj_diag_body(int b, const int* __restrict__ dims,
                                              const long long* __restrict__ off,
                                              const int* __restrict__ lds,
                                              const T* __restrict__ mats, T* Dws, int step,
                                              int ibase, int* info) {
    constexpr int kDiagCols = kGJB / W;
    const int n = dims[b];
    const int k0 = step * kGJB;
    if (k0 >= n || info[b]) return;
    const int kb = min(kGJB, n - k0);
    const int ld = lds[b];
    const T* A = mats + off[b] + k0 + (size_t)k0 * ld;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __shared__ T colbuf[2][kGJB];
    __shared__ int piv[kGJB];
    T a[kDiagCols];
#pragma unroll
    for (int q = 0; q < kDiagCols; ++q) {
        const int c = w + W * q;
        a[q] = (lane < kb && c < kb) ? A[lane + (size_t)c * ld] : (T)(c == lane ? 1 : 0);
    }
    if (w == 0) colbuf[0][lane] = a[0];
    __syncthreads();
    bool used = false;
    int my_step = 0;
    int bad = 0;
    // step k = kq W + kw: column k is register kq of wave kw, so with
    // kq unrolled every register index is a constant
#pragma unroll
    for (int kq = 0; kq < kDiagCols; ++kq) {
#pragma unroll 1
        for (int kw = 0; kw < W; ++kw) {
            const int k = kq * W + kw;
            const T f = colbuf[k & 1][lane];
            bool nz;
            const int p = __builtin_amdgcn_readfirstlane(
                wave_argmax_key(argmax_key(used ? (T)0 : fabs(f), lane), nz));
            if (!nz) {
                bad = k + 1;
                goto done;
            }
            const bool me = lane == p;
            const T d = recip(read_lane(f, p));
#pragma unroll
            for (int q = 0; q < kDiagCols; ++q) {
                const T prd = read_lane(a[q], p) * d;  // scaled pivot row
                a[q] = me ? prd : fma(-f, prd, a[q]);
            }
            if (w == kw) a[kq] = me ? d : -f * d;  // the pivot column itself
            if (me) {
                used = true;
                my_step = k;
                if (w == 0) piv[k] = lane;
            }
            // publish column k + 1
            if (kw + 1 < W) {
                if (w == kw + 1) colbuf[(k + 1) & 1][lane] = a[kq];
            } else if (kq + 1 < kDiagCols) {
                if (w == 0) colbuf[(k + 1) & 1][lane] = a[min(kq + 1, kDiagCols - 1)];
            }
            __syncthreads();
        }
    }
done:
    if (bad) {
        if (threadIdx.x == 0) info[b] = ibase + k0 + bad;
        return;
    }
    T* D = Dws + (size_t)b * kGJB * kGJB;
    if (my_step < kb) {
#pragma unroll
        for (int q = 0; q < kDiagCols; ++q) {
            const int c = w + W * q;
            const int pc = piv[c];
            if (c < kb && pc < kb) D[my_step + (size_t)pc * kGJB] = a[q];
        }
    }
    // rows and columns >= kb of the 64 x 64 workspace are zero
    for (int e = threadIdx.x; e < kGJB * kGJB; e += W * 64) {
        const int c = e >> 6, i = e & 63;
        if (c >= kb || i >= kb) D[e] = (T)0;
    }
}
template <class T>
__global__ void __launch_bounds__(kDiagWaves * 64) k_bgj_diag(const int* __restrict__ dims,
                                                             const long long* __restrict__ off,
                                                             const int* __restrict__ lds,
                                                             const T* __restrict__ mats, T* Dws,
                                                             int step, int ibase, int* info) {
    bgj_diag_body<T, kDiagWaves>(blockIdx.x, dims, off, lds, mats, Dws, step, ibase, info);
}

// acc(4x4 per thread) = sum_c L[c][row] * R[c][col] over a 64x64 tile.
template <class T>
__device__ __forceinline__ void tile_mm(const T (*L)[65], const T (*Rm)[65], int kb, T acc[4][4]) {
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    for (int c = 0; c < kb; ++c) {
        T a[4], bv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = L[c][tx + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; ++j) bv[j] = Rm[c][ty + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], bv[j], acc[i][j]);
    }
}

// R(:, j) = D A(K, j) for 64 columns j outside K.  Grid (column chunks, boxes).
template <class T>
__global__ void __launch_bounds__(256) k_bgj_row(const int* __restrict__ dims,
                                                 const long long* __restrict__ off,
                                                 const int* __restrict__ lds,
                                                 const T* __restrict__ mats,
                                                 const T* __restrict__ Dws, T* Rws,
                                                 const long long* __restrict__ r_off, int step,
                                                 const int* __restrict__ info) {
    const int b = blockIdx.y;
    const int n = dims[b];
    const int k0 = step * kGJB;
    if (k0 >= n || info[b]) return;
    const int kb = min(kGJB, n - k0);
    const int j0 = blockIdx.x * 64;
    if (j0 >= n) return;
    const int ld = lds[b];
    const T* A = mats + off[b];
    const T* D = Dws + (size_t)b * kGJB * kGJB;
    __shared__ T Ds[kGJB][65];  // Ds[c][i] = D(i, c)
    __shared__ T As[kGJB][65];  // As[c][j] = A(k0 + c, j0 + j)
    const int tid = threadIdx.x;
    for (int e = tid; e < kGJB * 64; e += 256) {
        const int i = e & 63, c = e >> 6;
        Ds[c][i] = D[i + (size_t)c * kGJB];
        const int cc = e & 63, jj = e >> 6;  // coalesced along K rows
        const int gj = j0 + jj;
        As[cc][jj] = (cc < kb && gj < n) ? A[k0 + cc + (size_t)gj * ld] : (T)0;
    }
    __syncthreads();
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (T)0;
    tile_mm<T>(Ds, As, kb, acc);
    const int tx = tid & 15, ty = tid >> 4;
    T* R = Rws + r_off[b];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int gj = j0 + ty + 16 * j;
        if (gj >= n) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int ii = tx + 16 * i;
            if (ii < kb) R[ii + (size_t)gj * kGJB] = acc[i][j];
        }
    }
}

// A(i, K) = -A(i, K) D for i outside K; A(K, K) = D.  Grid (row chunks, boxes).
// The same workgroup also sets A(K, j) = R(:, j) for the columns j of its
// chunk outside K (the two blocks are disjoint), which saves a launch per
// step.
template <class T>
__device__ __forceinline__ void bgj_col_body(int b, int chunk, const int* __restrict__ dims,
                                             const long long* __restrict__ off,
                                             const int* __restrict__ lds, T* mats,
                                             const T* __restrict__ Dws,
                                             const T* __restrict__ Rws,
                                             const long long* __restrict__ r_off, int step,
                                             const int* __restrict__ info) {
    const int n = dims[b];
    const int k0 = step * kGJB;
    if (k0 >= n || info[b]) return;
    const int kb = min(kGJB, n - k0);
    const int i0 = chunk * 64;
    if (i0 >= n) return;
    const int ld = lds[b];
    T* A = mats + off[b];
    if (i0 != k0) {
        // row copy: columns i0 .. i0 + 63 of the chunk
        const T* R = Rws + r_off[b];
        const int jn = min(64, n - i0);
        for (int e = threadIdx.x; e < kb * jn; e += 256) {
            const int c = e % kb, j = i0 + e / kb;
            A[k0 + c + (size_t)j * ld] = R[c + (size_t)j * kGJB];
        }
    }
    const T* D = Dws + (size_t)b * kGJB * kGJB;
    __shared__ T Ls[kGJB][65];  // Ls[c][i] = A(i0 + i, k0 + c)
    __shared__ T Dt[kGJB][65];  // Dt[c][j] = D(c, j)
    const int tid = threadIdx.x;
    for (int e = tid; e < kGJB * 64; e += 256) {
        const int i = e & 63, c = e >> 6;
        const int gi = i0 + i;
        Ls[c][i] = (gi < n && c < kb) ? A[gi + (size_t)(k0 + c) * ld] : (T)0;
        const int cc = e & 63, jj = e >> 6;
        Dt[cc][jj] = D[cc + (size_t)jj * kGJB];
    }
    __syncthreads();
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (T)0;
    tile_mm<T>(Ls, Dt, kb, acc);
    const int tx = tid & 15, ty = tid >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int jj = ty + 16 * j;
        if (jj >= kb) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int gi = i0 + tx + 16 * i;
            if (gi >= n) continue;
            T* e = A + gi + (size_t)(k0 + jj) * ld;
            if (gi >= k0 && gi < k0 + kb) *e = Dt[gi - k0][jj];
            else *e = -acc[i][j];
        }
    }
}
template <class T>
__global__ void __launch_bounds__(256) k_bgj_col(const int* __restrict__ dims,
                                                 const long long* __restrict__ off,
                                                 const int* __restrict__ lds, T* mats,
                                                 const T* __restrict__ Dws,
                                                 const T* __restrict__ Rws,
                                                 const long long* __restrict__ r_off, int step,
                                                 const int* __restrict__ info) {
    bgj_col_body<T>(blockIdx.y, blockIdx.x, dims, off, lds, mats, Dws, Rws, r_off, step, info);
}

// One launch for the column step of inner step col_step (< 0: none) and the
// diagonal block of step diag_step (< 0: none), 256 threads each: the column
// step writes block K's inverse and the column blocks A(I, K), I != K, and
// the row blocks A(K, J); the next diagonal block A(K', K') is none of them
// and was updated by the in-window GEMM launched before, so the two are
// independent.  Their 64 x 64 D workspaces are separate slots (Dcol, Ddiag).
// Workgroups [0, nchunk * nbig) are column chunks (chunk = id % nchunk, box
// = id / nchunk), the rest one diagonal block per box (four waves).
template <class T>
__global__ void __launch_bounds__(256) k_bgj_col_diag(int nchunk, int nbig,
                                                      const int* __restrict__ dims,
                                                      const long long* __restrict__ off,
                                                      const int* __restrict__ lds, T* mats,
                                                      const T* Dcol, T* Ddiag,
                                                      const T* __restrict__ Rws,
                                                      const long long* __restrict__ r_off,
                                                      int col_step, int diag_step, int ibase,
                                                      int* info) {
    const int id = (int)blockIdx.x;
    const int ncol = col_step >= 0 ? nchunk * nbig : 0;
    if (id < ncol) {
        bgj_col_body<T>(id / nchunk, id % nchunk, dims, off, lds, mats, Dcol, Rws, r_off,
                        col_step, info);
    } else {
        bgj_diag_body<T, 4>(id - ncol, dims, off, lds, mats, Ddiag, diag_step, ibase, info);
    }
}


constexpr int BM = 64, BN = 64, BK = 16;

// C(m x n) = alpha * op(A) * B + beta * C, column-major with per-box
// leading dimensions: op(A) = A (A stored m x k) or A^T (A stored k x m).
// One 64x64 output tile per workgroup over a flattened tile list.
template <class T, bool TRANSA, class TC = T>
__global__ void __launch_bounds__(256) k_gemm(int nbox, const int* __restrict__ tile_off,
                                              const GemmDesc* __restrict__ desc,
                                              const T* __restrict__ Ab, const T* __restrict__ Bb,
                                              TC* Cb, T alpha, T beta,
                                              const int* __restrict__ flags) {
    // a contiguous range of the tile list per XCD, as in k_gemm_mfma
    const int nwg = (int)gridDim.x, per = nwg >> 3, rem = nwg & 7;
    const int xcd = (int)blockIdx.x & 7;
    const int t = xcd * per + min(xcd, rem) + ((int)blockIdx.x >> 3);
    const int b = find_box(tile_off, nbox, t);
    const int local = t - tile_off[b];
    const GemmDesc dd = desc[b];
    const int m = dd.m, n = dd.n, k = dd.k;
    const int tm = (m + BM - 1) / BM;
    int i0, j0;
    if (dd.lower) {  // only the tiles on or below the diagonal are listed
        int J = 0, r = local;
        while (r >= tm - J) {
            r -= tm - J;
            ++J;
        }
        i0 = (J + r) * BM;
        j0 = J * BN;
    } else {
        i0 = (local % tm) * BM;
        j0 = (local / tm) * BN;
    }
    if (dd.flag >= 0 && flags && flags[dd.flag] == 0) return;
    const T* A = Ab + dd.offA;
    const T* B = Bb + dd.offB;
    TC* C = Cb + dd.offC;
    __shared__ T As[BK][BM + 1];
    __shared__ T Bs[BK][BN + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (T)0;
    // The next slab's operands are loaded into registers while the current
    // one is multiplied out of LDS.
    T ra[4], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = threadIdx.x + 256 * q;
            if (!TRANSA) {
                const int r = e & (BM - 1), c = e >> 6;  // A tile: 64 rows x 16 cols
                const int gi = i0 + r, gk = k0 + c;
                ra[q] = (gi < m && gk < k) ? A[gi + (size_t)gk * dd.lda] : (T)0;
            } else {
                const int c = e & (BK - 1), r = e >> 4;  // (A^T) tile from A(k0+c, i0+r)
                const int gi = i0 + r, gk = k0 + c;
                ra[q] = (gi < m && gk < k) ? A[gk + (size_t)gi * dd.lda] : (T)0;
            }
            const int r2 = e & (BK - 1), cb = e >> 4;  // B tile: 16 rows x 64 cols
            const int gkb = k0 + r2, gj = j0 + cb;
            rb[q] = (gkb < k && gj < n) ? B[gkb + (size_t)gj * dd.ldb] : (T)0;
        }
    };
    if (k > 0) gload(0);
    for (int k0 = 0; k0 < k; k0 += BK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = threadIdx.x + 256 * q;
            if (!TRANSA) As[e >> 6][e & (BM - 1)] = ra[q];
            else As[e & (BK - 1)][e >> 4] = ra[q];
            Bs[e & (BK - 1)][e >> 4] = rb[q];
        }
        __syncthreads();
        if (k0 + BK < k) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            T a[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], bv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int gj = j0 + ty + 16 * j;
        if (gj >= n) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int gi = i0 + tx + 16 * i;
            if (gi >= m) continue;
            TC* e = C + gi + (size_t)gj * dd.ldc;
            const TC v = (TC)(alpha * acc[i][j]);
            *e = (beta == (T)0) ? v : (TC)beta * *e + v;
        }
    }
}

// MFMA variant of k_gemm for the larger boxes.  Four waves in a 2x2 grid;
// each wave owns a (2 MT) x (2 MT) block of 2x2 MFMA tiles.
//   float : v_mfma_f32_32x32x2f32, MT = 32, workgroup tile 128 x 128
//   double: v_mfma_f64_16x16x4f64, MT = 16, workgroup tile  64 x  64
// Operand layout (lane l of a wave): A element (row l % MT, k l / MT),
// B element (k l / MT, col l % MT).  Accumulator element i of lane l is
// (row, col) = (8 (i / 4) + 4 (l / 32) + i % 4, l % 32) for 32x32x2 and
// (4 i + l / 16, l % 16) for the f64 16x16x4.
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// Matrix-core configuration per element type.  Each wave owns a
// (BLK*MT) x (BLK*MT) part of the 128 x 128 workgroup tile.  The products
// are issued transposed, D = op(B)^T op(A)^T = C^T per MT x MT block, so that
// the accumulator element i of lane l is C(row(l), col(i, l)) with the row
// running along the lanes: the epilogue then reads and writes whole column
// segments of C.
template <class T>
struct MfmaCfg;
template <>
struct MfmaCfg<float> {
    // v_mfma_f32_32x32x2f32: D row 8(i/4) + 4(l/32) + i%4, column l%32
    static constexpr int MT = 32, KT = 2, NACC = 16, BLK = 2;
    typedef floatx16 acc_t;
    __device__ static acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    __device__ static void pos(int i, int lane, int& r, int& c) {
        r = lane & 31;
        c = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
    }
};
template <>
struct MfmaCfg<double> {
    // v_mfma_f64_16x16x4f64: D row 4i + l/16, column l%16
    static constexpr int MT = 16, KT = 4, NACC = 4, BLK = 4;
    typedef doublex4 acc_t;
    __device__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static void pos(int i, int lane, int& r, int& c) {
        r = lane & 15;
        c = 4 * i + (lane >> 4);
    }
};


// Raw buffer loads: an offset at or past num_records bytes reads zero
// instead of faulting, which bounds the tile loads without branches.
constexpr int kBufWord3 = 0x00020000;  // raw (unformatted) buffer
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                             kBufWord3);
}
__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, unsigned off, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, unsigned off, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

// Split products (bf16x3): x = hi + lo with hi = bf16(x) and lo = bf16(x - hi),
// both rounded to nearest even; a b ~ hi_a hi_b + hi_a lo_b + lo_a hi_b, the
// lo_a lo_b term (2^-16 relative) dropped, summed in FP32 by
// v_mfma_f32_32x32x16_bf16.  A lane's 8 k of a slab are one operand of that
// instruction; its k order inside the instruction is the same for both
// operands, so any consistent assignment sums every k once.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ void split_bf16(const float* x, bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = (__bf16)x[j];
        hi[j] = h;
        lo[j] = (__bf16)(x[j] - (float)h);
    }
}

// LDS row pitch (elements) of a k-slab row: rows are read 32 bytes per lane
// (its 4 doubles or 8 floats of k), and the pitch spreads 16 consecutive rows
// over all 64 banks.
template <class T>
struct SlabPitch;
template <>
struct SlabPitch<double> {
    static constexpr int value = 18;  // 144 B: 36 words, 16 rows hit 16 distinct bank quads
};
template <>
struct SlabPitch<float> {
    static constexpr int value = 20;  // 80 B: 20 words, idem
};

// Tile rows per group in the order of a general descriptor's output tiles.
constexpr int kGemmGroupRows = 8;

// Workgroups per CU the matrix-core GEMM is compiled for.  The FP32
// 128 x 128 tile uses 40 KB of LDS, so four fit the CU's 160 KB; asking for
// four holds it to 128 VGPRs (four waves per SIMD instead of three).  The
// split-product variant needs more registers and keeps two.
template <class T, int TL, bool SPLIT>
constexpr int kGemmMinBlocks = (sizeof(T) == 4 && TL == 128 && !SPLIT) ? 4 : 2;

// C(m x n) = alpha op(A) B + beta C over a flattened list of 128 x 128 output
// tiles of every descriptor, 4 waves in a 2 x 2 arrangement.  The k dimension
// is walked in slabs of 16 held in two LDS buffers: the next slab is loaded
// into registers while the current one feeds the matrix cores, so global
// latency overlaps the products and each slab costs one barrier.
//
// Both operands sit in LDS as [row][k] (k contiguous).  The matrix-core
// instruction sums over its KT k-slots, one per lane group g = lane / MT;
// slot g of step s is mapped to k = g SPL + s (SPL = 16 / KT steps), so each
// lane reads its whole slab share of a row, SPL consecutive k, with two
// 128-bit LDS reads.  Tile rows and columns past m and n read whatever lies
// within the operand's buffer (those outputs are not stored); k past the
// end reads zero, from the buffer bound or an explicit select where the
// operand is k-contiguous.  The host keeps each operand under 2 GiB.
// SPLIT (float only): each slab's products are three bf16 matrix-core
// products of the split operands (split_bf16) instead of eight FP32 ones.
// TL: workgroup tile, 128 (the default) or 64.
template <class T, bool TRANSA, class TC = T, bool SPLIT = false, int TL = kMfmaTile>
__global__ void __launch_bounds__(256, (kGemmMinBlocks<T, TL, SPLIT>)) k_gemm_mfma(int nbox, const int* __restrict__ tile_off,
                                                      const GemmDesc* __restrict__ desc,
                                                      const T* __restrict__ Ab,
                                                      const T* __restrict__ Bb, TC* Cb, T alpha,
                                                      T beta, const int* __restrict__ flags,
                                                      int contig) {
    typedef MfmaCfg<T> C;
    constexpr int MT = C::MT, KT = C::KT, BLK = TL / (2 * MT), WT = BLK * MT, TILE = 2 * WT,
                  KB = 16;
    constexpr int SPL = KB / KT;           // matrix-core steps per slab
    constexpr int LDK = SlabPitch<T>::value;
    constexpr int NLD = TILE * KB / 256;  // elements per thread and operand per slab
    static_assert(TILE == TL && BLK >= 1 && TILE % 16 == 0, "tile");
    static_assert(SPL * sizeof(T) == 32, "a lane's slab share is 32 bytes");
    // Workgroups are dealt to the eight XCDs round-robin.  With contig each
    // XCD is given a contiguous range of the tile list (column-major tiles of
    // a box), so the tiles sharing an operand panel run in one XCD's L2.
    int t = (int)blockIdx.x;
    if (contig) {
        const int nwg = (int)gridDim.x, per = nwg >> 3, rem = nwg & 7;
        const int xcd = (int)blockIdx.x & 7;
        t = xcd * per + min(xcd, rem) + ((int)blockIdx.x >> 3);
    }
    const int b = find_box(tile_off, nbox, t);
    const int local = t - tile_off[b];
    const GemmDesc dd = desc[b];
    const int m = dd.m, n = dd.n, k = dd.k;
    const int tm = (m + TILE - 1) / TILE;
    int i0, j0;
    if (dd.lower) {
        // lower-triangle descriptors list only the tiles on or below the
        // diagonal, column after column (column J holds rows J .. tm - 1)
        int J = 0, r = local;
        while (r >= tm - J) {
            r -= tm - J;
            ++J;
        }
        i0 = (J + r) * TILE;
        j0 = J * TILE;
    } else {
        // groups of kGemmGroupRows tile rows, column-major within a group:
        // the tiles an XCD runs at once cover a few rows and a few columns
        // of the box, so they share their A row panels as well as their B
        // column panels in that XCD's L2
        const int tn = (n + TILE - 1) / TILE;
        const int gsz = kGemmGroupRows * tn, g = local / gsz, first = g * kGemmGroupRows;
        const int rows = min(kGemmGroupRows, tm - first), w = local - g * gsz;
        i0 = (first + w % rows) * TILE;
        j0 = (w / rows) * TILE;
    }
    if (dd.flag >= 0 && flags && flags[dd.flag] == 0) return;
    TC* Cm = Cb + dd.offC;
    // operand extents in bytes: A is m x k (lda) or k x m (lda, TRANSA), B is k x n (ldb)
    const __amdgpu_buffer_rsrc_t ra_rsrc =
        make_rsrc(Ab + dd.offA, (unsigned)((size_t)dd.lda * (TRANSA ? m : k) * sizeof(T)));
    const __amdgpu_buffer_rsrc_t rb_rsrc =
        make_rsrc(Bb + dd.offB, (unsigned)((size_t)dd.ldb * n * sizeof(T)));
    __shared__ __attribute__((aligned(16))) T As[2][TILE][LDK];
    __shared__ __attribute__((aligned(16))) T Bs[2][TILE][LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave & 1) * WT, wc = (wave >> 1) * WT;
    // Per-thread load slots.  k-contiguous operands (B, and A when TRANSA):
    // 16 lanes cover the 16 k of one row, e = tid + 256 q -> (k e % 16,
    // row e / 16).  A column-major with rows contiguous: a wave's load covers
    // 16 rows x 4 k (four 128-byte column segments), slot q of wave w and
    // lane l holding (row 16 q' + l % 16, k 4 w' + l / 16) with (q', w') the
    // slot's (row block, k block); that shape also lands each half-wave's LDS
    // write on 64 distinct banks.
    auto a_row = [&](int q) { return 16 * ((q * 4 + wave) % (TILE / 16)) + (lane & 15); };
    auto a_k = [&](int q) { return 4 * ((q * 4 + wave) / (TILE / 16)) + (lane >> 4); };
    // Each thread keeps one byte offset per operand (its slot q = 0) and
    // the slab's running offset; slot q adds a workgroup-uniform distance
    // (16 rows of a k-contiguous operand per slot; for A column-major, 64
    // rows for odd q and 4 k per pair of slots in the 128 tile, 4 k per slot
    // in the 64 tile).  All of it stays in the buffer offset, which the
    // bound check covers.
    static_assert(TILE == 128 || TILE == 64, "slot distances assume 64 or 128 tiles");
    const unsigned sz = sizeof(T);
    const unsigned offa0 =
        !TRANSA ? (unsigned)(((i0 + a_row(0)) + (size_t)a_k(0) * dd.lda) * sz)
                : (unsigned)(((tid % KB) + (size_t)(i0 + tid / KB) * dd.lda) * sz);
    const unsigned offb0 = (unsigned)(((tid % KB) + (size_t)(j0 + tid / KB) * dd.ldb) * sz);
    const unsigned lda_b = (unsigned)dd.lda * sz, ldb_b = (unsigned)dd.ldb * sz;
    auto a_delta = [&](int q) -> unsigned {
        if (TRANSA) return (unsigned)(q * (256 / KB)) * lda_b;
        if (TILE == 128) return (unsigned)(64 * (q & 1)) * sz + (unsigned)(4 * (q >> 1)) * lda_b;
        return (unsigned)(4 * q) * lda_b;
    };
    const unsigned stepa = TRANSA ? KB * sz : KB * lda_b;
    const unsigned stepb = KB * sz;
    unsigned va = offa0, vb = offb0;  // slot-0 offsets of the slab being loaded
    T ra[NLD], rb[NLD];
    auto gload = [&](int sl) {
        const int k0 = sl * KB;
        const bool tail = k0 + KB > k;  // uniform: only the last slab
#pragma unroll
        for (int q = 0; q < NLD; ++q) {
            ra[q] = buf_load(ra_rsrc, va + a_delta(q), T());
            rb[q] = buf_load(rb_rsrc, vb + (unsigned)(q * (256 / KB)) * ldb_b, T());
        }
        va += stepa;
        vb += stepb;
        if (tail && k0 + tid % KB >= k) {
            // k-contiguous operands read the next rows of their column past
            // k (a thread's k is the same in every slot): zero them (A
            // column-major reads zero from the bound)
#pragma unroll
            for (int q = 0; q < NLD; ++q) {
                rb[q] = (T)0;
                if (TRANSA) ra[q] = (T)0;
            }
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < NLD; ++q) {
            const int e = tid + 256 * q;
            if (!TRANSA) As[buf][a_row(q)][a_k(q)] = ra[q];
            else As[buf][e / KB][e % KB] = ra[q];
            Bs[buf][e / KB][e % KB] = rb[q];
        }
    };
    typename C::acc_t acc[BLK][BLK];
#pragma unroll
    for (int a = 0; a < BLK; ++a)
#pragma unroll
        for (int c = 0; c < BLK; ++c)
#pragma unroll
            for (int i = 0; i < C::NACC; ++i) acc[a][c][i] = (T)0;
    const int nsl = (k + KB - 1) / KB;  // k == 0: no loads, C = beta C
    if (nsl > 0) {
        gload(0);
        sstore(0);
    }
    __syncthreads();
    const int g = lane / MT, rr = lane % MT;
    for (int sl = 0; sl < nsl; ++sl) {
        const int cur = sl & 1;
        if (sl + 1 < nsl) gload(sl + 1);
        T bv[BLK][SPL];
#pragma unroll
        for (int c = 0; c < BLK; ++c) {
            const T* src = &Bs[cur][wc + c * MT + rr][g * SPL];
#pragma unroll
            for (int s2 = 0; s2 < SPL; ++s2) bv[c][s2] = src[s2];
        }
        bf16x8 bh[SPLIT ? BLK : 1], bl[SPLIT ? BLK : 1];
        if constexpr (SPLIT) {
            static_assert(sizeof(T) == 4 && SPL == 8, "split products: float slabs");
#pragma unroll
            for (int c = 0; c < BLK; ++c) split_bf16(bv[c], bh[c], bl[c]);
        }
        // The next slab goes to the other LDS buffer halfway through the
        // products (its loads have landed by then), so the slab ends with
        // the barrier alone instead of a store-and-wait.
#pragma unroll
        for (int a = 0; a < BLK; ++a) {
            if (a == BLK / 2 && sl + 1 < nsl) {
                __builtin_amdgcn_sched_barrier(0);
                sstore(cur ^ 1);
                __builtin_amdgcn_sched_barrier(0);
            }
            T av[SPL];
            const T* src = &As[cur][wr + a * MT + rr][g * SPL];
#pragma unroll
            for (int s2 = 0; s2 < SPL; ++s2) av[s2] = src[s2];
            if constexpr (SPLIT) {
                bf16x8 ah, al;
                split_bf16(av, ah, al);
#pragma unroll
                for (int c = 0; c < BLK; ++c) {
                    acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bl[c], ah, acc[a][c], 0, 0, 0);
                    acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[c], al, acc[a][c], 0, 0, 0);
                    acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[c], ah, acc[a][c], 0, 0, 0);
                }
            } else {
#pragma unroll
                for (int s2 = 0; s2 < SPL; ++s2)
#pragma unroll
                    for (int c = 0; c < BLK; ++c) acc[a][c] = C::mma(bv[c][s2], av[s2], acc[a][c]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < BLK; ++a)
#pragma unroll
        for (int c = 0; c < BLK; ++c)
#pragma unroll
            for (int i = 0; i < C::NACC; ++i) {
                int r, cc;
                C::pos(i, lane, r, cc);
                const int gi = i0 + wr + a * MT + r, gj = j0 + wc + c * MT + cc;
                if (gi < m && gj < n) {
                    TC* e = Cm + gi + (size_t)gj * dd.ldc;
                    const TC v = (TC)(alpha * acc[a][c][i]);
                    *e = (beta == (T)0) ? v : (TC)beta * *e + v;
                }
            }
}

__global__ void k_fill_int(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fold_info(int nbox, const int* __restrict__ info,
                                                   int* status) {
    __shared__ int best[256];
    int b0 = INT32_MAX;
    for (int b = threadIdx.x; b < nbox; b += 256)
        if (info[b] != 0) {
            b0 = b;
            break;
        }
    best[threadIdx.x] = b0;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) best[threadIdx.x] = min(best[threadIdx.x], best[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && best[0] != INT32_MAX && *status == 0) *status = info[best[0]];
}

__global__ void __launch_bounds__(256) k_count_nonzero(int n, const int* __restrict__ a,
                                                       int* count) {
    __shared__ int part[256];
    int c = 0;
    for (int i = threadIdx.x; i < n; i += 256) c += a[i] != 0;
    part[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = part[0];
}

// info[bix[j]] = cinfo[j]: blocked-path status back to the caller's boxes.
__global__ void k_scatter_info(int nbig, const int* __restrict__ bix,
                               const int* __restrict__ cinfo, int* info) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nbig) info[bix[j]] = cinfo[j];
}

// Strided block copies dst(rows x cols, ldd) = src(rows x cols, lds) over a
// descriptor list; src is src1 if sel else src0.  Grid (descriptors, chunks of
// 4 columns).
template <class T>
__global__ void __launch_bounds__(256) k_copy2d(const Copy2D* __restrict__ desc,
                                                const T* __restrict__ src0,
                                                const T* __restrict__ src1, T* dst) {
    const Copy2D d = desc[blockIdx.x];
    const T* src = d.sel ? src1 : src0;
    const int c0 = blockIdx.y * 4;
    if (c0 >= d.cols) return;
    const int c1 = min(c0 + 4, d.cols);
    for (int c = c0; c < c1; ++c)
        for (int i = threadIdx.x; i < d.rows; i += 256)
            dst[d.dst + i + (size_t)c * d.ldd] = src[d.src + i + (size_t)c * d.lds];
}

// SymCopy descriptors over a flattened list of 64 x 64 destination tiles.
// Transposed tiles pass through LDS so that both the reads and the writes
// run along columns.
template <class T>
__global__ void __launch_bounds__(256) k_sym_copy(int ndesc, const long long* __restrict__ toff,
                                                  const SymCopy* __restrict__ desc, T* A, T* SW) {
    const long long blk = blockIdx.x;
    const int b = find_box(toff, ndesc, blk);
    const SymCopy d = desc[b];
    const int tm = (d.rows + 63) / 64;
    const int local = (int)(blk - toff[b]);
    const int i0 = (local % tm) * 64, j0 = (local / tm) * 64;
    if (d.upper == 1 && i0 >= j0 + 63) return;  // no element with i < j
    if (d.upper == 2 && j0 >= i0 + 63) return;  // no element with i > j
    T* dst = (d.dsel ? SW : A) + d.dst;
    const T* src = (d.ssel ? SW : A) + d.src;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    __shared__ T t[64][65];
    if (d.trans) {
        // t[i - i0][j - j0] = src(j, i), read along src's columns
        for (int c = ty; c < 64; c += 4) {
            const int j = j0 + tx, i = i0 + c;
            t[c][tx] = (j < d.cols && i < d.rows) ? src[j + (size_t)i * d.lds] : (T)0;
        }
        __syncthreads();
    }
    if (d.avg == 2) {
        // symmetrize in place (dst == src, trans, upper == 1): both triangles
        // take (a_ij + a_ji) / 2.  Each thread leaves its mean in the LDS
        // slot it read, and the lower triangle is written from there so
        // that its writes run along columns too.
        for (int c = ty; c < 64; c += 4) {
            const int i = i0 + tx, j = j0 + c;
            if (i >= d.rows || j >= d.cols || i >= j) continue;
            T* e = dst + i + (size_t)j * d.ldd;
            const T v = (T)0.5 * (*e + t[tx][c]);
            *e = v;
            t[tx][c] = v;
        }
        __syncthreads();
        for (int c = ty; c < 64; c += 4) {
            const int j = j0 + tx, i = i0 + c;  // lower element (j, i), j > i
            if (i >= d.rows || j >= d.cols || i >= j) continue;
            dst[j + (size_t)i * d.ldd] = t[c][tx];
        }
        return;
    }
    for (int c = ty; c < 64; c += 4) {
        const int i = i0 + tx, j = j0 + c;
        if (i >= d.rows || j >= d.cols || (d.upper == 1 && i >= j) || (d.upper == 2 && i <= j))
            continue;
        const T v = d.trans ? t[tx][c] : src[i + (size_t)j * d.lds];
        const int sg = j >= d.neg_from ? -d.s0 : d.s0;
        T* e = dst + i + (size_t)j * d.ldd;
        const T w = sg > 0 ? v : -v;
        *e = d.avg ? (T)0.5 * (*e + w) : w;
    }
}

// dst(box) = (D) src(rows r0.., cols c0.. of a ld-strided block); dst packed
// column-major with ld = rows.
template <class D>
__global__ void k_copy_blocks(int nbox, const long long* __restrict__ dst_off,
                              const CopyDesc* __restrict__ desc, const double* __restrict__ src,
                              D* __restrict__ dst, long long total) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; e < total; e += stride) {
        const int b = find_box(dst_off, nbox, e);
        const CopyDesc d = desc[b];
        const long long l = e - dst_off[b];
        const int i = (int)(l % d.rows), j = (int)(l / d.rows);
        dst[e] = (D)src[d.off + d.r0 + i + (size_t)(d.c0 + j) * d.ld];
    }
}

// Cp = C(perm, perm) for each face (n x n, ld n).
__global__ void k_permute_sym(int nbox, const long long* __restrict__ off,
                              const int* __restrict__ dn, const int* __restrict__ perm,
                              const int* __restrict__ perm_off, const double* __restrict__ C,
                              double* __restrict__ Cp, long long total) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; e < total; e += stride) {
        const int b = find_box(off, nbox, e);
        const int n = dn[b];
        const long long l = e - off[b];
        const int i = (int)(l % n), j = (int)(l / n);
        const int* p = perm + perm_off[b];
        Cp[e] = C[off[b] + p[i] + (size_t)p[j] * n];
    }
}

// Interpolative decomposition through the Gram matrix G = M^T M (n x n) of a
// tall face matrix, of which only the lower triangle is read.  Pivoted
// Cholesky of G chooses, step by step, the column with the largest remaining
// squared norm -- the column-pivoted QR rule -- and its rows are the rows of
// that QR's R (up to signs).  It stops at the first pivot with
// |R_ii| <= eps |R_00|; the remaining columns are redundant.
//
// Blocked as LAPACK's dpstrf: a panel of kGramNB pivots is chosen by one
// workgroup per face (left-looking inside the panel: row i of R is the pivot
// row of the trailing matrix minus the panel's earlier rows), then the
// trailing matrix takes the panel's rank-kGramNB update as a batched GEMM.
// The pivot order, remaining diagonal and stop state persist in global
// memory between panels.  Row i of R is stored at R[i n + col]; the panel's
// rows are also kept column by column in P (P[l + col kGramNB]) for the GEMM.
// Dynamic shared memory: d[n] (double), jp[n] (int).
__global__ void __launch_bounds__(256) k_gram_panel(
    const int* __restrict__ dn, const unsigned char* __restrict__ mode, const double* Gb,
    const long long* __restrict__ offG, double* Rb, double* Pb, const long long* __restrict__ offP,
    double* dgb, int* perm, const int* __restrict__ perm_off, int* kout, int* active, double* tolv,
    double eps, int j0) {
    constexpr int NB = kGramNB;
    const int b = blockIdx.x;
    if (!mode[b] || !active[b]) return;
    const int n = dn[b];
    const double* G = Gb + offG[b];
    double* R = Rb + offG[b];
    double* P = Pb + offP[b];
    double* dg = dgb + perm_off[b];
    int* jpg = perm + perm_off[b];
    extern __shared__ unsigned char smem[];
    double* d = reinterpret_cast<double*>(smem);
    int* jp = reinterpret_cast<int*>(d + n);
    __shared__ double redv[4];
    __shared__ int redi[4];
    __shared__ double s_pv, s_tol;
    __shared__ int s_pi, s_stop;
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int j = tid; j < n; j += bs) {
        d[j] = j0 == 0 ? G[j + (size_t)j * n] : dg[j];
        jp[j] = j0 == 0 ? j : jpg[j];
    }
    if (tid == 0) {
        s_stop = -1;
        s_tol = j0 == 0 ? 0.0 : tolv[b];
    }
    __syncthreads();
    const int i1 = min(j0 + NB, n);
    for (int i = j0; i < i1; ++i) {
        double best = -1.0;
        int bi = n;
        for (int pos = i + tid; pos < n; pos += bs) {
            const double v = d[jp[pos]];
            if (v > best) { best = v; bi = pos; }
        }
        block_argmax(best, bi, redv, redi, &s_pv, &s_pi);
        const int pos = s_pi;
        const double rii = sqrt(s_pv > 0.0 ? s_pv : 0.0);
        if (i == 0 && tid == 0) s_tol = rii * eps;
        __syncthreads();
        if (!(rii > s_tol)) {
            if (tid == 0) s_stop = i;
            break;
        }
        if (tid == 0 && pos != i) {
            const int t = jp[pos];
            jp[pos] = jp[i];
            jp[i] = t;
        }
        __syncthreads();
        const int p = jp[i], li = i - j0;
        const double* pp = P + (size_t)p * NB;
        const double inv = 1.0 / rii;
        for (int q = i + 1 + tid; q < n; q += bs) {
            const int c = jp[q];
            // G holds its lower triangle: G(p, c) = G(max, min)
            double v = p >= c ? G[p + (size_t)c * n] : G[c + (size_t)p * n];
            double* pc = P + (size_t)c * NB;
            for (int l = 0; l < li; ++l) v -= pp[l] * pc[l];
            v *= inv;
            R[(size_t)i * n + c] = v;
            pc[li] = v;
            const double nd = d[c] - v * v;
            d[c] = nd > 0.0 ? nd : 0.0;
        }
        // the panel row is zero on the columns pivoted so far, rii on its own
        for (int q = tid; q <= i; q += bs) P[li + (size_t)jp[q] * NB] = (q == i) ? rii : 0.0;
        if (tid == 0) R[(size_t)i * n + p] = rii;
        __syncthreads();
    }
    __syncthreads();
    for (int j = tid; j < n; j += bs) {
        dg[j] = d[j];
        jpg[j] = jp[j];
    }
    if (tid == 0) {
        tolv[b] = s_tol;
        if (s_stop >= 0 || i1 == n) {
            kout[b] = s_stop >= 0 ? s_stop : n;
            active[b] = 0;
        }
    }
}

// The same panel with the panel rows held in registers: thread t owns columns
// c = t + u BS (u < CPT) and keeps their kGramNB panel entries and remaining
// diagonal there, so a pivot step reads only G(p, c) from memory and the
// pivot column's earlier entries from LDS.  The arithmetic per column is the
// same sequence of operations as in k_gram_panel, so both give the same bits.
// Pivot choice: the largest remaining diagonal, ties to the lowest position.
// Faces need n <= BS CPT.  Dynamic shared memory: jp[n], pinv[n] (int).
template <int BS, int CPT>
__global__ void __launch_bounds__(BS) k_gram_panel_reg(
    const int* __restrict__ dn, const unsigned char* __restrict__ mode, const double* Gb,
    const long long* __restrict__ offG, double* Rb, double* Pb, const long long* __restrict__ offP,
    double* dgb, int* perm, const int* __restrict__ perm_off, int* kout, int* active, double* tolv,
    double eps, int j0) {
    constexpr int NB = kGramNB;
    constexpr int NW = BS / kWave;
    const int b = blockIdx.x;
    if (!mode[b] || !active[b]) return;
    const int n = dn[b];
    const double* G = Gb + offG[b];
    double* R = Rb + offG[b];
    double* P = Pb + offP[b];
    double* dg = dgb + perm_off[b];
    int* jpg = perm + perm_off[b];
    extern __shared__ unsigned char smem[];
    int* jp = reinterpret_cast<int*>(smem);
    int* pinv = jp + n;
    __shared__ double prow[NB];
    __shared__ double redv[NW];
    __shared__ int redk[NW];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    for (int q = tid; q < n; q += BS) {
        const int c = j0 == 0 ? q : jpg[q];
        jp[q] = c;
        pinv[c] = q;
    }
    __syncthreads();
    double pc[CPT][NB];
    double d[CPT];
    bool live[CPT];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int c = tid + u * BS;
        live[u] = c < n && pinv[c] >= j0;
        d[u] = c < n ? (j0 == 0 ? G[c + (size_t)c * n] : dg[c]) : 0.0;
#pragma unroll
        for (int l = 0; l < NB; ++l) pc[u][l] = 0.0;
    }
    double tol = j0 == 0 ? 0.0 : tolv[b];
    int stop = -1;
    const int i1 = min(j0 + NB, n);
#pragma unroll 1
    for (int li = 0; li < NB; ++li) {
        const int i = j0 + li;
        if (i >= i1) break;
        // candidate: largest d, then lowest position; key = position << 16 | column
        double best = -1.0;
        unsigned key = 0xffffffffu;
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int c = tid + u * BS;
            if (!live[u]) continue;
            const unsigned k = ((unsigned)pinv[c] << 16) | (unsigned)c;
            if (d[u] > best || (d[u] == best && k < key)) {
                best = d[u];
                key = k;
            }
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const double ov = __shfl_xor(best, o, kWave);
            const unsigned ok = (unsigned)__shfl_xor((int)key, o, kWave);
            if (ov > best || (ov == best && ok < key)) {
                best = ov;
                key = ok;
            }
        }
        if (lane == 0) {
            redv[wid] = best;
            redk[wid] = (int)key;
        }
        __syncthreads();
        best = redv[0];
        key = (unsigned)redk[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) {
            const double ov = redv[w];
            const unsigned ok = (unsigned)redk[w];
            if (ov > best || (ov == best && ok < key)) {
                best = ov;
                key = ok;
            }
        }
        const double rii = sqrt(best > 0.0 ? best : 0.0);
        if (i == 0) tol = rii * eps;
        if (!(rii > tol)) {
            stop = i;
            break;
        }
        const int pos = (int)(key >> 16), p = (int)(key & 0xffffu);
        // G(p, c) from the lower triangle, issued before the pivot row is shared
        double g[CPT];
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int c = tid + u * BS;
            g[u] = 0.0;
            if (live[u] && c != p) g[u] = p >= c ? G[p + (size_t)c * n] : G[c + (size_t)p * n];
        }
        // The pivot column's owner shares its panel entries (zero from
        // position li on, so the sums below may run over the whole panel).
        if (tid == (p % BS)) {
            const int up = p / BS;
#pragma unroll
            for (int u = 0; u < CPT; ++u) {
                if (u != up) continue;
#pragma unroll
                for (int l = 0; l < NB; ++l) {
                    prow[l] = pc[u][l];
                    if (l == li) pc[u][l] = rii;
                }
                live[u] = false;
            }
            R[(size_t)i * n + p] = rii;
        }
        if (tid == 0) {
            const int c = jp[i];
            jp[pos] = c;
            pinv[c] = pos;
            jp[i] = p;
            pinv[p] = i;
        }
        __syncthreads();
        const double inv = 1.0 / rii;
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            if (!live[u]) continue;
            const int c = tid + u * BS;
            double v = g[u];
            // terms l >= li are 0 * 0 and leave v unchanged
#pragma unroll
            for (int l = 0; l < NB; ++l) v -= prow[l] * pc[u][l];
            v *= inv;
            R[(size_t)i * n + c] = v;
#pragma unroll
            for (int l = 0; l < NB; ++l)
                if (l == li) pc[u][l] = v;
            const double nd = d[u] - v * v;
            d[u] = nd > 0.0 ? nd : 0.0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int c = tid + u * BS;
        if (c >= n) continue;
        dg[c] = d[u];
        double* pcol = P + (size_t)c * NB;
#pragma unroll
        for (int l = 0; l < NB; ++l) pcol[l] = pc[u][l];
    }
    for (int q = tid; q < n; q += BS) jpg[q] = jp[q];
    if (tid == 0) {
        tolv[b] = tol;
        if (stop >= 0 || i1 == n) {
            kout[b] = stop >= 0 ? stop : n;
            active[b] = 0;
        }
    }
}

// [R11 | R12] (k x n, ld k) of each Gram face from the rows of R in pivot
// order: R11(i, j) = R(i, jp[j]) for j >= i, zero below the diagonal.
__global__ void k_gram_gather_r(int nface, const long long* __restrict__ toff,
                                const GramRDesc* __restrict__ desc, const double* __restrict__ Rb,
                                const int* __restrict__ perm, double* out, long long total) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int f = find_box(toff, nface, e);
        const GramRDesc d = desc[f];
        const long long l = e - toff[f];
        const int i = (int)(l % d.k), j = (int)(l / d.k);
        const int c = perm[d.perm + j];
        out[d.out + l] = (j < d.k && j < i) ? 0.0 : Rb[d.r + (size_t)i * d.n + c];
    }
}

// Column-pivoted QR + interpolative decomposition of one face per workgroup.
// Dynamic shared memory: vn1[n], vn2[n] (double), jp[n], pos[n] (int).
// LDSM: the face matrix is staged in LDS (after the norm and index arrays)
// and the factorization runs there, then R is written back; the operations
// and their order are the same either way.
template <bool LDSM>
__global__ void __launch_bounds__(256) k_interp_decomp(const int* __restrict__ dm,
                                                       const int* __restrict__ dn,
                                                       const long long* __restrict__ offM,
                                                       double* Mb, double eps, int* perm,
                                                       const int* __restrict__ perm_off,
                                                       int* kout, double* Wb,
                                                       const long long* __restrict__ offW,
                                                       const unsigned char* __restrict__ mode,
                                                       int* jpvt_out, double* tau_out) {
    const int b = blockIdx.x;
    if (mode && mode[b]) return;
    const int m = dm[b], n = dn[b];
    if (n == 0) {
        if (threadIdx.x == 0) kout[b] = 0;
        return;
    }
    const int lda = m > 1 ? m : 1;
    double* Mg = Mb + offM[b];
    extern __shared__ unsigned char smem[];
    double* vn1 = reinterpret_cast<double*>(smem);
    double* vn2 = vn1 + n;
    int* jp = reinterpret_cast<int*>(vn2 + n);
    int* pos = jp + n;
    double* M = Mg;
    const int mtot = m > 0 ? lda * n : 0;
    if constexpr (LDSM) {
        M = reinterpret_cast<double*>(pos + n);
        for (int e = threadIdx.x; e < mtot; e += blockDim.x) M[e] = Mg[e];
        __syncthreads();
    }
    __shared__ double red[16];
    __shared__ double redv[16];
    __shared__ int redi[16];
    __shared__ double s_pv, s_tau;
    __shared__ int s_pi, s_k;
    const int tid = threadIdx.x, bs = blockDim.x;
    const int lane = tid & (kWave - 1), wid = tid / kWave, nw = (bs + kWave - 1) / kWave;
    const double tol3z = sqrt(0x1p-53);

    for (int j = wid; j < n; j += nw) {
        double s = 0.0;
        for (int r = lane; r < m; r += kWave) {
            double v = M[r + (size_t)j * lda];
            s += v * v;
        }
        s = wave_sum(s);
        if (lane == 0) {
            vn1[j] = vn2[j] = sqrt(s);
            jp[j] = j;
        }
    }
    __syncthreads();
    const int mn = m < n ? m : n;
    for (int i = 0; i < mn; ++i) {
        // pivot: first index of the largest partial norm
        double best = -1.0;
        int bi = n;
        for (int j = i + tid; j < n; j += bs)
            if (vn1[j] > best) { best = vn1[j]; bi = j; }
        block_argmax(best, bi, redv, redi, &s_pv, &s_pi);
        const int pvt = s_pi;
        if (pvt != i) {
            for (int r = tid; r < m; r += bs) {
                double t = M[r + (size_t)pvt * lda];
                M[r + (size_t)pvt * lda] = M[r + (size_t)i * lda];
                M[r + (size_t)i * lda] = t;
            }
            __syncthreads();
            if (tid == 0) {
                int t = jp[pvt]; jp[pvt] = jp[i]; jp[i] = t;
                vn1[pvt] = vn1[i];
                vn2[pvt] = vn2[i];
            }
        }
        // Householder reflector for column i, rows i..m-1
        double* col = M + (size_t)i * lda;
        double ss = 0.0;
        for (int r = i + 1 + tid; r < m; r += bs) ss += col[r] * col[r];
        double xn2 = block_sum(ss, red);
        if (tid == 0) {
            double tau = 0.0;
            if (i < m - 1 && xn2 > 0.0) {
                double alpha = col[i];
                double xnorm = sqrt(xn2);
                double beta = -copysign(hypot(alpha, xnorm), alpha);
                tau = (beta - alpha) / beta;
                s_pv = 1.0 / (alpha - beta);  // scale for v
                col[i] = beta;
            }
            s_tau = tau;
        }
        __syncthreads();
        const double tau = s_tau;
        if (tau_out && tid == 0) tau_out[perm_off[b] + i] = tau;
        if (tau != 0.0) {
            const double scal = s_pv;
            for (int r = i + 1 + tid; r < m; r += bs) col[r] *= scal;
        }
        __syncthreads();
        // apply H to trailing columns and downdate partial norms (one wave per column)
        for (int j = i + 1 + wid; j < n; j += nw) {
            double* cj = M + (size_t)j * lda;
            double mij = cj[i];
            double sq = 0.0;
            if (tau != 0.0) {
                double d = 0.0;
                for (int r = i + 1 + lane; r < m; r += kWave) d += cj[r] * col[r];
                d = wave_sum(d) + mij;
                double f = tau * d;
                for (int r = i + 1 + lane; r < m; r += kWave) {
                    double v = cj[r] - f * col[r];
                    cj[r] = v;
                    sq += v * v;
                }
                mij -= f;
                if (lane == 0) cj[i] = mij;
            } else {
                for (int r = i + 1 + lane; r < m; r += kWave) sq += cj[r] * cj[r];
            }
            sq = wave_sum(sq);
            if (lane == 0 && vn1[j] != 0.0) {
                double temp = fabs(mij) / vn1[j];
                temp = 1.0 - temp * temp;
                temp = temp > 0.0 ? temp : 0.0;
                double q = vn1[j] / vn2[j];
                double temp2 = temp * q * q;
                if (temp2 <= tol3z) {
                    double nv = (i < m - 1) ? sqrt(sq) : 0.0;
                    vn1[j] = nv;
                    vn2[j] = nv;
                } else {
                    vn1[j] *= sqrt(temp);
                }
            }
        }
        __syncthreads();
    }
    // Rank rule: column i is a skeleton iff |R(i,i)| > eps * |R(0,0)|.  For
    // i >= m the reference reads element i*(lda+1) of the buffer.
    // With m == 0 the reference's buffer is all zeros, so every read is 0.
    const size_t total = m > 0 ? (size_t)lda * n : 0;
    auto Rv = [&](size_t row, size_t c) -> double {
        size_t idx = row + c * lda;
        return idx < total ? M[idx] : 0.0;
    };
    if (tid == 0) {
        double tol = fabs(Rv(0, 0)) * eps;
        int k = 0;
        for (int i = 0; i < n; ++i)
            if (fabs(Rv(i, i)) > tol) pos[k++] = i;
        int q = 0;
        for (int i = 0; i < n; ++i)
            if (!(fabs(Rv(i, i)) > tol)) pos[k + q++] = i;
        s_k = k;
        kout[b] = k;
    }
    __syncthreads();
    if constexpr (LDSM)
        for (int e = tid; e < mtot; e += bs) Mg[e] = M[e];
    const int k = s_k, q = n - k;
    int* pb = perm + perm_off[b];
    for (int t = tid; t < n; t += bs) pb[t] = jp[pos[t]];
    if (jpvt_out)
        for (int t = tid; t < n; t += bs) jpvt_out[perm_off[b] + t] = jp[t];
    // W = R11^{-1} Wfull(:, red), Wfull(i, j) = R(skel[i], j) for j >= i else 0
    double* W = Wb + offW[b];
    for (int jj = tid; jj < q; jj += bs) {
        int jcol = pos[k + jj];
        double* y = W + (size_t)jj * k;
        for (int i = 0; i < k; ++i) y[i] = (jcol >= i) ? Rv(pos[i], jcol) : 0.0;
        for (int i = k - 1; i >= 0; --i) {
            if (y[i] != 0.0) {
                y[i] /= Rv(pos[i], i);
                double t = y[i];
                for (int r = 0; r < i; ++r) y[r] -= t * Rv(pos[r], i);
            }
        }
    }
}

template <class S, class D>
__global__ void k_convert(const S* __restrict__ src, D* __restrict__ dst, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = (D)src[i];
}

// dst[dst_off[b] + e] = src[src_off[b] + e] for e < dst_off[b+1] - dst_off[b]
template <class D>
__global__ void k_pack_boxes(int nbox, const long long* __restrict__ dst_off,
                             const long long* __restrict__ src_off,
                             const double* __restrict__ src, D* __restrict__ dst,
                             long long total) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; e < total; e += stride) {
        int b = find_box(dst_off, nbox, e);
        dst[e] = (D)src[src_off[b] + (e - dst_off[b])];
    }
}

// G(i, j) = sum_s P_s(i, j) for i >= j; grid (chunks of 256 columns' worth
// of elements, descriptors)
__global__ void __launch_bounds__(256) k_sum_slices_lower(const SliceSum* __restrict__ desc,
                                                          double* G) {
    const SliceSum d = desc[blockIdx.y];
    const long long nn = (long long)d.n * d.n;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < nn;
         e += (long long)gridDim.x * 256) {
        const int i = (int)(e % d.n), j = (int)(e / d.n);
        if (i < j) continue;
        const double* src = G + d.src + e;
        double acc = src[0];
        for (int q = 1; q < d.slices; ++q) acc += src[q * nn];
        G[d.dst + e] = acc;
    }
}

int grid_for(long long n, int block = 256, int cap = 1 << 16) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace

template <class T>
void batched_gj_inverse(int nbox, const int* h_n, const long long* h_off, const int* d_n,
                        const long long* d_off, T* d_A, int* d_info, GJWork& ws, GemmWork& gws,
                        hipStream_t s, bool sym) {
    if (nbox <= 0) return;
    const int panel = ws.panel > 0 ? ws.panel : kGJPanel;
    // Boxes of at most kSmallGJ rows: one workgroup each.  Larger boxes are
    // listed in slots 0..nbig-1 for the blocked path.
    std::vector<int> bix;
    bool any_wave = false, any_small = false;
    int max_n = 0;
    for (int b = 0; b < nbox; ++b) {
        const int n = h_n[b];
        if (n > kSmallGJ) {
            bix.push_back(b);
            max_n = std::max(max_n, n);
        } else if (n > kWaveGJ) {
            any_small = true;
        } else if (n > 0) {
            any_wave = true;
        }
    }
    if (any_wave) {
        k_gj_wave<T><<<(nbox + 3) / 4, 256, 0, s>>>(nbox, d_n, d_off, d_A, d_info);
        HIFDE_LAUNCH_CHECK();
    }
    if (any_small) {
        k_gj_small<T><<<nbox, 256, 0, s>>>(d_n, d_off, d_A, d_info);
        HIFDE_LAUNCH_CHECK();
    }
    const int nbig = (int)bix.size();
    if (nbig == 0) return;
    const int osteps = (max_n + panel - 1) / panel;
    // Per outer step and slot: window size, offset of A_KK, leading dimension.
    std::vector<int> wn((size_t)osteps * nbig, 0), wld((size_t)osteps * nbig, 1);
    std::vector<long long> woff((size_t)osteps * nbig, 0), r_off(nbig);
    std::vector<long long> ro_off(nbig + 1, 0), cw_off(nbig + 1, 0);
    for (int j = 0; j < nbig; ++j) {
        const int b = bix[j], n = h_n[b];
        r_off[j] = (long long)j * kGJB * panel;
        const bool outer = n > panel;
        ro_off[j + 1] = ro_off[j] + (outer ? (long long)panel * n : 0);
        cw_off[j + 1] = cw_off[j] + (outer ? (long long)n * panel : 0);
        for (int so = 0; so * panel < n; ++so) {
            const int K0 = so * panel;
            const size_t w = (size_t)so * nbig + j;
            wn[w] = std::min(panel, n - K0);
            woff[w] = h_off[b] + K0 + (long long)K0 * n;
            wld[w] = n;
        }
    }
    ws.wn.upload(wn, s);
    ws.wld.upload(wld, s);
    ws.woff.upload(woff, s);
    ws.b_off.upload(r_off, s);
    ws.bix.upload(bix, s);
    ws.info.reserve(nbig);
    k_fill_int<<<(nbig + 255) / 256, 256, 0, s>>>(ws.info.p, nbig, 0);
    ws.B.reserve((size_t)nbig * kGJB * panel * sizeof(T));
    // two D slots: step st uses slot st & 1 (a column step and the next
    // diagonal block run in one launch)
    ws.D.reserve(2 * (size_t)nbig * kGJB * kGJB * sizeof(T));
    // symmetric outer level: per box W (n x panel, ld n) and V^T
    // (panel x n); the general one: RO and CW
    std::vector<long long> sw_off(nbig + 1, 0);
    for (int j = 0; j < nbig; ++j) {
        const int n = h_n[bix[j]];
        sw_off[j + 1] = sw_off[j] + (sym && n > panel ? 2LL * n * panel : 0);
    }
    if (sym) {
        ws.SW.reserve((size_t)std::max<long long>(sw_off[nbig], 1) * sizeof(T));
    } else {
        ws.RO.reserve((size_t)std::max<long long>(ro_off[nbig], 1) * sizeof(T));
        ws.CW.reserve((size_t)std::max<long long>(cw_off[nbig], 1) * sizeof(T));
    }
    T* Rws = reinterpret_cast<T*>(ws.B.p);
    T* Dws = reinterpret_cast<T*>(ws.D.p);
    T* RO = reinterpret_cast<T*>(ws.RO.p);
    T* CW = reinterpret_cast<T*>(ws.CW.p);
    T* SW = reinterpret_cast<T*>(ws.SW.p);
    int* cinfo = ws.info.p;
    // The same walk runs twice: first it records every GEMM and strided copy
    // (uploaded in one copy each), then it launches them in order between
    // the window kernels.
    GemmPlan& plan = ws.plan;
    plan.clear();
    plan.esize = (int)sizeof(T);
    std::vector<std::pair<int, int>> gh;  // GEMM launch ranges in walk order
    std::vector<Copy2D> cps;
    std::vector<int> cfirst, cnum, cmaxc;  // per outer step: copy range and widest copy
    std::vector<GemmDesc> upd, g1, g2, g3;
    // symmetric-path copies: recorded as launch ranges in the first walk
    std::vector<SymCopy> scps;
    std::vector<long long> stoff;
    struct SRange {
        int d0, n;
        long long t0, tiles;
    };
    std::vector<SRange> srng;
    constexpr int kAll = 1 << 30;  // neg_from: no column negated
    for (int exec = 0; exec < 2; ++exec) {
        size_t gi = 0, si = 0;
        auto scopy = [&](const std::vector<SymCopy>& d) {
            if (!exec) {
                SRange r{(int)scps.size(), (int)d.size(), (long long)stoff.size(), 0};
                stoff.push_back(0);
                for (const SymCopy& c : d) {
                    r.tiles += (long long)((c.rows + 63) / 64) * ((c.cols + 63) / 64);
                    stoff.push_back(r.tiles);
                    scps.push_back(c);
                }
                srng.push_back(r);
                return;
            }
            const SRange& r = srng[si++];
            if (r.n == 0 || r.tiles == 0) return;
            k_sym_copy<T><<<(unsigned)r.tiles, 256, 0, s>>>(r.n, ws.stile.p + r.t0,
                                                            ws.sdesc.p + r.d0, d_A, SW);
            HIFDE_LAUNCH_CHECK();
        };
        auto gemm = [&](const std::vector<GemmDesc>& d, const T* A, const T* B, T* C, T alpha,
                        T beta, bool transA = false, bool one_launch = false) {
            if (!exec) gh.push_back(plan.add(d, transA, one_launch));
            else gemm_run<T, T>(plan, gh[gi++], A, B, C, alpha, beta, s);
        };
        for (int so = 0; so < osteps; ++so) {
            const int K0 = so * panel;
            const size_t w0 = (size_t)so * nbig;
            const int* dwn = ws.wn.p + w0;
            const int* dwld = ws.wld.p + w0;
            const long long* dwoff = ws.woff.p + w0;
            int max_w = 0;
            for (int j = 0; j < nbig; ++j) max_w = std::max(max_w, wn[w0 + j]);
            if (sym) {
                // only lower triangles are current: the window's upper
                // triangle is rebuilt from its lower one before inversion
                // (a box of one window is read whole, as given)
                std::vector<SymCopy> mir;
                for (int j = 0; j < nbig && !exec; ++j) {
                    const int kb = wn[w0 + j];
                    if (kb == 0 || h_n[bix[j]] <= panel) continue;
                    const long long wo = woff[w0 + j];
                    const int ld = wld[w0 + j];
                    mir.push_back({wo, wo, kb, kb, ld, ld, 1, 1, kAll, 1, 0, 0});
                }
                scopy(mir);
            }
            // inner level: every window inverted in place
            const int isteps = (max_w + kGJB - 1) / kGJB;
            const dim3 gch((max_w + 63) / 64, nbig);
            const int nchunk = (int)gch.x;
            const size_t dslot = (size_t)nbig * kGJB * kGJB;
            auto Dst = [&](int st) { return Dws + (st & 1) * dslot; };
            for (int st = 0; st < isteps; ++st) {
                if (exec) {
                    if (!ws.fused_steps)
                        k_bgj_diag<T><<<nbig, kDiagWaves * 64, 0, s>>>(dwn, dwoff, dwld, d_A,
                                                                      Dst(st), st, K0, cinfo);
                    else if (st == 0)
                        k_bgj_col_diag<T><<<nbig, 256, 0, s>>>(nchunk, nbig, dwn, dwoff, dwld,
                                                               d_A, nullptr, Dst(0), Rws,
                                                               ws.b_off.p, -1, 0, K0, cinfo);
                    k_bgj_row<T><<<gch, 256, 0, s>>>(dwn, dwoff, dwld, d_A, Dst(st), Rws,
                                                      ws.b_off.p, st, cinfo);
                }
                // A(:, J) -= A(:, K) R(:, J) over the window's columns left and
                // right of the block; the block rows are then set to R.
                upd.clear();
                for (int j = 0; j < nbig && !exec; ++j) {
                    const size_t w = w0 + j;
                    const int nw = wn[w], k0 = st * kGJB;
                    if (k0 >= nw) continue;
                    const int kb = std::min(kGJB, nw - k0), ld = wld[w];
                    const long long o = woff[w], ro = r_off[j];
                    if (k0 > 0) upd.push_back({o + (long long)k0 * ld, ro, o, nw, k0, kb, ld, kGJB, ld});
                    const int j1 = k0 + kb;
                    if (j1 < nw)
                        upd.push_back({o + (long long)k0 * ld, ro + (long long)j1 * kGJB,
                                       o + (long long)j1 * ld, nw, nw - j1, kb, ld, kGJB, ld});
                }
                gemm(upd, d_A, Rws, d_A, (T)-1, (T)1, false, true);
                if (exec) {
                    if (!ws.fused_steps) {
                        k_bgj_col<T><<<gch, 256, 0, s>>>(dwn, dwoff, dwld, d_A, Dst(st), Rws,
                                                          ws.b_off.p, st, cinfo);
                    } else {
                        const bool next = st + 1 < isteps;
                        k_bgj_col_diag<T><<<nchunk * nbig + (next ? nbig : 0), 256, 0, s>>>(
                            nchunk, nbig, dwn, dwoff, dwld, d_A, Dst(st), Dst(st + 1), Rws,
                            ws.b_off.p, st, next ? st + 1 : -1, K0, cinfo);
                    }
                }
            }
            if (sym) {
                // Symmetric outer level.  With P the windows already
                // processed, the in-place iterate Z keeps Z_PP and Z_QQ
                // symmetric and Z_QP = -Z_PQ^T, so its lower triangle holds
                // everything.  Step K, with V = Z(:, K) over the rows != K
                // and W = V D (D = Z_KK after the inner level), is
                //   Z_IJ -= s_J W_I V_J^T    (I >= J; s_J = -1 in P, +1 after K)
                //   Z(I, K) = -W_I,  Z(K, J) = -W_J^T  (J in P)
                // -- the general level's rank-kb update on the lower triangle
                // only.  Both halves of V come from the lower triangle as
                // Vt = (sigma V)^T with sigma = +1 in P, -1 after K:
                //   Vt(:, J in P)  = -Z(K, J)       (a row block of Z)
                //   Vt(:, J > K)   = -Z(J, K)^T     (a column block, transposed)
                // and Wt = Vt^T D = sigma W, so the update is +Wt_I Vt_J for
                // I in P and -Wt_I Vt_J below K.
                // D, the inverted window, is symmetric in exact arithmetic and
                // the update below relies on D^T = D: it is made exactly
                // symmetric, (D + D^T) / 2, which is also nearer the true
                // inverse, before use.
                std::vector<SymCopy> sy;
                for (int j = 0; j < nbig && !exec; ++j) {
                    const int kb = wn[w0 + j];
                    if (kb == 0 || h_n[bix[j]] <= panel) continue;
                    const long long wo = woff[w0 + j];
                    const int ld = wld[w0 + j];
                    sy.push_back({wo, wo, kb, kb, ld, ld, 1, 1, kAll, 1, 0, 0, 2});
                }
                std::vector<SymCopy> gv, wb;
                std::vector<GemmDesc> gw, tl, bq;
                for (int j = 0; j < nbig && !exec; ++j) {
                    const int n = h_n[bix[j]];
                    const int kb = wn[w0 + j];
                    if (kb == 0 || n <= panel) continue;
                    const long long o = h_off[bix[j]];
                    const long long swo = sw_off[j], svt = swo + (long long)n * panel;
                    const int K1 = K0 + kb, nq = n - K1, ldv = n;
                    if (K0 > 0) gv.push_back({svt, o + K0, kb, K0, panel, n, 0, 0, 0, 1, 1, 0});
                    if (nq > 0)
                        gv.push_back({svt + (long long)K0 * panel, o + K1 + (long long)K0 * n,
                                      kb, nq, panel, n, 1, 0, 0, 1, 1, 0});
                    gw.push_back({svt, o + K0 + (long long)K0 * n, swo, n - kb, kb, kb, panel,
                                  n, ldv});
                    if (K0 > 0) {
                        GemmDesc d{swo, svt, o, K0, K0, kb, ldv, panel, n};
                        d.lower = 1;
                        tl.push_back(d);
                    }
                    if (K0 > 0 && nq > 0)
                        bq.push_back({swo + K0, svt, o + K1, nq, K0, kb, ldv, panel, n});
                    if (nq > 0) {
                        GemmDesc d{swo + K0, svt + (long long)K0 * panel,
                                   o + K1 + (long long)K1 * n, nq, nq, kb, ldv, panel, n};
                        d.lower = 1;
                        bq.push_back(d);
                        wb.push_back({o + K1 + (long long)K0 * n, swo + K0, nq, kb, n, ldv, 0, 0,
                                      kAll, 1, 0, 1});
                    }
                    if (K0 > 0) wb.push_back({o + K0, swo, kb, K0, n, ldv, 1, 0, 0, 1, 0, 1});
                }
                // one launch: the symmetrization touches only the window's
                // diagonal block, the gathers read the blocks beside it and
                // write the V^T workspace
                sy.insert(sy.end(), gv.begin(), gv.end());
                scopy(sy);
                gemm(gw, SW, d_A, SW, (T)1, (T)0, true);
                gemm(tl, SW, SW, d_A, (T)1, (T)1);
                gemm(bq, SW, SW, d_A, (T)-1, (T)1);
                scopy(wb);
                continue;
            }
            // outer level, boxes wider than one window
            g1.clear();
            g2.clear();
            g3.clear();
            if (!exec) {
                cfirst.push_back((int)cps.size());
                int maxc = 0;
                for (int j = 0; j < nbig; ++j) {
                    const int n = h_n[bix[j]];
                    const int kb = wn[w0 + j];
                    if (kb == 0 || n <= panel) continue;
                    const long long o = h_off[bix[j]], dk = woff[w0 + j];
                    const long long ro = ro_off[j], co = cw_off[j];
                    const int j1 = K0 + kb;
                    const int lo[2] = {0, j1}, len[2] = {K0, n - j1};
                    for (int q = 0; q < 2; ++q) {
                        if (len[q] == 0) continue;
                        // R(:, J) = D A(K, J)
                        g1.push_back({dk, o + K0 + (long long)lo[q] * n,
                                      ro + (long long)lo[q] * panel, kb, len[q], kb, n, n,
                                      panel});
                        // C(I, :) = -A(I, K) D
                        g3.push_back({o + lo[q] + (long long)K0 * n, dk, co + lo[q], len[q], kb,
                                      kb, n, n, n});
                        // A(I, K) = C(I, :);  A(K, J) = R(:, J)
                        cps.push_back({co + lo[q], o + lo[q] + (long long)K0 * n, len[q], kb, n,
                                       n, 0});
                        cps.push_back({ro + (long long)lo[q] * panel,
                                       o + K0 + (long long)lo[q] * n, kb, len[q], panel, n, 1});
                        maxc = std::max(maxc, std::max(kb, len[q]));
                        // A(I, J) -= A(I, K) R(:, J)
                        for (int r = 0; r < 2; ++r) {
                            if (len[r] == 0) continue;
                            g2.push_back({o + lo[r] + (long long)K0 * n,
                                          ro + (long long)lo[q] * panel,
                                          o + lo[r] + (long long)lo[q] * n, len[r], len[q], kb, n,
                                          panel, n});
                        }
                    }
                }
                cnum.push_back((int)cps.size() - cfirst.back());
                cmaxc.push_back(maxc);
            }
            gemm(g1, d_A, d_A, RO, (T)1, (T)0);
            gemm(g2, d_A, RO, d_A, (T)-1, (T)1);
            gemm(g3, d_A, d_A, CW, (T)-1, (T)0);
            if (exec && cnum[so] > 0) {
                k_copy2d<T><<<dim3((unsigned)cnum[so], (cmaxc[so] + 3) / 4), 256, 0, s>>>(
                    ws.cdesc.p + cfirst[so], CW, RO, d_A);
                HIFDE_LAUNCH_CHECK();
            }
        }
        if (sym) {
            // every window processed: Z = A^{-1}, symmetric
            std::vector<SymCopy> mir;
            for (int j = 0; j < nbig && !exec; ++j) {
                const int n = h_n[bix[j]];
                if (n <= panel) continue;
                const long long o = h_off[bix[j]];
                mir.push_back({o, o, n, n, n, n, 1, 1, kAll, 1, 0, 0});
            }
            scopy(mir);
        }
        if (!exec) {
            plan.upload(s);
            ws.cdesc.upload(cps, s);
            if (sym) {
                ws.sdesc.upload(scps, s);
                ws.stile.upload(stoff, s);
            }
        }
    }
    k_scatter_info<<<(nbig + 255) / 256, 256, 0, s>>>(nbig, ws.bix.p, cinfo, d_info);
    HIFDE_LAUNCH_CHECK();
}

void fold_info(int nbox, const int* d_info, int* d_status, hipStream_t s) {
    if (nbox <= 0) return;
    if (!d_info || !d_status) throw Error(-4, "fold_info: buffer not allocated");
    k_fold_info<<<1, 256, 0, s>>>(nbox, d_info, d_status);
    HIFDE_LAUNCH_CHECK();
}

void fold_active(int n, const int* d_active, int* d_count, hipStream_t s) {
    if (!d_active || !d_count) throw Error(-4, "fold_active: buffer not allocated");
    k_count_nonzero<<<1, 256, 0, s>>>(n, d_active, d_count);
    HIFDE_LAUNCH_CHECK();
}

void fill_int(int* p, int n, int v, hipStream_t s) {
    if (n <= 0) return;
    k_fill_int<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
    HIFDE_LAUNCH_CHECK();
}

template <class T>
constexpr const char* type_name() {
    return sizeof(T) == sizeof(float) ? "float" : "double";
}

// Output computed on 128 x 128 tiles over that on 64 x 64 tiles above which a
// product of depth above 64 takes the 64 x 64 tiles (GemmPlan::tile64 >= 1).
constexpr double kTile64Waste = 1.25;

namespace {
std::atomic<const signed char*> g_tune_table{nullptr};
thread_local GemmRecorder* t_recorder = nullptr;
int bin_of(long long v, const long long* edges, int nedge) {
    int b = 0;
    while (b < nedge && v > edges[b]) ++b;
    return b;
}
}  // namespace
GemmRecorder* active_gemm_recorder() { return t_recorder; }
void set_active_gemm_recorder(GemmRecorder* r) { t_recorder = r; }

int gemm_tune_key(int esize, int transA, int lower, int m, int n, int k, int nb) {
    static const long long kE[kTuneKBins - 1] = {16, 32, 64, 128, 256, 512, 1024};
    static const long long mnE[kTuneMNBins - 1] = {31, 63, 127, 255, 511};
    static const long long nbE[kTuneNBBins - 1] = {1, 7, 63, 511};
    const int kb = bin_of(k, kE, kTuneKBins - 1);
    const int mb = bin_of(std::min(m, n), mnE, kTuneMNBins - 1);
    const int bb = bin_of(nb, nbE, kTuneNBBins - 1);
    return ((((((esize == 4 ? 0 : 1) * 2 + (transA ? 1 : 0)) * 2 + (lower ? 1 : 0)) * kTuneKBins +
              kb) * kTuneMNBins + mb) * kTuneNBBins) + bb;
}
int gemm_tune_lookup(int key) {
    const signed char* t = g_tune_table.load(std::memory_order_acquire);
    if (!t || key < 0 || key >= kTuneEntries) return -1;
    return t[key];
}
// Installs a routing table (kTuneEntries kinds); the table lives until the
// process ends (a replaced one is kept too: plans on other threads may be
// reading it).
void gemm_tune_install(const signed char* kinds) {
    signed char* t = new signed char[kTuneEntries];
    std::copy(kinds, kinds + kTuneEntries, t);
    g_tune_table.store(t, std::memory_order_release);
}

std::pair<int, int> GemmPlan::add(const std::vector<GemmDesc>& d, bool transA, bool one_launch) {
    constexpr int TILE = kMfmaTile;
    const int first = (int)launches.size();
    // kinds: 0 k_gemm (FMA, 64 x 64 tiles), 1 k_gemm_mfma; -1 empty.
    // Classified once, then one launch per kind in descriptor order.
    std::vector<signed char> kinds(d.size());
    int count[3] = {0, 0, 0};
    // output tiles a launch computes (lower: tiles with column <= row)
    auto tiles_of = [](const GemmDesc& g, int tile) {
        const long long tm = (g.m + tile - 1) / tile, tn = (g.n + tile - 1) / tile;
        if (!g.lower) return tm * tn;
        long long c = 0;
        for (long long j = 0; j < tn; ++j) c += std::max(0LL, tm - j);
        return c;
    };
    for (size_t i = 0; i < d.size(); ++i) {
        const GemmDesc& g = d[i];
        if (g.m <= 0 || g.n <= 0) {
            kinds[i] = -1;
            continue;
        }
        // Products of depth 64 or less (the block inverse's in-window
        // updates, the Gram panels' trailing updates) are bound by the output
        // tile's read and write, or by latency where a level has few boxes:
        // the FMA kernel's 64 x 64 tiles, with the next slab prefetched,
        // finish them sooner than the 128 x 128 matrix-core tiles.
        const bool small = g.k <= 64;
        // the MFMA kernel addresses each operand with 32-bit byte offsets:
        // keep operand extents (8-byte elements, plus a slab and a tile of
        // overrun) under 2 GiB
        const double ext_a = (double)g.lda * (transA ? g.m : g.k) + 16.0 * g.lda + 256;
        const double ext_b = (double)g.ldb * g.n + 16.0 * g.ldb + 256;
        const bool fits = 8.0 * std::max(ext_a, ext_b) < 2147483648.0;
        const bool large = g.m >= TILE / 2 && g.n >= TILE / 2 && !small && fits;
        int kind = large ? 1 : 0;
        // 64 x 64 matrix-core tiles: where the 128 x 128 tiles would compute
        // over a quarter more output (tile1), and also for the products of
        // depth 64 or less with both sides of at least 32 (tile2)
        if (tile64 >= 1 && large &&
            tiles_of(g, TILE) * TILE * TILE >
                kTile64Waste * tiles_of(g, TILE / 2) * (TILE / 2) * (TILE / 2))
            kind = 2;
        if (tile64 >= 2 && small && g.k > 0 && g.m >= 32 && g.n >= 32 && fits) kind = 2;
        // 3: every product with both output sides of at least 32
        if (tile64 >= 3 && g.k > 0 && g.m >= 32 && g.n >= 32 && fits) kind = 2;
        // kinds the matrix-core kernel has been validated for: both output
        // sides of at least 32, a positive depth, operands within its
        // 32-bit offsets
        const bool core_ok = fits && g.k > 0 && g.m >= 32 && g.n >= 32;
        int want = force_kind;
        if (want < 0 && use_table)
            want = gemm_tune_lookup(
                gemm_tune_key(esize, transA, g.lower, g.m, g.n, g.k, (int)d.size()));
        if (want == 0 || (want > 0 && core_ok)) kind = want;
        kinds[i] = (signed char)kind;
        ++count[kind];
        if (GemmRecorder* rec = t_recorder)
            rec->log.insert(rec->log.end(), {rec->calls, esize, transA ? 1 : 0, g.lower, g.m,
                                             g.n, g.k, (int)d.size(), kind});
    }
    if (t_recorder) ++t_recorder->calls;
    // A latency-bound call of few descriptors (the block inverse's in-window
    // update of a few large boxes has a left and a right part of different
    // shapes): when its descriptors were routed to several kernels, they all
    // go to the one that carries most of its flops (where that kernel takes
    // every one of them), so the call is one launch.
    if (one_launch && (int)d.size() <= kUnifyMaxDesc && force_kind < 0 &&
        (count[0] > 0) + (count[1] > 0) + (count[2] > 0) > 1) {
        double fl[3] = {0, 0, 0};
        for (size_t i = 0; i < d.size(); ++i)
            if (kinds[i] >= 0) fl[kinds[i]] += (double)d[i].m * d[i].n * d[i].k;
        const int best = fl[0] >= fl[1] ? (fl[0] >= fl[2] ? 0 : 2) : (fl[1] >= fl[2] ? 1 : 2);
        bool ok = true;
        for (size_t i = 0; i < d.size() && best > 0; ++i) {
            const GemmDesc& g = d[i];
            if (kinds[i] < 0) continue;
            const double ext_a = (double)g.lda * (transA ? g.m : g.k) + 16.0 * g.lda + 256;
            const double ext_b = (double)g.ldb * g.n + 16.0 * g.ldb + 256;
            const bool fits = 8.0 * std::max(ext_a, ext_b) < 2147483648.0;
            ok = ok && fits && g.k > 0 && g.m >= 32 && g.n >= 32;
        }
        if (ok)
            for (size_t i = 0; i < d.size(); ++i)
                if (kinds[i] >= 0 && kinds[i] != best) {
                    --count[kinds[i]];
                    kinds[i] = (signed char)best;
                    ++count[best];
                }
    }
    desc.reserve(desc.size() + d.size());
    tiles.reserve(tiles.size() + d.size() + 3);
    for (int pass = 0; pass < 3; ++pass) {
        if (count[pass] == 0) continue;
        const int d0 = (int)desc.size(), t0 = (int)tiles.size();
        const int tile = pass == 1 ? TILE : pass == 2 ? TILE / 2 : BM;
        Launch L{pass, transA ? 1 : 0, 0, d0, t0, 0, 0.0, 0.0, 0.0, pass >= 1 && split ? 1 : 0};
        tiles.push_back(0);
        for (size_t i = 0; i < d.size(); ++i) {
            if (kinds[i] != pass) continue;
            const GemmDesc& g = d[i];
            const int tm = (g.m + tile - 1) / tile, tn = (g.n + tile - 1) / tile;
            desc.push_back(g);
            // both kernels list only a lower descriptor's tiles on or below
            // the diagonal
            tiles.push_back(tiles.back() + tiles_of(g, tile));
            ++L.nb;
            const double cells = g.lower ? 0.5 * g.m * (g.n + 1.0) : (double)g.m * g.n;
            L.flops += 2.0 * cells * g.k;
            L.cells += cells;
            L.ab_elems += (double)g.m * g.k + (double)g.k * g.n;
        }
        L.total = tiles.back();
        L.contig = pass >= 1 ? 1 : 0;
        launches.push_back(L);
    }
    return {first, (int)launches.size()};
}

template <class T, class TC>
void gemm_run(const GemmPlan& p, std::pair<int, int> h, const T* d_A, const T* d_B, TC* d_C,
              T alpha, T beta, hipStream_t s, const int* d_flags) {
    KernelTimer* kt = active_timer();
    for (int li = h.first; li < h.second; ++li) {
        const GemmPlan::Launch& L = p.launches[li];
        const GemmDesc* dd = p.d_desc.p + L.d0;
        const int* tt = p.d_tiles.p + L.t0;
        char name[96];
        const char* tn = type_name<T>();
        const char* cn = type_name<TC>();
        const char* ta = L.transA ? "true" : "false";
        if (kt) {
            // the kernel's name as the profilers print it
            if (L.mfma)
                std::snprintf(name, sizeof name, "k_gemm_mfma<%s, %s, %s, %s, %d>", tn, ta, cn,
                              L.split ? "true" : "false", L.mfma == 2 ? 64 : kMfmaTile);
            else
                std::snprintf(name, sizeof name, "k_gemm<%s, %s, %s>", tn, ta, cn);
            kt->begin(s, name);
        }
        if (!L.mfma) {
            if (L.transA)
                k_gemm<T, true, TC><<<L.total, 256, 0, s>>>(L.nb, tt, dd, d_A, d_B, d_C, alpha,
                                                            beta, d_flags);
            else
                k_gemm<T, false, TC><<<L.total, 256, 0, s>>>(L.nb, tt, dd, d_A, d_B, d_C, alpha,
                                                             beta, d_flags);
        } else if (L.mfma == 2) {
            if (L.split) {
                if constexpr (std::is_same<T, float>::value) {
                    if (L.transA)
                        k_gemm_mfma<T, true, TC, true, 64><<<L.total, 256, 0, s>>>(
                            L.nb, tt, dd, d_A, d_B, d_C, alpha, beta, d_flags, L.contig);
                    else
                        k_gemm_mfma<T, false, TC, true, 64><<<L.total, 256, 0, s>>>(
                            L.nb, tt, dd, d_A, d_B, d_C, alpha, beta, d_flags, L.contig);
                } else {
                    throw Error(-1, "split (bf16x3) products need FP32 operands");
                }
            } else if (L.transA) {
                k_gemm_mfma<T, true, TC, false, 64><<<L.total, 256, 0, s>>>(
                    L.nb, tt, dd, d_A, d_B, d_C, alpha, beta, d_flags, L.contig);
            } else {
                k_gemm_mfma<T, false, TC, false, 64><<<L.total, 256, 0, s>>>(
                    L.nb, tt, dd, d_A, d_B, d_C, alpha, beta, d_flags, L.contig);
            }
        } else if (L.split) {
            if constexpr (std::is_same<T, float>::value) {
                if (L.transA)
                    k_gemm_mfma<T, true, TC, true><<<L.total, 256, 0, s>>>(
                        L.nb, tt, dd, d_A, d_B, d_C, alpha, beta, d_flags, L.contig);
                else
                    k_gemm_mfma<T, false, TC, true><<<L.total, 256, 0, s>>>(
                        L.nb, tt, dd, d_A, d_B, d_C, alpha, beta, d_flags, L.contig);
            } else {
                throw Error(-1, "split (bf16x3) products need FP32 operands");
            }
        } else {
            if (L.transA)
                k_gemm_mfma<T, true, TC><<<L.total, 256, 0, s>>>(L.nb, tt, dd, d_A, d_B, d_C,
                                                                 alpha, beta, d_flags, L.contig);
            else
                k_gemm_mfma<T, false, TC><<<L.total, 256, 0, s>>>(L.nb, tt, dd, d_A, d_B, d_C,
                                                                  alpha, beta, d_flags, L.contig);
        }
        HIFDE_LAUNCH_CHECK();
        if (kt)
            kt->end(s, L.flops,
                    L.ab_elems * sizeof(T) + L.cells * sizeof(TC) * (beta == (T)0 ? 1 : 2));
    }
}

template <class T, class TC>
void batched_gemm_desc(bool transA, const std::vector<GemmDesc>& desc, const T* d_A, const T* d_B,
                       TC* d_C, T alpha, T beta, GemmWork& ws, hipStream_t s,
                       const int* d_flags) {
    if (desc.empty()) return;
    ws.plan.clear();
    ws.plan.esize = (int)sizeof(T);
    const std::pair<int, int> h = ws.plan.add(desc, transA);
    if (h.first == h.second) return;
    ws.plan.upload(s);
    gemm_run<T, TC>(ws.plan, h, d_A, d_B, d_C, alpha, beta, s, d_flags);
}



template <class T>
void batched_gemm(bool transA, int nbox, const int* h_m, const int* h_n, const int* h_k,
                  const long long* h_offA, const long long* h_offB, const long long* h_offC,
                  const T* d_A, const T* d_B, T* d_C, T alpha, GemmWork& ws, hipStream_t s,
                  bool lower) {
    std::vector<GemmDesc> desc(nbox);
    for (int b = 0; b < nbox; ++b) {
        GemmDesc& d = desc[b];
        d.offA = h_offA[b];
        d.offB = h_offB[b];
        d.offC = h_offC[b];
        d.m = h_m[b];
        d.n = h_n[b];
        d.k = h_k[b];
        d.lda = std::max(transA ? d.k : d.m, 1);
        d.ldb = std::max(d.k, 1);
        d.ldc = std::max(d.m, 1);
        d.lower = lower && d.m == d.n ? 1 : 0;
    }
    batched_gemm_desc<T>(transA, desc, d_A, d_B, d_C, alpha, (T)0, ws, s);
}

template <class D>
void copy_blocks(const std::vector<CopyDesc>& desc, const std::vector<long long>& dst_off,
                 const double* d_src, D* d_dst, GemmWork& ws, hipStream_t s) {
    const int nbox = (int)desc.size();
    if (nbox <= 0 || dst_off.back() <= 0) return;
    ws.cdesc.upload(desc, s);
    ws.off.upload(dst_off, s);
    k_copy_blocks<D><<<grid_for(dst_off.back()), 256, 0, s>>>(nbox, ws.off.p, ws.cdesc.p, d_src,
                                                              d_dst, dst_off.back());
    HIFDE_LAUNCH_CHECK();
}

void permute_sym(int nbox, const long long* d_off, const int* d_n, const int* d_perm,
                 const int* d_perm_off, const double* d_C, double* d_Cp, long long total,
                 hipStream_t s) {
    if (nbox <= 0 || total <= 0) return;
    k_permute_sym<<<grid_for(total), 256, 0, s>>>(nbox, d_off, d_n, d_perm, d_perm_off, d_C, d_Cp,
                                                  total);
    HIFDE_LAUNCH_CHECK();
}

void gram_pivchol(int nface, const int* h_n, const unsigned char* h_mode, const int* d_n,
                  const unsigned char* d_mode, double* d_G, const long long* h_offG,
                  const long long* d_offG, double* d_R, double eps, int* d_perm,
                  const int* h_perm_off, const int* d_perm_off, int* d_k, GramWork& ws,
                  GemmWork& gws, hipStream_t s) {
    int max_n = 0;
    std::vector<long long> offP(nface + 1, 0);
    std::vector<int> act(nface, 0);
    std::vector<GemmDesc> upd;
    for (int f = 0; f < nface; ++f) {
        const bool g = h_mode[f] && h_n[f] > 0;
        offP[f + 1] = offP[f] + (g ? (long long)h_n[f] * kGramNB : 0);
        if (!g) continue;
        act[f] = 1;
        max_n = std::max(max_n, h_n[f]);
        // trailing update G -= P^T P over the lower triangle
        GemmDesc u{offP[f], offP[f], h_offG[f], h_n[f], h_n[f], kGramNB, kGramNB, kGramNB,
                   h_n[f]};
        u.lower = 1;
        u.flag = f;
        upd.push_back(u);
    }
    if (max_n == 0) return;
    const size_t shm = (size_t)max_n * (sizeof(double) + sizeof(int));
    if (shm > 150 * 1024) throw Error(-3, "gram_pivchol: face too large for LDS");
    ws.P.reserve(std::max<long long>(offP.back(), 1));
    ws.offP.upload(offP, s);
    ws.dg.reserve(std::max(h_perm_off[nface], 1));
    ws.tol.reserve(nface);
    ws.active.upload(act, s);
    ws.count.reserve(1);
    ws.host_active.resize(1);
    // the trailing update has the same descriptors after every panel
    ws.plan.clear();
    ws.plan.esize = 8;
    const std::pair<int, int> uh = ws.plan.add(upd, true);
    ws.plan.upload(s);
    const size_t shm_reg = (size_t)max_n * 2 * sizeof(int);
    for (int j0 = 0; j0 < max_n; j0 += kGramNB) {
#define HIFDE_GRAM_REG(BS, CPT)                                                                   \
    k_gram_panel_reg<BS, CPT><<<nface, BS, shm_reg, s>>>(                                         \
        d_n, d_mode, d_G, d_offG, d_R, ws.P.p, ws.offP.p, ws.dg.p, d_perm, d_perm_off, d_k,       \
        ws.active.p, ws.tol.p, eps, j0)
        if (!ws.reg_panel || max_n > 1536)
            k_gram_panel<<<nface, 256, shm, s>>>(d_n, d_mode, d_G, d_offG, d_R, ws.P.p, ws.offP.p,
                                                 ws.dg.p, d_perm, d_perm_off, d_k, ws.active.p,
                                                 ws.tol.p, eps, j0);
        else if (max_n <= 64) HIFDE_GRAM_REG(64, 1);
        else if (max_n <= 256) HIFDE_GRAM_REG(256, 1);
        else if (max_n <= 512) HIFDE_GRAM_REG(512, 1);
        else if (max_n <= 1024) HIFDE_GRAM_REG(512, 2);
        else HIFDE_GRAM_REG(512, 3);
#undef HIFDE_GRAM_REG
        HIFDE_LAUNCH_CHECK();
        if (j0 + kGramNB >= max_n) break;
        gemm_run<double, double>(ws.plan, uh, ws.P.p, ws.P.p, d_G, -1.0, 1.0, s, ws.active.p);
        // every fourth panel, stop once every face has found its rank
        if ((j0 / kGramNB) % 4 == 3) {
            fold_active(nface, ws.active.p, ws.count.p, s);
            HIFDE_CHECK(hipMemcpyAsync(ws.host_active.data(), ws.count.p, sizeof(int),
                                       hipMemcpyDeviceToHost, s));
            HIFDE_CHECK(hipStreamSynchronize(s));
            if (ws.host_active[0] == 0) break;
        }
    }
}

void sum_slices_lower(const std::vector<SliceSum>& desc, DevBuf<SliceSum>& d_desc, double* d_G,
                      hipStream_t s) {
    if (desc.empty()) return;
    long long max_nn = 0;
    for (const SliceSum& d : desc) max_nn = std::max(max_nn, (long long)d.n * d.n);
    d_desc.upload(desc, s);
    const dim3 grid((unsigned)std::min<long long>((max_nn + 255) / 256, 1024), (unsigned)desc.size());
    k_sum_slices_lower<<<grid, 256, 0, s>>>(d_desc.p, d_G);
    HIFDE_LAUNCH_CHECK();
}

void gram_w(const std::vector<GramRDesc>& desc, const std::vector<long long>& toff,
            const double* d_R, const int* d_perm, const std::vector<long long>& offW,
            double* d_W, GramWork& ws, GJWork& gj, GemmWork& gws, hipStream_t s) {
    const int nb = (int)desc.size();
    if (nb == 0 || toff.back() == 0) return;
    ws.rdesc.upload(desc, s);
    ws.roff.upload(toff, s);
    ws.RK.reserve(toff.back());
    k_gram_gather_r<<<grid_for(toff.back()), 256, 0, s>>>(nb, ws.roff.p, ws.rdesc.p, d_R, d_perm,
                                                         ws.RK.p, toff.back());
    HIFDE_LAUNCH_CHECK();
    // R11^{-1} in place (the leading k x k block of each [R11 | R12])
    std::vector<int> hk(nb);
    std::vector<long long> ho(nb);
    for (int t = 0; t < nb; ++t) {
        hk[t] = desc[t].k;
        ho[t] = toff[t];
    }
    ws.kd.upload(hk, s);
    ws.ko.upload(ho, s);
    ws.info.reserve(nb);
    fill_int(ws.info.p, nb, 0, s);
    batched_gj_inverse<double>(nb, hk.data(), ho.data(), ws.kd.p, ws.ko.p, ws.RK.p, ws.info.p, gj,
                               gws, s);
    // W = R11^{-1} R12 (k x q, ld k)
    std::vector<GemmDesc> g;
    for (int t = 0; t < nb; ++t) {
        const int k = desc[t].k, q = desc[t].n - desc[t].k;
        if (k == 0 || q == 0) continue;
        g.push_back({toff[t], toff[t] + (long long)k * k, offW[t], k, q, k, k, k, k});
    }
    batched_gemm_desc<double>(false, g, ws.RK.p, ws.RK.p, d_W, 1.0, 0.0, gws, s);
}

void batched_interp_decomp(int nface, const int* d_m, const int* d_n, const long long* d_offM,
                           double* d_M, double eps, int* d_perm, const int* d_perm_off, int* d_k,
                           double* d_W, const long long* d_offW, int max_n, int max_m,
                           const unsigned char* d_mode, hipStream_t s, int* d_jpvt,
                           double* d_tau) {
    if (nface <= 0) return;
    size_t shm = (size_t)max_n * (2 * sizeof(double) + 2 * sizeof(int));
    if (shm > 150 * 1024) throw Error(-3, "interp_decomp: face too large for LDS");
    int threads = (max_m <= 128 && max_n <= 16) ? 64 : 256;
    // Small faces (the level-0 faces) run in LDS: their steps are chains of
    // dependent reductions, which global-memory latency would dominate.
    const size_t shm_m = (size_t)std::max(max_m, 1) * max_n * sizeof(double);
    if (shm + shm_m <= 32 * 1024)
        k_interp_decomp<true><<<nface, threads, shm + shm_m, s>>>(
            d_m, d_n, d_offM, d_M, eps, d_perm, d_perm_off, d_k, d_W, d_offW, d_mode, d_jpvt,
            d_tau);
    else
        k_interp_decomp<false><<<nface, threads, shm, s>>>(d_m, d_n, d_offM, d_M, eps, d_perm,
                                                          d_perm_off, d_k, d_W, d_offW, d_mode,
                                                          d_jpvt, d_tau);
    HIFDE_LAUNCH_CHECK();
}

// B := R^{-1} B for upper-triangular R, one box per workgroup, one column of
// B per thread (TriangularSolveWrapper, src/InterpDecomp.cpp:213-231: dtrsm
// side L, upper, no transpose, non-unit diagonal).  Back substitution in the
// reference BLAS order: an entry that is exactly zero is skipped, a zero
// diagonal divides by zero as dtrsm does.
__global__ void __launch_bounds__(256) k_trsm_upper(const int* __restrict__ dm,
                                                    const int* __restrict__ dn,
                                                    const long long* __restrict__ offR,
                                                    const double* __restrict__ Rb,
                                                    const long long* __restrict__ offB,
                                                    double* Bb) {
    const int b = blockIdx.y;
    const int m = dm[b], n = dn[b];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || m <= 0) return;
    const double* R = Rb + offR[b];
    double* y = Bb + offB[b] + (size_t)j * m;
    for (int i = m - 1; i >= 0; --i) {
        if (y[i] != 0.0) {
            y[i] /= R[i + (size_t)i * m];
            const double t = y[i];
            for (int r = 0; r < i; ++r) y[r] -= t * R[r + (size_t)i * m];
        }
    }
}

void batched_trsm_upper(int nbox, const int* d_m, const int* d_n, const long long* d_offR,
                        const double* d_R, const long long* d_offB, double* d_B, int max_n,
                        hipStream_t s) {
    if (nbox <= 0 || max_n <= 0) return;
    dim3 grid((max_n + 255) / 256, nbox);
    k_trsm_upper<<<grid, 256, 0, s>>>(d_m, d_n, d_offR, d_R, d_offB, d_B);
    HIFDE_LAUNCH_CHECK();
}

template <class D>
void pack_boxes(int nbox, const long long* d_dst_off, const long long* d_src_off,
                const double* d_src, D* d_dst, long long total, hipStream_t s) {
    if (nbox <= 0 || total <= 0) return;
    k_pack_boxes<D><<<grid_for(total), 256, 0, s>>>(nbox, d_dst_off, d_src_off, d_src, d_dst,
                                                   total);
    HIFDE_LAUNCH_CHECK();
}
template void pack_boxes<double>(int, const long long*, const long long*, const double*, double*,
                                 long long, hipStream_t);
template void pack_boxes<float>(int, const long long*, const long long*, const double*, float*,
                                long long, hipStream_t);

void convert_d2f(const double* src, float* dst, long long n, hipStream_t s) {
    if (n <= 0) return;
    k_convert<double, float><<<grid_for(n), 256, 0, s>>>(src, dst, n);
    HIFDE_LAUNCH_CHECK();
}

template void batched_gj_inverse<double>(int, const int*, const long long*, const int*,
                                         const long long*, double*, int*, GJWork&, GemmWork&,
                                         hipStream_t, bool);
template void batched_gj_inverse<float>(int, const int*, const long long*, const int*,
                                        const long long*, float*, int*, GJWork&, GemmWork&,
                                        hipStream_t, bool);

template void gemm_run<double, double>(const GemmPlan&, std::pair<int, int>, const double*,
                                       const double*, double*, double, double, hipStream_t,
                                       const int*);
template void gemm_run<float, float>(const GemmPlan&, std::pair<int, int>, const float*,
                                     const float*, float*, float, float, hipStream_t, const int*);
template void gemm_run<float, double>(const GemmPlan&, std::pair<int, int>, const float*,
                                      const float*, double*, float, float, hipStream_t,
                                      const int*);
template void batched_gemm_desc<double, double>(bool, const std::vector<GemmDesc>&,
                                                const double*, const double*, double*, double,
                                                double, GemmWork&, hipStream_t, const int*);
template void batched_gemm_desc<float, float>(bool, const std::vector<GemmDesc>&, const float*,
                                              const float*, float*, float, float, GemmWork&,
                                              hipStream_t, const int*);
template void batched_gemm_desc<float, double>(bool, const std::vector<GemmDesc>&, const float*,
                                               const float*, double*, float, float, GemmWork&,
                                               hipStream_t, const int*);
template void batched_gemm<double>(bool, int, const int*, const int*, const int*,
                                   const long long*, const long long*, const long long*,
                                   const double*, const double*, double*, double, GemmWork&,
                                   hipStream_t, bool);
template void batched_gemm<float>(bool, int, const int*, const int*, const int*,
                                  const long long*, const long long*, const long long*,
                                  const float*, const float*, float*, float, GemmWork&,
                                  hipStream_t, bool);
template void copy_blocks<double>(const std::vector<CopyDesc>&, const std::vector<long long>&,
                                  const double*, double*, GemmWork&, hipStream_t);
template void copy_blocks<float>(const std::vector<CopyDesc>&, const std::vector<long long>&,
                                 const double*, float*, GemmWork&, hipStream_t);

}  // namespace hifde
