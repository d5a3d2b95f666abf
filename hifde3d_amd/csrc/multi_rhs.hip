// Several right-hand sides at once: the apply sweeps as batched GEMMs on
// packed per-box blocks, and the vector kernels of the multi-RHS PCG.
//
// The single-vector apply (apply_pcg.hip) is a GEMV per box.  With R
// right-hand sides each box's step is a GEMM: the inputs of every box are
// gathered into a packed column-major block (nin x R), multiplied by the box's
// factor block on the matrix cores, and the outputs scattered back (added,
// written, or parked in the per-pass side buffers of the cell sweep).  The
// factor blocks are read once for all R vectors.
//
// A block of R vectors is addressed through a layout {sg, sc}: entry (dof g,
// column c) at U[g * sg + c * sc].  The caller's blocks are column-major
// ({1, ldu}); the sweeps run on an interleaved working copy ({R, 1}), where
// the R values of one DOF are contiguous, so every gather and scatter of a
// box's DOFs reads and writes whole cache lines.

#include "hifde_internal.hpp"

namespace hifde {

namespace {

template <class O>
__device__ __forceinline__ int find_slot(const O* __restrict__ off, int n, long long e) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((long long)off[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

inline unsigned grid1(long long n) {
    long long g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > (1LL << 20)) g = 1LL << 20;
    return (unsigned)g;
}

// pack[off[b] R + i + c nin_b] = U(gid[off[b] + i], c); one thread per
// entry, all R columns
template <class T>
__global__ void k_rhs_gather(int nbox, const int* __restrict__ off, const int* __restrict__ gid,
                             const double* __restrict__ U, long long sg, long long sc, int R,
                             T* pack) {
    const long long ne = off[nbox];
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ne;
         t += (long long)gridDim.x * blockDim.x) {
        const int b = find_slot(off, nbox, t);
        const long long nb = off[b + 1] - off[b];
        T* dst = pack + (long long)off[b] * R + (t - off[b]);
        const double* src = U + gid[t] * sg;
        for (int c = 0; c < R; ++c) dst[c * nb] = (T)src[c * sc];
    }
}

// mode 0: U(g, c) += alpha * pack;  1: U(g, c) = pack;
// 2: side(pass flag, g, c) = alpha * pack, the side buffer laid out like U,
// one block per pass flag, `pstride` apart
template <class T>
__global__ void k_rhs_scatter(int nbox, const int* __restrict__ off, const int* __restrict__ gid,
                              const T* __restrict__ pack, int R, double* U, long long sg,
                              long long sc, double alpha, int mode,
                              const unsigned char* __restrict__ sp, long long pstride) {
    const long long ne = off[nbox];
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ne;
         t += (long long)gridDim.x * blockDim.x) {
        const int b = find_slot(off, nbox, t);
        const long long nb = off[b + 1] - off[b];
        const T* src = pack + (long long)off[b] * R + (t - off[b]);
        double* dst = U + gid[t] * sg;
        if (mode == 0) {
            for (int c = 0; c < R; ++c) dst[c * sc] += alpha * (double)src[c * nb];
        } else if (mode == 1) {
            for (int c = 0; c < R; ++c) dst[c * sc] = (double)src[c * nb];
        } else {
            dst += (long long)sp[t] * pstride;
            for (int c = 0; c < R; ++c) dst[c * sc] = alpha * (double)src[c * nb];
        }
    }
}

// Fused gather -> product -> scatter for boxes of moderate size.  One
// workgroup forms 64 outputs of one box for CB = 4 CPT right-hand sides
// (columns c0 = blockIdx.y CB ...): the box's input entries are staged
// through LDS 64 at a time together with the matching 64 x 64 block of M
// (read along its contiguous dimension), and every thread accumulates one
// output for CPT columns in FP64.  op as in k_apply_tiles; op 4 writes U in
// place, which needs every box to fit one tile (the host checks r <= 64).
// With side non-null the outputs go to the pass-flag side buffers (laid out
// like U, one block per flag, pstride apart).  When the columns of a DOF are
// contiguous (sc == 1) the staging loads walk the columns first.
template <class T, int CPT>
__global__ void __launch_bounds__(256) k_apply_multi(
    int nbox, const int* __restrict__ tile_off, const int* __restrict__ dr,
    const int* __restrict__ ds, const int* __restrict__ red_off,
    const int* __restrict__ skel_off, const int* __restrict__ red_gid,
    const int* __restrict__ skel_gid, const unsigned char* __restrict__ skel_pass,
    const long long* __restrict__ off, const T* __restrict__ Mb, int op, double alpha, double* U,
    long long sg, long long sc, int R, double* side, long long pstride) {
    constexpr int CB = 4 * CPT;
    const int t = blockIdx.x;
    const int b = find_slot(tile_off, nbox, t);
    const int tile = t - tile_off[b];
    const int c0 = blockIdx.y * CB;
    const int r = dr[b], s = ds[b];
    if (r == 0 || (op != 4 && s == 0)) return;
    const T* M = Mb + off[b];
    const bool in_red = (op == 0 || op == 3 || op == 4);
    const bool out_red = (op == 1 || op == 2 || op == 4);
    const bool trans = (op == 0 || op == 2);
    const int nin = in_red ? r : s, nout = out_red ? r : s;
    const int* ig = in_red ? red_gid + red_off[b] : skel_gid + skel_off[b];
    const int* og = out_red ? red_gid + red_off[b] : skel_gid + skel_off[b];
    const int o0 = tile * 64;
    const int cnt = min(64, nout - o0);
    __shared__ double xs[CB][65];  // padded: column-first staging stores hit distinct banks
    __shared__ T Ms[64][65];  // Ms[output][input]
    const int tid = threadIdx.x, o = tid & 63, g = tid >> 6;
    double acc[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) acc[q] = 0.0;
    for (int j0 = 0; j0 < nin; j0 += 64) {
        const int kb = min(64, nin - j0);
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int e = tid + 256 * q;
            const int j = sc == 1 ? e / CB : e & 63, c = sc == 1 ? e % CB : e >> 6;
            double v = 0.0;
            if (j < kb && c0 + c < R) v = U[ig[j0 + j] * sg + (long long)(c0 + c) * sc];
            xs[c][j] = v;
        }
        if (!trans) {  // M: nout x nin, leading dim nout
#pragma unroll 4
            for (int q = 0; q < 16; ++q) {
                const int e = tid + 256 * q;
                const int oo = e & 63, jj = e >> 6;
                Ms[oo][jj] = (oo < cnt && jj < kb) ? M[(size_t)(j0 + jj) * nout + o0 + oo] : (T)0;
            }
        } else {  // M: nin x nout, leading dim nin
#pragma unroll 4
            for (int q = 0; q < 16; ++q) {
                const int e = tid + 256 * q;
                const int jj = e & 63, oo = e >> 6;
                Ms[oo][jj] = (oo < cnt && jj < kb) ? M[(size_t)(o0 + oo) * nin + j0 + jj] : (T)0;
            }
        }
        __syncthreads();
        for (int jj = 0; jj < kb; ++jj) {
            const double m = (double)Ms[o][jj];
#pragma unroll
            for (int q = 0; q < CPT; ++q) acc[q] += m * xs[g * CPT + q][jj];
        }
        __syncthreads();
    }
    if (o >= cnt) return;
    const long long dof = (long long)og[o0 + o] * sg;
    double* dst = side ? side + (long long)skel_pass[skel_off[b] + o0 + o] * pstride + dof
                       : U + dof;
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = c0 + g * CPT + q;
        if (c >= R) break;
        if (op == 4) dst[c * sc] = acc[q];
        else if (side) dst[c * sc] = alpha * acc[q];
        else dst[c * sc] += alpha * acc[q];
    }
}

// U(g, c) += side(p, g, c) for the entries with pass flag p (side laid out
// like U, pstride per flag); one thread per entry, all columns
__global__ void k_rhs_side_add(long long ne, const int* __restrict__ gid,
                               const unsigned char* __restrict__ sp,
                               const double* __restrict__ side, long long pstride, int R, int p,
                               double* U, long long sg, long long sc) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ne;
         t += (long long)gridDim.x * blockDim.x) {
        if ((int)sp[t] != p) continue;
        const long long g = (long long)gid[t] * sg;
        const double* src = side + (long long)p * pstride + g;
        double* dst = U + g;
        for (int c = 0; c < R; ++c) dst[c * sc] += src[c * sc];
    }
}

// Column-major (n x R, leading dim ld) <-> interleaved (W[g R + c]) through
// a 64-DOF x R LDS tile (R <= 64 per launch column chunk): both sides are
// read and written along their contiguous direction.
__global__ void __launch_bounds__(256) k_to_interleaved(long long n, int R, const double* U,
                                                        long long ld, double* W) {
    __shared__ double tile[64][65];
    const long long g0 = (long long)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    const int cn = min(64, R - c0);
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int i = e & 63, c = e >> 6;
        if (c < cn && g0 + i < n) tile[c][i] = U[g0 + i + (long long)(c0 + c) * ld];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * cn; e += 256) {
        const int c = e % cn, i = e / cn;
        if (g0 + i < n) W[(g0 + i) * R + c0 + c] = tile[c][i];
    }
}
__global__ void __launch_bounds__(256) k_from_interleaved(long long n, int R, const double* W,
                                                          double* U, long long ld) {
    __shared__ double tile[64][65];
    const long long g0 = (long long)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    const int cn = min(64, R - c0);
    for (int e = threadIdx.x; e < 64 * cn; e += 256) {
        const int c = e % cn, i = e / cn;
        if (g0 + i < n) tile[c][i] = W[(g0 + i) * R + c0 + c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int i = e & 63, c = e >> 6;
        if (c < cn && g0 + i < n) U[g0 + i + (long long)(c0 + c) * ld] = tile[c][i];
    }
}

// zero the rows of the boundary DOFs (a coordinate 0) of an interleaved block
__global__ void k_zero_boundary_il(int NC, int R, double* W) {
    const long long n = (long long)NC * NC * NC;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n * R;
         e += (long long)gridDim.x * blockDim.x) {
        const long long g = e / R;
        const int i = (int)(g % NC), j = (int)((g / NC) % NC), k = (int)(g / ((long long)NC * NC));
        if (i == 0 || j == 0 || k == 0) W[e] = 0.0;
    }
}

// Y(:, c) = A X(:, c) for R columns (CSR A, n rows)
__global__ void k_csr_spmm(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ v, const double* __restrict__ X, long long ldx,
                           double* Y, long long ldy, int R) {
    const long long total = (long long)n * R;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e / n);
        const int i = (int)(e - (long long)c * n);
        const double* x = X + c * ldx;
        double acc = 0.0;
        for (int q = rp[i]; q < rp[i + 1]; ++q) acc += v[q] * x[ci[q]];
        Y[i + c * ldy] = acc;
    }
}

// part[c * nblk + blk] = partial sum of X(:, c) . Y(:, c)
__global__ void __launch_bounds__(256) k_dot_cols(long long n, const double* __restrict__ X,
                                                  long long ldx, const double* __restrict__ Y,
                                                  long long ldy, double* part) {
    const int c = blockIdx.y;
    const double* x = X + c * ldx;
    const double* y = Y + c * ldy;
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        acc += x[i] * y[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[(long long)c * gridDim.x + blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ void k_sum_cols(const double* __restrict__ part, int nblk, int R, double* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= R) return;
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += part[(long long)c * nblk + b];
    out[c] = acc;
}

// Dots per column and segment: part[(c S + s) nblk + block] sums x_i y_i
// over the entries with seg[i] == s (entries of segment 255 belong to none).
template <int S>
__global__ void __launch_bounds__(256) k_dot_cols_seg(long long n, const double* __restrict__ X,
                                                      long long ldx,
                                                      const double* __restrict__ Y,
                                                      long long ldy,
                                                      const unsigned char* __restrict__ seg,
                                                      int nseg, double* part) {
    const int c = blockIdx.y;
    const double* x = X + c * ldx;
    const double* y = Y + c * ldy;
    double acc[S];
#pragma unroll
    for (int t = 0; t < S; ++t) acc[t] = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int sg = seg[i];
        const double v = x[i] * y[i];
#pragma unroll
        for (int t = 0; t < S; ++t) acc[t] += sg == t ? v : 0.0;
    }
    __shared__ double ws[4][S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
        double a = acc[t];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][t] = a;
    }
    __syncthreads();
    if (threadIdx.x < nseg)
        part[((long long)c * nseg + threadIdx.x) * gridDim.x + blockIdx.x] =
            ws[0][threadIdx.x] + ws[1][threadIdx.x] + ws[2][threadIdx.x] + ws[3][threadIdx.x];
}

// As k_axpy_cols with one scalar per column and segment, a[c S + seg[i]];
// entries of no segment are left unchanged.
__global__ void k_axpy_cols_seg(long long n, const double* __restrict__ a, const double* X,
                                long long ldx, double* Y, long long ldy, int R, int xpby,
                                const unsigned char* __restrict__ seg, int nseg) {
    const long long total = n * R;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e / n);
        const long long i = e - (long long)c * n;
        const int sg = seg[i];
        if (sg >= nseg) continue;
        const double al = a[c * nseg + sg];
        const double x = X[i + c * ldx];
        double& y = Y[i + c * ldy];
        y = xpby ? x + al * y : y + al * x;
    }
}

// Y(:, c) = Y(:, c) + a[c] X(:, c)   (xpby == 0)
// Y(:, c) = X(:, c) + a[c] Y(:, c)   (xpby == 1)
__global__ void k_axpy_cols(long long n, const double* __restrict__ a, const double* X,
                            long long ldx, double* Y, long long ldy, int R, int xpby) {
    const long long total = n * R;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e / n);
        const long long i = e - (long long)c * n;
        const double x = X[i + c * ldx];
        double& y = Y[i + c * ldy];
        y = xpby ? x + a[c] * y : y + a[c] * x;
    }
}

}  // namespace

template <class T>
void rhs_gather(int nbox, long long nentries, const int* d_off, const int* d_gid, const double* U,
                RhsLayout ly, int R, T* pack, hipStream_t s) {
    if (nbox <= 0 || nentries <= 0) return;
    k_rhs_gather<T><<<grid1(nentries), 256, 0, s>>>(nbox, d_off, d_gid, U, ly.sg, ly.sc, R, pack);
    HIFDE_LAUNCH_CHECK();
}

template <class T>
void rhs_scatter(int nbox, long long nentries, const int* d_off, const int* d_gid, const T* pack,
                 int R, double* U, RhsLayout ly, double alpha, int mode, const unsigned char* d_sp,
                 long long pstride, hipStream_t s) {
    if (nbox <= 0 || nentries <= 0) return;
    k_rhs_scatter<T><<<grid1(nentries), 256, 0, s>>>(nbox, d_off, d_gid, pack, R, U, ly.sg, ly.sc,
                                                     alpha, mode, d_sp, pstride);
    HIFDE_LAUNCH_CHECK();
}

void to_interleaved(long long n, int R, const double* U, long long ld, double* W, hipStream_t s) {
    if (n <= 0 || R <= 0) return;
    k_to_interleaved<<<dim3((unsigned)((n + 63) / 64), (R + 63) / 64), 256, 0, s>>>(n, R, U, ld, W);
    HIFDE_LAUNCH_CHECK();
}

void from_interleaved(long long n, int R, const double* W, double* U, long long ld,
                      hipStream_t s) {
    if (n <= 0 || R <= 0) return;
    k_from_interleaved<<<dim3((unsigned)((n + 63) / 64), (R + 63) / 64), 256, 0, s>>>(n, R, W, U,
                                                                                     ld);
    HIFDE_LAUNCH_CHECK();
}

void zero_boundary_il(int NC, int R, double* W, hipStream_t s) {
    k_zero_boundary_il<<<grid1((long long)NC * NC * NC * R), 256, 0, s>>>(NC, R, W);
    HIFDE_LAUNCH_CHECK();
}

template <class T>
void apply_multi_fused(int nbox, int ntiles, const int* d_tile_off, const int* d_r, const int* d_s,
                       const int* d_red_off, const int* d_skel_off, const int* d_red_gid,
                       const int* d_skel_gid, const unsigned char* d_skel_pass,
                       const long long* d_off, const T* d_M, int op, double alpha, double* U,
                       RhsLayout ly, int R, double* side, long long pstride, double flops,
                       double bytes, hipStream_t s) {
    if (nbox <= 0 || ntiles <= 0 || R <= 0) return;
    KernelTimer* kt = active_timer();
    if (kt)
        kt->begin(s, sizeof(T) == sizeof(float) ? "k_apply_multi<float>" : "k_apply_multi<double>");
    const int cpt = R <= 4 ? 1 : R <= 8 ? 2 : R <= 16 ? 4 : 8;
    const dim3 grid(ntiles, (R + 4 * cpt - 1) / (4 * cpt));
#define HIFDE_AM(CPT)                                                                           \
    k_apply_multi<T, CPT><<<grid, 256, 0, s>>>(nbox, d_tile_off, d_r, d_s, d_red_off,           \
                                               d_skel_off, d_red_gid, d_skel_gid, d_skel_pass, \
                                               d_off, d_M, op, alpha, U, ly.sg, ly.sc, R,  \
                                               side, pstride)
    if (cpt == 1) HIFDE_AM(1);
    else if (cpt == 2) HIFDE_AM(2);
    else if (cpt == 4) HIFDE_AM(4);
    else HIFDE_AM(8);
#undef HIFDE_AM
    HIFDE_LAUNCH_CHECK();
    if (kt) kt->end(s, flops, bytes);
}

void rhs_side_add(long long nentries, const int* d_gid, const unsigned char* d_sp,
                  const double* side, long long pstride, int R, double* U, RhsLayout ly,
                  hipStream_t s) {
    if (nentries <= 0) return;
    for (int p = 0; p < 2; ++p) {
        k_rhs_side_add<<<grid1(nentries), 256, 0, s>>>(nentries, d_gid, d_sp, side, pstride, R, p,
                                                       U, ly.sg, ly.sc);
        HIFDE_LAUNCH_CHECK();
    }
}

void csr_spmm(int n, const int* d_rp, const int* d_ci, const double* d_v, const double* X,
              long long ldx, double* Y, long long ldy, int R, hipStream_t s) {
    k_csr_spmm<<<grid1((long long)n * R), 256, 0, s>>>(n, d_rp, d_ci, d_v, X, ldx, Y, ldy, R);
    HIFDE_LAUNCH_CHECK();
}

void dot_cols(long long n, const double* X, long long ldx, const double* Y, long long ldy, int R,
              double* d_part, int nblk, double* d_out, hipStream_t s) {
    k_dot_cols<<<dim3(nblk, R), 256, 0, s>>>(n, X, ldx, Y, ldy, d_part);
    HIFDE_LAUNCH_CHECK();
    k_sum_cols<<<(R + 63) / 64, 64, 0, s>>>(d_part, nblk, R, d_out);
    HIFDE_LAUNCH_CHECK();
}

void dot_cols_seg(long long n, const double* X, long long ldx, const double* Y, long long ldy,
                  int R, const unsigned char* d_seg, int nseg, double* d_part, int nblk,
                  double* d_out, hipStream_t s) {
    if (nseg < 1 || nseg > 8) throw Error(-1, "dot_cols_seg: 1 to 8 segments");
    k_dot_cols_seg<8><<<dim3(nblk, R), 256, 0, s>>>(n, X, ldx, Y, ldy, d_seg, nseg, d_part);
    HIFDE_LAUNCH_CHECK();
    k_sum_cols<<<(R * nseg + 63) / 64, 64, 0, s>>>(d_part, nblk, R * nseg, d_out);
    HIFDE_LAUNCH_CHECK();
}

void axpy_cols_seg(long long n, const double* d_a, const double* X, long long ldx, double* Y,
                   long long ldy, int R, bool xpby, const unsigned char* d_seg, int nseg,
                   hipStream_t s) {
    k_axpy_cols_seg<<<grid1(n * R), 256, 0, s>>>(n, d_a, X, ldx, Y, ldy, R, xpby ? 1 : 0, d_seg,
                                                 nseg);
    HIFDE_LAUNCH_CHECK();
}

void axpy_cols(long long n, const double* d_a, const double* X, long long ldx, double* Y,
               long long ldy, int R, bool xpby, hipStream_t s) {
    k_axpy_cols<<<grid1(n * R), 256, 0, s>>>(n, d_a, X, ldx, Y, ldy, R, xpby ? 1 : 0);
    HIFDE_LAUNCH_CHECK();
}

template void rhs_gather<float>(int, long long, const int*, const int*, const double*, RhsLayout,
                                int, float*, hipStream_t);
template void rhs_gather<double>(int, long long, const int*, const int*, const double*, RhsLayout,
                                 int, double*, hipStream_t);
template void apply_multi_fused<float>(int, int, const int*, const int*, const int*, const int*,
                                       const int*, const int*, const int*, const unsigned char*,
                                       const long long*, const float*, int, double, double*,
                                       RhsLayout, int, double*, long long, double, double,
                                       hipStream_t);
template void apply_multi_fused<double>(int, int, const int*, const int*, const int*, const int*,
                                        const int*, const int*, const int*, const unsigned char*,
                                        const long long*, const double*, int, double, double*,
                                        RhsLayout, int, double*, long long, double, double,
                                        hipStream_t);
template void rhs_scatter<float>(int, long long, const int*, const int*, const float*, int,
                                 double*, RhsLayout, double, int, const unsigned char*, long long,
                                 hipStream_t);
template void rhs_scatter<double>(int, long long, const int*, const int*, const double*, int,
                                  double*, RhsLayout, double, int, const unsigned char*, long long,
                                  hipStream_t);

}  // namespace hifde
