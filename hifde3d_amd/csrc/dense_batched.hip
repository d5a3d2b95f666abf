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
// Explicit inverse by Gauss-Jordan elimination with row partial pivoting (the
// pivot of column c is the first row >= c of largest magnitude, as in getrf).
// In-place GJ keeps [A | I] in one array: once column c has been eliminated
// it holds column c of the accumulated elimination matrix, and the final
// column interchanges (in reverse pivot order) turn that into A^{-1}.
//
// Boxes with n <= kSmallGJ are inverted in LDS, one workgroup per box.
// Larger boxes use a blocked form: for each panel K of kGJB columns
//   1. k_bgj_panel   eliminates the panel columns for all rows (one WG/box),
//                    recording pivots; the panel then holds G(:, K), the
//                    panel's elimination matrix.
//   2. k_bgj_swap    applies the panel's row interchanges to every other
//                    column and copies the pivot rows B = A(K, J) aside.
//   3. k_bgj_update  A(i, J) = [i not in K] A(i, J) + G(i, K) B  for J != K,
//                    a rank-kGJB update tiled over the whole GPU.
// followed by k_bgj_colswap for the column interchanges.
// ---------------------------------------------------------------------------
constexpr int kSmallGJ = 64;
constexpr int kGJB = 32;

template <class T>
__global__ void __launch_bounds__(128) k_gj_small(const int* __restrict__ dims,
                                                  const long long* __restrict__ off, T* mats,
                                                  int* info) {
    const int b = blockIdx.x;
    const int n = dims[b];
    if (n == 0 || n > kSmallGJ) return;
    T* A = mats + off[b];
    __shared__ T S[kSmallGJ][kSmallGJ + 1];  // S[col][row]
    __shared__ T colk[kSmallGJ], rowk[kSmallGJ];
    __shared__ int piv[kSmallGJ];
    __shared__ T redv[4];
    __shared__ int redi[4];
    __shared__ T s_pv;
    __shared__ int s_pi;
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int e = tid; e < n * n; e += bs) S[e / n][e % n] = A[e];
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        T best = (T)-1;
        int bi = n;
        for (int i = k + tid; i < n; i += bs) {
            T v = fabs(S[k][i]);
            if (v > best) { best = v; bi = i; }
        }
        block_argmax(best, bi, redv, redi, &s_pv, &s_pi);
        const int p = s_pi;
        if (!(s_pv > (T)0)) {
            if (tid == 0) info[b] = k + 1;
            return;
        }
        if (p != k)
            for (int c = tid; c < n; c += bs) {
                T t = S[c][k];
                S[c][k] = S[c][p];
                S[c][p] = t;
            }
        if (tid == 0) piv[k] = p;
        __syncthreads();
        const T d = (T)1 / S[k][k];
        for (int i = tid; i < n; i += bs) {
            colk[i] = S[k][i];
            rowk[i] = (i == k) ? d : S[i][k] * d;
        }
        __syncthreads();
        for (int e = tid; e < n * n; e += bs) {
            int c = e / n, i = e - c * n;
            if (i == k) {
                S[c][i] = rowk[c];
            } else {
                T a = (c == k) ? (T)0 : S[c][i];
                S[c][i] = a - colk[i] * rowk[c];
            }
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        int p = piv[k];
        if (p != k)
            for (int i = tid; i < n; i += bs) {
                T t = S[k][i];
                S[k][i] = S[p][i];
                S[p][i] = t;
            }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += bs) A[e] = S[e / n][e % n];
    if (tid == 0) info[b] = 0;
}

// Panel elimination, one workgroup per large box.  Dynamic shared memory:
// max_n T (pivot column) .
template <class T>
__global__ void __launch_bounds__(256) k_bgj_panel(const int* __restrict__ dims,
                                                   const long long* __restrict__ off, T* mats,
                                                   int* piv, const int* __restrict__ piv_off,
                                                   int step, int* info) {
    const int b = blockIdx.x;
    const int n = dims[b];
    const int k0 = step * kGJB;
    if (n <= kSmallGJ || k0 >= n || info[b]) return;
    const int kb = min(kGJB, n - k0);
    T* A = mats + off[b];
    int* pv = piv + piv_off[b];
    extern __shared__ unsigned char smem[];
    T* colk = reinterpret_cast<T*>(smem);
    __shared__ T rowk[kGJB];
    __shared__ T redv[4];
    __shared__ int redi[4];
    __shared__ T s_pv;
    __shared__ int s_pi;
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int c = k0; c < k0 + kb; ++c) {
        T best = (T)-1;
        int bi = n;
        for (int i = c + tid; i < n; i += bs) {
            T v = fabs(A[i + (size_t)c * n]);
            if (v > best) { best = v; bi = i; }
        }
        block_argmax(best, bi, redv, redi, &s_pv, &s_pi);
        const int p = s_pi;
        if (!(s_pv > (T)0)) {
            if (tid == 0) info[b] = c + 1;
            return;
        }
        if (p != c && tid < kb) {
            T* cj = A + (size_t)(k0 + tid) * n;
            T t = cj[c];
            cj[c] = cj[p];
            cj[p] = t;
        }
        if (tid == 0) pv[c] = p;
        __syncthreads();
        const T d = (T)1 / A[c + (size_t)c * n];
        for (int i = tid; i < n; i += bs) colk[i] = A[i + (size_t)c * n];
        if (tid < kb) {
            int j = k0 + tid;
            rowk[tid] = (j == c) ? d : A[c + (size_t)j * n] * d;
        }
        __syncthreads();
        for (int jj = 0; jj < kb; ++jj) {
            const int j = k0 + jj;
            T* cj = A + (size_t)j * n;
            const T rk = rowk[jj];
            for (int i = tid; i < n; i += bs) {
                if (i == c) {
                    cj[i] = rk;
                } else {
                    T a = (j == c) ? (T)0 : cj[i];
                    cj[i] = a - colk[i] * rk;
                }
            }
        }
        __syncthreads();
    }
}

// Row interchanges of the panel applied to column j (j outside the panel),
// then B(:, j) = A(K, j).  Grid: x over columns, y over boxes.
template <class T>
__global__ void __launch_bounds__(256) k_bgj_swap(const int* __restrict__ dims,
                                                  const long long* __restrict__ off, T* mats,
                                                  const int* __restrict__ piv,
                                                  const int* __restrict__ piv_off, int step,
                                                  T* Bws, const long long* __restrict__ b_off,
                                                  const int* __restrict__ info) {
    const int b = blockIdx.y;
    const int n = dims[b];
    const int k0 = step * kGJB;
    if (n <= kSmallGJ || k0 >= n || info[b]) return;
    const int kb = min(kGJB, n - k0);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || (j >= k0 && j < k0 + kb)) return;
    T* cj = mats + off[b] + (size_t)j * n;
    const int* pv = piv + piv_off[b];
    for (int c = k0; c < k0 + kb; ++c) {
        int p = pv[c];
        if (p != c) {
            T t = cj[c];
            cj[c] = cj[p];
            cj[p] = t;
        }
    }
    T* B = Bws + b_off[b] + (size_t)j * kGJB;
    for (int c = 0; c < kb; ++c) B[c] = cj[k0 + c];
}

// A(i, j) = [i not in K] A(i, j) + sum_c G(i, c) B(c, j), j not in K.
template <class T>
__global__ void __launch_bounds__(256) k_bgj_update(int nbox, const int* __restrict__ tile_off,
                                                    const int* __restrict__ dims,
                                                    const long long* __restrict__ off, T* mats,
                                                    const T* __restrict__ Bws,
                                                    const long long* __restrict__ b_off,
                                                    int step, const int* __restrict__ info) {
    const int t = blockIdx.x;
    const int b = find_box(tile_off, nbox, t);
    const int n = dims[b];
    const int k0 = step * kGJB;
    if (n <= kSmallGJ || k0 >= n || info[b]) return;
    const int kb = min(kGJB, n - k0);
    const int local = t - tile_off[b];
    const int tm = (n + 63) / 64;
    const int i0 = (local % tm) * 64, j0 = (local / tm) * 64;
    if (j0 + 64 > k0 && j0 < k0 + kb && j0 >= k0 && j0 + 64 <= k0 + kb) return;
    T* A = mats + off[b];
    const T* B = Bws + b_off[b];
    __shared__ T Gs[kGJB][64 + 1];
    __shared__ T Bs[kGJB][64 + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < 64 * kGJB; e += 256) {
        int r = e & 63, c = e >> 6;
        int gi = i0 + r;
        Gs[c][r] = (gi < n && c < kb) ? A[gi + (size_t)(k0 + c) * n] : (T)0;
        int cb = e & (kGJB - 1), jb = e / kGJB;
        int gj = j0 + jb;
        bool skip = gj >= n || (gj >= k0 && gj < k0 + kb) || cb >= kb;
        Bs[cb][jb] = skip ? (T)0 : B[cb + (size_t)gj * kGJB];
    }
    __syncthreads();
    const int tx = tid & 15, ty = tid >> 4;
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (T)0;
    for (int c = 0; c < kb; ++c) {
        T a[4], bv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Gs[c][tx + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; ++j) bv[j] = Bs[c][ty + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], bv[j], acc[i][j]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int gj = j0 + ty + 16 * j;
        if (gj >= n || (gj >= k0 && gj < k0 + kb)) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int gi = i0 + tx + 16 * i;
            if (gi >= n) continue;
            T* e = A + gi + (size_t)gj * n;
            const bool inK = gi >= k0 && gi < k0 + kb;
            *e = (inK ? (T)0 : *e) + acc[i][j];
        }
    }
}

// Column interchanges in reverse pivot order; one thread per row.
template <class T>
__global__ void __launch_bounds__(256) k_bgj_colswap(const int* __restrict__ dims,
                                                     const long long* __restrict__ off, T* mats,
                                                     const int* __restrict__ piv,
                                                     const int* __restrict__ piv_off,
                                                     int* info) {
    const int b = blockIdx.y;
    const int n = dims[b];
    if (n <= kSmallGJ || info[b]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T* A = mats + off[b];
    const int* pv = piv + piv_off[b];
    for (int k = n - 1; k >= 0; --k) {
        int p = pv[k];
        if (p != k) {
            T t = A[i + (size_t)k * n];
            A[i + (size_t)k * n] = A[i + (size_t)p * n];
            A[i + (size_t)p * n] = t;
        }
    }
}

// C(m x n) = alpha * A(m x k) * B(k x n), all column-major with ld = rows.
constexpr int BM = 64, BN = 64, BK = 16;

template <class T>
__global__ void __launch_bounds__(256) k_gemm_nn(int nbox, const int* __restrict__ tile_off,
                                                 const int* __restrict__ dm,
                                                 const int* __restrict__ dn,
                                                 const int* __restrict__ dk,
                                                 const long long* __restrict__ offA,
                                                 const long long* __restrict__ offB,
                                                 const long long* __restrict__ offC,
                                                 const T* __restrict__ Ab,
                                                 const T* __restrict__ Bb, T* __restrict__ Cb,
                                                 T alpha) {
    int t = blockIdx.x;
    int b = find_box(tile_off, nbox, t);
    int local = t - tile_off[b];
    int m = dm[b], n = dn[b], k = dk[b];
    int tm = (m + BM - 1) / BM;
    int i0 = (local % tm) * BM, j0 = (local / tm) * BN;
    const T* A = Ab + offA[b];
    const T* B = Bb + offB[b];
    T* C = Cb + offC[b];
    __shared__ T As[BK][BM + 1];
    __shared__ T Bs[BK][BN + 1];
    int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (T)0;
    for (int k0 = 0; k0 < k; k0 += BK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int e = threadIdx.x + 256 * q;
            int r = e & (BM - 1), c = e >> 6;  // A tile: 64 rows x 16 cols
            int gi = i0 + r, gk = k0 + c;
            As[c][r] = (gi < m && gk < k) ? A[gi + (size_t)gk * m] : (T)0;
            int rb = e & (BK - 1), cb = e >> 4;  // B tile: 16 rows x 64 cols
            int gkb = k0 + rb, gj = j0 + cb;
            Bs[rb][cb] = (gkb < k && gj < n) ? B[gkb + (size_t)gj * k] : (T)0;
        }
        __syncthreads();
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
        int gj = j0 + ty + 16 * j;
        if (gj >= n) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int gi = i0 + tx + 16 * i;
            if (gi < m) C[gi + (size_t)gj * m] = alpha * acc[i][j];
        }
    }
}

// Column-pivoted QR + interpolative decomposition of one face per workgroup.
// Dynamic shared memory: vn1[n], vn2[n] (double), jp[n], pos[n] (int).
__global__ void __launch_bounds__(256) k_interp_decomp(const int* __restrict__ dm,
                                                       const int* __restrict__ dn,
                                                       const long long* __restrict__ offM,
                                                       double* Mb, double eps, int* perm,
                                                       const int* __restrict__ perm_off,
                                                       int* kout, double* Wb,
                                                       const long long* __restrict__ offW) {
    const int b = blockIdx.x;
    const int m = dm[b], n = dn[b];
    if (n == 0) {
        if (threadIdx.x == 0) kout[b] = 0;
        return;
    }
    const int lda = m > 1 ? m : 1;
    double* M = Mb + offM[b];
    extern __shared__ unsigned char smem[];
    double* vn1 = reinterpret_cast<double*>(smem);
    double* vn2 = vn1 + n;
    int* jp = reinterpret_cast<int*>(vn2 + n);
    int* pos = jp + n;
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
    const int k = s_k, q = n - k;
    int* pb = perm + perm_off[b];
    for (int t = tid; t < n; t += bs) pb[t] = jp[pos[t]];
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

// Blocks of Rot^T C Rot with Rot = I - E_{skel,red} W, C = A(F,F) (n x n):
//   A12 = C_sr - C_ss W                  (k x q)
//   A21 = C_rs - W^T C_ss                (q x k)
//   A22 = C_rr - C_rs W - W^T A12        (q x q)
// Computed in FP64; Tsr (FP64 scratch, k x q) holds A12 for the A22 step.
template <class T>
__global__ void __launch_bounds__(256) k_face_transform(
    const int* __restrict__ dn, const int* __restrict__ dk, const int* __restrict__ perm,
    const int* __restrict__ perm_off, const double* __restrict__ Cb,
    const long long* __restrict__ offC, const double* __restrict__ Wb,
    const long long* __restrict__ offW, double* Tsrb, T* A22b, T* A21b, T* A12b,
    const long long* __restrict__ off_rr, const long long* __restrict__ off_rs) {
    const int b = blockIdx.x;
    const int n = dn[b], k = dk[b], q = n - k;
    if (n == 0) return;
    const int* p = perm + perm_off[b];
    const int* sk = p;
    const int* rd = p + k;
    const double* C = Cb + offC[b];
    const double* W = Wb + offW[b];
    double* Tsr = Tsrb + offW[b];
    T* A22 = A22b + off_rr[b];
    T* A21 = A21b + off_rs[b];
    T* A12 = A12b + off_rs[b];
    const int tid = threadIdx.x, bs = blockDim.x;
    auto Cv = [&](int i, int j) { return C[i + (size_t)j * n]; };
    for (int e = tid; e < k * q; e += bs) {
        int i = e % k, jj = e / k;
        double v = Cv(sk[i], rd[jj]);
        for (int l = 0; l < k; ++l) v -= Cv(sk[i], sk[l]) * W[l + (size_t)jj * k];
        Tsr[e] = v;
        A12[e] = (T)v;
    }
    for (int e = tid; e < q * k; e += bs) {
        int ii = e % q, j = e / q;
        double v = Cv(rd[ii], sk[j]);
        for (int l = 0; l < k; ++l) v -= W[l + (size_t)ii * k] * Cv(sk[l], sk[j]);
        A21[e] = (T)v;
    }
    __syncthreads();
    for (int e = tid; e < q * q; e += bs) {
        int ii = e % q, jj = e / q;
        double v = Cv(rd[ii], rd[jj]);
        for (int l = 0; l < k; ++l) v -= Cv(rd[ii], sk[l]) * W[l + (size_t)jj * k];
        for (int l = 0; l < k; ++l) v -= W[l + (size_t)ii * k] * Tsr[l + (size_t)jj * k];
        A22[e] = (T)v;
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

int grid_for(long long n, int block = 256, int cap = 1 << 16) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace

template <class T>
void batched_gj_inverse(int nbox, const int* h_n, const int* d_n, const long long* d_off, T* d_A,
                        int* d_info, GJWork& ws, hipStream_t s) {
    if (nbox <= 0) return;
    int max_n = 0;
    bool any_small = false;
    std::vector<int> piv_off(nbox + 1, 0), tile_off(nbox + 1, 0);
    std::vector<long long> b_off(nbox + 1, 0);
    for (int b = 0; b < nbox; ++b) {
        int n = h_n[b];
        max_n = std::max(max_n, n);
        bool big = n > kSmallGJ;
        any_small |= (n > 0 && !big);
        piv_off[b + 1] = piv_off[b] + (big ? n : 0);
        b_off[b + 1] = b_off[b] + (big ? (long long)kGJB * n : 0);
        int tm = big ? (n + 63) / 64 : 0;
        tile_off[b + 1] = tile_off[b] + tm * tm;
    }
    if (any_small) {
        k_gj_small<T><<<nbox, 128, 0, s>>>(d_n, d_off, d_A, d_info);
        HIFDE_LAUNCH_CHECK();
    }
    if (max_n <= kSmallGJ) return;
    ws.piv.reserve(piv_off[nbox]);
    ws.piv_off.upload(piv_off, s);
    ws.tile_off.upload(tile_off, s);
    ws.b_off.upload(b_off, s);
    ws.B.reserve((size_t)b_off[nbox] * sizeof(T));
    T* Bws = reinterpret_cast<T*>(ws.B.p);
    const int steps = (max_n + kGJB - 1) / kGJB;
    const size_t shm = (size_t)max_n * sizeof(T);
    if (shm > 150 * 1024) throw Error(-3, "gj_inverse: box too large");
    dim3 gsw((max_n + 255) / 256, nbox);
    for (int st = 0; st < steps; ++st) {
        k_bgj_panel<T><<<nbox, 256, shm, s>>>(d_n, d_off, d_A, ws.piv.p, ws.piv_off.p, st, d_info);
        k_bgj_swap<T><<<gsw, 256, 0, s>>>(d_n, d_off, d_A, ws.piv.p, ws.piv_off.p, st, Bws,
                                          ws.b_off.p, d_info);
        k_bgj_update<T><<<tile_off[nbox], 256, 0, s>>>(nbox, ws.tile_off.p, d_n, d_off, d_A, Bws,
                                                       ws.b_off.p, st, d_info);
    }
    k_bgj_colswap<T><<<gsw, 256, 0, s>>>(d_n, d_off, d_A, ws.piv.p, ws.piv_off.p, d_info);
    HIFDE_LAUNCH_CHECK();
}

template <class T>
void batched_gemm_nn(int nbox, const int* h_m, const int* h_n, const int* h_k, const int* d_m,
                     const int* d_n, const int* d_k, const long long* d_offA,
                     const long long* d_offB, const long long* d_offC, const T* d_A,
                     const T* d_B, T* d_C, T alpha, DevBuf<int>& tile_scratch, hipStream_t s) {
    if (nbox <= 0) return;
    std::vector<int> toff(nbox + 1, 0);
    for (int b = 0; b < nbox; ++b) {
        long long tiles = (h_m[b] > 0 && h_n[b] > 0)
                              ? (long long)((h_m[b] + BM - 1) / BM) * ((h_n[b] + BN - 1) / BN)
                              : 0;
        toff[b + 1] = toff[b] + (int)tiles;
    }
    int total = toff[nbox];
    if (total == 0) return;
    tile_scratch.upload(toff, s);
    k_gemm_nn<T><<<total, 256, 0, s>>>(nbox, tile_scratch.p, d_m, d_n, d_k, d_offA, d_offB,
                                       d_offC, d_A, d_B, d_C, alpha);
    HIFDE_LAUNCH_CHECK();
}

void batched_interp_decomp(int nface, const int* d_m, const int* d_n, const long long* d_offM,
                           double* d_M, double eps, int* d_perm, const int* d_perm_off, int* d_k,
                           double* d_W, const long long* d_offW, int max_n, int max_m,
                           hipStream_t s) {
    if (nface <= 0) return;
    size_t shm = (size_t)max_n * (2 * sizeof(double) + 2 * sizeof(int));
    if (shm > 150 * 1024) throw Error(-3, "interp_decomp: face too large for LDS");
    int threads = (max_m <= 128 && max_n <= 16) ? 64 : 256;
    k_interp_decomp<<<nface, threads, shm, s>>>(d_m, d_n, d_offM, d_M, eps, d_perm, d_perm_off,
                                                d_k, d_W, d_offW);
    HIFDE_LAUNCH_CHECK();
}

template <class T>
void batched_face_transform(int nface, const int* d_n, const int* d_k, const int* d_perm,
                            const int* d_perm_off, const double* d_C, const long long* d_offC,
                            const double* d_W, const long long* d_offW, double* d_Tsr, T* d_A22,
                            T* d_A21, T* d_A12, const long long* d_off_rr,
                            const long long* d_off_rs, hipStream_t s) {
    if (nface <= 0) return;
    k_face_transform<T><<<nface, 256, 0, s>>>(d_n, d_k, d_perm, d_perm_off, d_C, d_offC, d_W,
                                              d_offW, d_Tsr, d_A22, d_A21, d_A12, d_off_rr,
                                              d_off_rs);
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

template void batched_gj_inverse<double>(int, const int*, const int*, const long long*, double*,
                                         int*, GJWork&, hipStream_t);
template void batched_gj_inverse<float>(int, const int*, const int*, const long long*, float*, int*,
                                        GJWork&, hipStream_t);
template void batched_gemm_nn<double>(int, const int*, const int*, const int*, const int*,
                                      const int*, const int*, const long long*, const long long*,
                                      const long long*, const double*, const double*, double*,
                                      double, DevBuf<int>&, hipStream_t);
template void batched_gemm_nn<float>(int, const int*, const int*, const int*, const int*,
                                     const int*, const int*, const long long*, const long long*,
                                     const long long*, const float*, const float*, float*, float,
                                     DevBuf<int>&, hipStream_t);
template void batched_face_transform<double>(int, const int*, const int*, const int*, const int*,
                                             const double*, const long long*, const double*,
                                             const long long*, double*, double*, double*,
                                             double*, const long long*, const long long*,
                                             hipStream_t);
template void batched_face_transform<float>(int, const int*, const int*, const int*, const int*,
                                            const double*, const long long*, const double*,
                                            const long long*, double*, float*, float*, float*,
                                            const long long*, const long long*, hipStream_t);

}  // namespace hifde
