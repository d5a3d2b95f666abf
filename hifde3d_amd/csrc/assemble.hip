// Device problem assembly: the 7-point stencil of SetupStencil
// (include/setup_stencil.hpp:6-76) built as CSR directly in HBM from device
// nodal tensors, and the synthetic-input generators of the workloads
// (SplitMix64 uniforms, Box-Muller normals, a = 10^U, the lognormal Gaussian
// random field).  The reference fills its tensors by hand with a = 1, V = 0
// (tests/HIFDE3D/main.cpp:59-74) and has no generators; these follow the
// definitions of hifde3d_amd/problems.py value for value.
//
// Integer and ordering work (row counts, CSR layout, entry order) is exact;
// the arithmetic of the stencil is the host assembly's, operation for
// operation, with contraction into FMAs disabled so that the CSR values are
// bitwise those of the host path.  The generators' transcendental functions
// (log, sin, cos, pow, exp) are the device library's, within an ulp or two of
// the host's.

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "hifde_internal.hpp"

namespace hifde {
namespace {

constexpr unsigned long long kGolden = 0x9E3779B97F4A7C15ull;

// t-th SplitMix64 output (t >= 1) of a seed (the generator's published
// finaliser on seed + t * golden gamma).
__device__ __forceinline__ unsigned long long splitmix(unsigned long long seed,
                                                       unsigned long long t) {
    unsigned long long z = seed + t * kGolden;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// 53-bit uniform in [0, 1) from output number idx + 1 (idx 0-based)
__device__ __forceinline__ double unif(unsigned long long seed, unsigned long long idx) {
    return (double)(splitmix(seed, idx + 1) >> 11) * (1.0 / 9007199254740992.0);
}
// Box-Muller normal number e of the sequence: pair e / 2 from uniforms
// 2p (radius, through 1 - u) and 2p + 1 (angle); even e takes the cosine
__device__ __forceinline__ double normal_at(unsigned long long seed, unsigned long long e) {
#pragma clang fp contract(off)
    const unsigned long long p = e >> 1;
    const double u1 = 1.0 - unif(seed, 2 * p);
    const double u2 = unif(seed, 2 * p + 1);
    const double rad = sqrt(-2.0 * log(u1));
    const double ang = (2.0 * M_PI) * u2;
    return (e & 1) ? rad * sin(ang) : rad * cos(ang);
}

__global__ void k_gen_uniform(double* out, long long n, unsigned long long seed, double lo,
                              double hi, int pow10) {
#pragma clang fp contract(off)
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const double s = lo + (hi - lo) * unif(seed, (unsigned long long)t);
        out[t] = pow10 ? pow(10.0, s) : s;
    }
}

__global__ void k_gen_normal(double* out, long long n, unsigned long long seed) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        out[t] = normal_at(seed, (unsigned long long)t);
}

// Right-hand sides: N(0,1) on the interior DOFs in linear-index order
// (column c takes numbers c * N^3 ...), zero on the boundary planes.
__global__ void k_gen_rhs(double* out, int NC, int nrhs, long long ld, unsigned long long seed) {
    const long long n3 = (long long)NC * NC * NC;
    const int N = NC - 1;
    const long long nint = (long long)N * N * N;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n3 * nrhs;
         t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / n3, p = t - c * n3;
        const int i = (int)(p % NC), j = (int)((p / NC) % NC), k = (int)(p / ((long long)NC * NC));
        double v = 0.0;
        if (i > 0 && j > 0 && k > 0) {
            const long long rank = (i - 1) + (long long)N * ((j - 1) + (long long)N * (k - 1));
            v = normal_at(seed, (unsigned long long)(c * nint + rank));
        }
        out[c * ld + p] = v;
    }
}

// --- stencil -------------------------------------------------------------
__device__ __forceinline__ int stencil_row_len(int i, int j, int k, int N) {
    if (i < 1 || j < 1 || k < 1 || i > N || j > N || k > N) return 0;
    return 1 + (i > 1) + (i < N) + (j > 1) + (j < N) + (k > 1) + (k < N);
}

__global__ void k_stencil_count(int NC, int* len) {
    const long long n3 = (long long)NC * NC * NC;
    const int N = NC - 1;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n3;
         p += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(p % NC), j = (int)((p / NC) % NC), k = (int)(p / ((long long)NC * NC));
        len[p] = stencil_row_len(i, j, k, N);
    }
}

// One row per thread, in the host assembly's order: diagonal, then the
// neighbours i-1, i+1, j-1, j+1, k-1, k+1 that are unknowns.  Nodal tensors
// have m = NC + 1 points per side, index i + m (j + m k).
__global__ void k_stencil_fill(int NC, const double* __restrict__ a,
                               const double* __restrict__ V, int symmetric, double hh,
                               const int* __restrict__ rp, int* __restrict__ ci,
                               double* __restrict__ val, int* __restrict__ rr) {
#pragma clang fp contract(off)
    const long long n3 = (long long)NC * NC * NC;
    const int N = NC - 1;
    const long long m = NC + 1;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n3;
         p += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(p % NC), j = (int)((p / NC) % NC), k = (int)(p / ((long long)NC * NC));
        if (stencil_row_len(i, j, k, N) == 0) continue;
        auto T = [&](const double* t, int x, int y, int z) { return t[x + m * (y + m * z)]; };
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                              {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        const double ap = T(a, i, j, k);
        double d = 0.0;
        if (!symmetric) {
            d += (T(a, i - 1, j, k) + T(a, i + 1, j, k)) / hh;
            d += (T(a, i, j - 1, k) + T(a, i, j + 1, k)) / hh;
            d += (T(a, i, j, k - 1) + T(a, i, j, k + 1)) / hh;
        } else {
            for (int t = 0; t < 6; ++t) d += 0.5 * (ap + T(a, nb[t][0], nb[t][1], nb[t][2])) / hh;
        }
        d += T(V, i, j, k);
        int q = rp[p];
        ci[q] = (int)p;
        val[q] = d;
        rr[q] = (int)p;
        ++q;
        for (int t = 0; t < 6; ++t) {
            const int x = nb[t][0], y = nb[t][1], z = nb[t][2];
            if (x < 1 || x > N || y < 1 || y > N || z < 1 || z > N) continue;
            const double aq = T(a, x, y, z);
            ci[q] = x + NC * (y + NC * z);
            val[q] = symmetric ? -0.5 * (ap + aq) / hh : -aq / hh;
            rr[q] = (int)p;
            ++q;
        }
    }
}

// flag = 0 when some entry differs bitwise from its transpose
__global__ void k_csr_symmetric(long long nd, const int* __restrict__ rp,
                                const int* __restrict__ ci, const double* __restrict__ val,
                                int* flag) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nd;
         i += (long long)gridDim.x * blockDim.x) {
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            const int j = ci[q];
            bool found = false;
            for (int r = rp[j]; r < rp[j + 1]; ++r)
                if (ci[r] == (int)i) {
                    found = val[r] == val[q];
                    break;
                }
            if (!found) atomicAnd(flag, 0);
        }
    }
}

// --- lognormal field -----------------------------------------------------
// kk[d] = (1/P) sum_q s1(q) cos(2 pi q d / P): the inverse DFT of the square
// root of the 1-D factor of the Gaussian covariance's spectral density,
// s1(q) = exp(-corr^2 w_q^2 / 4) with w_q the angular frequency of bin q on
// spacing h (numpy.fft.fftfreq order).  K1(i, x) = kk[(i - x) mod P] is the
// circulant factor of the separable filter; the kernel writes K1 for rows
// i < m (m x P, ld m) and its transpose (P x m, ld P).
__global__ void k_grf_filter(int P, int m, double h, double corr, double* K1, double* K1T) {
    extern __shared__ double kk[];
    for (int d = threadIdx.x; d < P; d += blockDim.x) {
        double acc = 0.0;
        for (int q = 0; q < P; ++q) {
            const int fq = q < (P + 1) / 2 ? q : q - P;
            const double w = (double)fq / (P * h) * (2.0 * M_PI);
            const double s1 = exp(-0.25 * corr * corr * w * w);
            const long long ph = ((long long)q * d) % P;
            acc += s1 * cos((2.0 * M_PI) * (double)ph / (double)P);
        }
        kk[d] = acc / P;
    }
    __syncthreads();
    for (long long e = threadIdx.x; e < (long long)m * P; e += blockDim.x) {
        const int i = (int)(e % m), x = (int)(e / m);
        const double v = kk[((i - x) % P + P) % P];
        K1[i + (long long)m * x] = v;
        K1T[x + (long long)P * i] = v;
    }
}

constexpr int kRedThreads = 256;
// per-block partial sums of x (mode 0) or of (x - mean)^2 (mode 1)
__global__ void __launch_bounds__(kRedThreads) k_partial_sum(const double* __restrict__ x,
                                                             long long n, int mode,
                                                             const double* __restrict__ mean,
                                                             double* part) {
    __shared__ double sh[kRedThreads];
    double s = 0.0;
    const double mu = mode ? *mean : 0.0;
    for (long long t = blockIdx.x * (long long)kRedThreads + threadIdx.x; t < n;
         t += (long long)gridDim.x * kRedThreads) {
        const double v = x[t] - mu;
        s += mode ? v * v : v;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
// out[0] = (sum of part) / n, in a fixed order
__global__ void k_final_mean(const double* __restrict__ part, int np, long long n, double* out) {
    __shared__ double sh[kRedThreads];
    double s = 0.0;
    for (int t = threadIdx.x; t < np; t += kRedThreads) s += part[t];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0] / (double)n;
}
// out = exp((g - mean) / sqrt(var)) (lognormal) or (g - mean) / sqrt(var)
__global__ void k_normalize(const double* __restrict__ g, long long n,
                            const double* __restrict__ mv, int lognormal, double* out) {
#pragma clang fp contract(off)
    const double mu = mv[0], sd = sqrt(mv[1]);
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const double v = (g[t] - mu) / sd;
        out[t] = lognormal ? exp(v) : v;
    }
}

// --- packed batches ----------------------------------------------------
// nprob independent problems of grid NC (unknowns 1..NC-1 per axis) stored
// as the octants of one grid of M = 2 NC points per side: octant o holds
// problem o at offset NC * (o & 1, o >> 1 & 1, o >> 2 & 1).  The planes at
// coordinate NC separate the octants; their points, and every point of an
// octant past nprob, carry identity rows with no couplings.
struct PackPoint {
    int kind;  // 0 boundary (no row), 1 identity row, 2 unknown of problem o
    int o, i, j, k;
};
__device__ __forceinline__ PackPoint pack_point(long long p, int NC, int nprob) {
    const int M = 2 * NC;
    const int x = (int)(p % M), y = (int)((p / M) % M), z = (int)(p / ((long long)M * M));
    PackPoint q{0, -1, 0, 0, 0};
    if (x == 0 || y == 0 || z == 0) return q;
    if (x == NC || y == NC || z == NC) {
        q.kind = 1;
        return q;
    }
    q.o = (x > NC) + 2 * (y > NC) + 4 * (z > NC);
    if (q.o >= nprob) {
        q.kind = 1;
        q.o = -1;
        return q;
    }
    q.kind = 2;
    q.i = x > NC ? x - NC : x;
    q.j = y > NC ? y - NC : y;
    q.k = z > NC ? z - NC : z;
    return q;
}

__global__ void k_pack_count(int NC, int nprob, int* len, unsigned char* seg) {
    const int M = 2 * NC;
    const long long n3 = (long long)M * M * M;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n3;
         p += (long long)gridDim.x * blockDim.x) {
        const PackPoint q = pack_point(p, NC, nprob);
        len[p] = q.kind == 0 ? 0 : q.kind == 1 ? 1 : stencil_row_len(q.i, q.j, q.k, NC - 1);
        seg[p] = q.kind == 2 ? (unsigned char)q.o : (unsigned char)255;
    }
}

// Rows in the standalone assembly's order and arithmetic (k_stencil_fill),
// columns relabelled into the packed grid; problem o's nodal tensors at
// a[o], V[o] ((NC+1)^3 each).
__global__ void k_pack_fill(int NC, int nprob, const double* const* __restrict__ a,
                            const double* const* __restrict__ V, int symmetric, double hh,
                            const int* __restrict__ rp, int* __restrict__ ci,
                            double* __restrict__ val, int* __restrict__ rr) {
#pragma clang fp contract(off)
    const int M = 2 * NC, N = NC - 1;
    const long long n3 = (long long)M * M * M;
    const long long m = NC + 1;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n3;
         p += (long long)gridDim.x * blockDim.x) {
        const PackPoint q = pack_point(p, NC, nprob);
        if (q.kind == 0) continue;
        int w = rp[p];
        if (q.kind == 1) {
            ci[w] = (int)p;
            val[w] = 1.0;
            rr[w] = (int)p;
            continue;
        }
        const double* A = a[q.o];
        const double* Vo = V[q.o];
        const int i = q.i, j = q.j, k = q.k;
        const int ox = (int)(p % M) - i, oy = (int)((p / M) % M) - j,
                  oz = (int)(p / ((long long)M * M)) - k;
        auto T = [&](const double* t, int x, int y, int z) { return t[x + m * (y + m * z)]; };
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                              {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        const double ap = T(A, i, j, k);
        double d = 0.0;
        if (!symmetric) {
            d += (T(A, i - 1, j, k) + T(A, i + 1, j, k)) / hh;
            d += (T(A, i, j - 1, k) + T(A, i, j + 1, k)) / hh;
            d += (T(A, i, j, k - 1) + T(A, i, j, k + 1)) / hh;
        } else {
            for (int t = 0; t < 6; ++t) d += 0.5 * (ap + T(A, nb[t][0], nb[t][1], nb[t][2])) / hh;
        }
        d += T(Vo, i, j, k);
        ci[w] = (int)p;
        val[w] = d;
        rr[w] = (int)p;
        ++w;
        for (int t = 0; t < 6; ++t) {
            const int x = nb[t][0], y = nb[t][1], z = nb[t][2];
            if (x < 1 || x > N || y < 1 || y > N || z < 1 || z > N) continue;
            const double aq = T(A, x, y, z);
            ci[w] = (x + ox) + M * ((y + oy) + M * (z + oz));
            val[w] = symmetric ? -0.5 * (ap + aq) / hh : -aq / hh;
            rr[w] = (int)p;
            ++w;
        }
    }
}

// dir 0: packed column c <- problem o's vector c (sub + (o R + c) ld_sub,
// standalone layout) on each problem's unknowns, 0 elsewhere.  dir 1: the
// reverse, 0 on each problem's boundary planes.
__global__ void k_pack_vectors(int NC, int nprob, int R, double* sub, long long ld_sub,
                               double* packed, long long ld_packed, int dir) {
    const int M = 2 * NC;
    const long long n3 = (long long)M * M * M, s3 = (long long)NC * NC * NC;
    if (dir == 0) {
        for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n3 * R;
             e += (long long)gridDim.x * blockDim.x) {
            const int c = (int)(e / n3);
            const long long p = e - c * n3;
            const PackPoint q = pack_point(p, NC, nprob);
            packed[c * ld_packed + p] =
                q.kind == 2 ? sub[((long long)q.o * R + c) * ld_sub + q.i +
                                  (long long)NC * (q.j + (long long)NC * q.k)]
                            : 0.0;
        }
    } else {
        for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
             e < s3 * R * nprob; e += (long long)gridDim.x * blockDim.x) {
            const long long oc = e / s3;
            const long long l = e - oc * s3;
            const int o = (int)(oc / R), c = (int)(oc - (long long)o * R);
            const int i = (int)(l % NC), j = (int)((l / NC) % NC), k = (int)(l / ((long long)NC * NC));
            double v = 0.0;
            if (i > 0 && j > 0 && k > 0) {
                const int x = i + NC * (o & 1), y = j + NC * ((o >> 1) & 1),
                          z = k + NC * ((o >> 2) & 1);
                v = packed[c * ld_packed + x + (long long)M * (y + (long long)M * z)];
            }
            sub[oc * ld_sub + l] = v;
        }
    }
}

int grid_of(long long n) {
    const long long b = (n + 255) / 256;
    return (int)std::max<long long>(1, std::min<long long>(b, 65536));
}

}  // namespace

void gen_uniform(double* d_out, long long n, unsigned long long seed, double lo, double hi,
                 bool pow10, hipStream_t s) {
    if (n <= 0) return;
    k_gen_uniform<<<grid_of(n), 256, 0, s>>>(d_out, n, seed, lo, hi, pow10 ? 1 : 0);
    HIFDE_LAUNCH_CHECK();
}

void gen_normal(double* d_out, long long n, unsigned long long seed, hipStream_t s) {
    if (n <= 0) return;
    k_gen_normal<<<grid_of(n), 256, 0, s>>>(d_out, n, seed);
    HIFDE_LAUNCH_CHECK();
}

void gen_rhs(double* d_out, int NC, int nrhs, long long ld, unsigned long long seed,
             hipStream_t s) {
    const long long n = (long long)NC * NC * NC * nrhs;
    if (n <= 0) return;
    k_gen_rhs<<<grid_of(n), 256, 0, s>>>(d_out, NC, nrhs, ld, seed);
    HIFDE_LAUNCH_CHECK();
}

// Gaussian random field on the (NC+1)^3 nodes, as problems.gaussian_field:
// white noise on the doubled grid P = 2(NC+1), filtered by the square root
// of the Gaussian spectral density (a separable circulant filter, applied as
// three FP64 GEMMs, one per axis, restricted to the first m = NC+1 points of
// each), then shifted to mean 0 and scaled to unit (population) variance.
// With lognormal, exp of the field.  Temporary buffers are freed before
// return, so the stream is synchronized.
void gen_gaussian_field(double* d_out, int NC, unsigned long long seed, double corr,
                        bool lognormal, hipStream_t s) {
    const int m = NC + 1, P = 2 * m;
    const long long P3 = (long long)P * P * P, m3 = (long long)m * m * m;
    DevBuf<double> W, T1, T2, G, K1, K1T, part, mv;
    W.reserve(P3);
    T1.reserve((size_t)m * P * P);
    T2.reserve((size_t)m * m * P);
    G.reserve(m3);
    K1.reserve((size_t)m * P);
    K1T.reserve((size_t)m * P);
    gen_normal(W.p, P3, seed, s);
    const size_t shm = (size_t)P * sizeof(double);
    if (shm > 64 * 1024) throw Error(-3, "gaussian field: filter too long for LDS");
    k_grf_filter<<<1, 256, shm, s>>>(P, m, 1.0 / NC, corr, K1.p, K1T.p);
    HIFDE_LAUNCH_CHECK();
    GemmWork gws;
    // x: T1 (m x P^2) = K1 (m x P) W (P x P^2)
    batched_gemm_desc<double>(false, {{0, 0, 0, m, P * P, P, m, P, m}}, K1.p, W.p, T1.p, 1.0, 0.0,
                              gws, s);
    // y: per z, T2_z (m x m) = T1_z (m x P) K1^T (P x m)
    std::vector<GemmDesc> dy;
    for (int z = 0; z < P; ++z)
        dy.push_back({(long long)z * m * P, 0, (long long)z * m * m, m, m, P, m, P, m});
    batched_gemm_desc<double>(false, dy, T1.p, K1T.p, T2.p, 1.0, 0.0, gws, s);
    // z: G (m^2 x m) = T2 (m^2 x P) K1^T (P x m)
    batched_gemm_desc<double>(false, {{0, 0, 0, m * m, m, P, m * m, P, m * m}}, T2.p, K1T.p, G.p,
                              1.0, 0.0, gws, s);
    const int nb = 1024;
    part.reserve(nb);
    mv.reserve(2);
    k_partial_sum<<<nb, kRedThreads, 0, s>>>(G.p, m3, 0, nullptr, part.p);
    k_final_mean<<<1, kRedThreads, 0, s>>>(part.p, nb, m3, mv.p);
    k_partial_sum<<<nb, kRedThreads, 0, s>>>(G.p, m3, 1, mv.p, part.p);
    k_final_mean<<<1, kRedThreads, 0, s>>>(part.p, nb, m3, mv.p + 1);
    k_normalize<<<grid_of(m3), 256, 0, s>>>(G.p, m3, mv.p, lognormal ? 1 : 0, d_out);
    HIFDE_LAUNCH_CHECK();
    HIFDE_CHECK(hipStreamSynchronize(s));
}

// CSR of the stencil (rows in linear order, entries in the host assembly's
// order) from device nodal tensors; returns nnz.  rp: NC^3 + 1; ci, val, rr
// grown to nnz.
long long stencil_csr(int NC, const double* d_a, const double* d_V, int symmetric,
                      DevBuf<int>& rp, DevBuf<int>& ci, DevBuf<double>& val, DevBuf<int>& rr,
                      hipStream_t s) {
    const long long n3 = (long long)NC * NC * NC;
    DevBuf<int> len;
    len.reserve(n3 + 1);
    rp.reserve(n3 + 1);
    HIFDE_CHECK(hipMemsetAsync(len.p + n3, 0, sizeof(int), s));
    k_stencil_count<<<grid_of(n3), 256, 0, s>>>(NC, len.p);
    HIFDE_LAUNCH_CHECK();
    size_t tmp_bytes = 0;
    HIFDE_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, len.p, rp.p, (int)(n3 + 1),
                                                 s));
    DevBuf<unsigned char> tmp;
    tmp.reserve(tmp_bytes);
    HIFDE_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, len.p, rp.p, (int)(n3 + 1), s));
    int nnz = 0;
    HIFDE_CHECK(hipMemcpyAsync(&nnz, rp.p + n3, sizeof(int), hipMemcpyDeviceToHost, s));
    HIFDE_CHECK(hipStreamSynchronize(s));
    ci.reserve(std::max(nnz, 1));
    val.reserve(std::max(nnz, 1));
    rr.reserve(std::max(nnz, 1));
    const double h = 1.0 / NC;
    k_stencil_fill<<<grid_of(n3), 256, 0, s>>>(NC, d_a, d_V, symmetric, h * h, rp.p, ci.p, val.p,
                                               rr.p);
    HIFDE_LAUNCH_CHECK();
    return nnz;
}

long long pack_stencil_csr(int NC, int nprob, const double* const* d_a_ptrs,
                           const double* const* d_V_ptrs, int symmetric, DevBuf<int>& rp,
                           DevBuf<int>& ci, DevBuf<double>& val, DevBuf<int>& rr,
                           DevBuf<unsigned char>& seg, hipStream_t s) {
    const long long M = 2LL * NC, n3 = M * M * M;
    DevBuf<int> len;
    len.reserve(n3 + 1);
    rp.reserve(n3 + 1);
    seg.reserve(n3);
    HIFDE_CHECK(hipMemsetAsync(len.p + n3, 0, sizeof(int), s));
    k_pack_count<<<grid_of(n3), 256, 0, s>>>(NC, nprob, len.p, seg.p);
    HIFDE_LAUNCH_CHECK();
    size_t tmp_bytes = 0;
    HIFDE_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, len.p, rp.p, (int)(n3 + 1),
                                                 s));
    DevBuf<unsigned char> tmp;
    tmp.reserve(tmp_bytes);
    HIFDE_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, len.p, rp.p, (int)(n3 + 1), s));
    int nnz = 0;
    HIFDE_CHECK(hipMemcpyAsync(&nnz, rp.p + n3, sizeof(int), hipMemcpyDeviceToHost, s));
    HIFDE_CHECK(hipStreamSynchronize(s));
    ci.reserve(std::max(nnz, 1));
    val.reserve(std::max(nnz, 1));
    rr.reserve(std::max(nnz, 1));
    const double h = 1.0 / NC;
    k_pack_fill<<<grid_of(n3), 256, 0, s>>>(NC, nprob, d_a_ptrs, d_V_ptrs, symmetric, h * h, rp.p,
                                            ci.p, val.p, rr.p);
    HIFDE_LAUNCH_CHECK();
    return nnz;
}

void pack_vectors(int NC, int nprob, int R, double* sub, long long ld_sub, double* packed,
                  long long ld_packed, bool unpack, hipStream_t s) {
    const long long M = 2LL * NC;
    const long long n = unpack ? (long long)NC * NC * NC * R * nprob : M * M * M * R;
    if (n <= 0) return;
    k_pack_vectors<<<grid_of(n), 256, 0, s>>>(NC, nprob, R, sub, ld_sub, packed, ld_packed,
                                              unpack ? 1 : 0);
    HIFDE_LAUNCH_CHECK();
}

bool csr_symmetric(long long nd, const int* d_rp, const int* d_ci, const double* d_val,
                   hipStream_t s) {
    DevBuf<int> flag;
    flag.reserve(1);
    const int one = 1;
    HIFDE_CHECK(hipMemcpyAsync(flag.p, &one, sizeof(int), hipMemcpyHostToDevice, s));
    k_csr_symmetric<<<grid_of(nd), 256, 0, s>>>(nd, d_rp, d_ci, d_val, flag.p);
    HIFDE_LAUNCH_CHECK();
    int h = 0;
    HIFDE_CHECK(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIFDE_CHECK(hipStreamSynchronize(s));
    return h != 0;
}

}  // namespace hifde
