// Apply sweeps of the factorization and the vector kernels of the PCG driver.
//
// Apply restates HIFFactor::Apply (src/Factor.cpp:992-1081): every box of a
// level is a gather -> small GEMV -> scatter (GetSkeletonVector /
// UpdateSkeleton / UpdateRedundant / ApplyA22, :798-975).  Boxes of one
// level touch disjoint DOFs except the skeleton DOFs of a cell that lie on a
// face shared with the neighbouring cell; those writes are split into two
// passes (max-face DOFs first, then min-face DOFs) so every DOF receives its
// two contributions in the reference's box order.
//
// The factor blocks may be stored in float or double; the vector and all
// accumulations are double.

#include <algorithm>

#include "hifde_internal.hpp"

namespace hifde {

namespace {

constexpr int kWave = 64;
constexpr int kXcds = 8;

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

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

// One tile of 64 RW outputs of one box per workgroup (RW = 2: ops 1, 3 and
// 4 only, each lane forming two outputs 64 rows apart).
// op 0: u_s[j] += alpha * sum_i M(i,j) u_r[i]   M: r x s   (X^H u_red)
// op 1: u_r[i] += alpha * sum_j M(i,j) u_s[j]   M: r x s   (X u_skel)
// op 2: u_r[j] += alpha * sum_i M(i,j) u_s[i]   M: s x r   (W^H u_skel)
// op 3: u_s[i] += alpha * sum_j M(i,j) u_r[j]   M: s x r   (W u_red)
// op 4: out[r_i] = sum_j M(i,j) u_r[j]          M: r x r   (A22 or A22^{-1})
// Ops 0-3 read and write disjoint DOF classes, so they update u in place;
// op 4 reads u_in and writes u_out (a different vector).  pass >= 0 restricts
// op 0/3 writes to skeleton DOFs whose pass flag equals pass.
template <class T, int RW>
__global__ void __launch_bounds__(256) k_apply_tiles(
    int nbox, const int* __restrict__ tile_off, const int* __restrict__ dr,
    const int* __restrict__ ds, const int* __restrict__ red_off,
    const int* __restrict__ skel_off, const int* __restrict__ red_gid,
    const int* __restrict__ skel_gid, const unsigned char* __restrict__ skel_pass,
    const long long* __restrict__ off, const T* __restrict__ Mb, int op, int pass, double alpha,
    const double* u_in, double* u_out, long long side_n) {
    // Workgroups are dealt to the 8 XCDs round-robin; consecutive tiles (the
    // row tiles of one box, whose boundary cache lines they share) are given
    // to workgroups of the same XCD, so those lines are fetched into one L2.
    const int nb = (int)gridDim.x, per = nb / kXcds, rem = nb % kXcds;
    const int x = (int)blockIdx.x % kXcds;
    const int t = x * per + min(x, rem) + (int)blockIdx.x / kXcds;
    const int b = find_box(tile_off, nbox, t);
    const int tile = t - tile_off[b];
    const int r = dr[b], s = ds[b];
    if (r == 0 || (op != 4 && s == 0)) return;  // src/Factor.cpp:898-900, 937-939, 965-967
    const T* M = Mb + off[b];
    const int* rg = red_gid + red_off[b];
    const int* sg = skel_gid + skel_off[b];
    const unsigned char* sp = skel_pass ? skel_pass + skel_off[b] : nullptr;
    const bool in_red = (op == 0 || op == 3 || op == 4);
    const bool out_red = (op == 1 || op == 2 || op == 4);
    const bool trans = (op == 0 || op == 2);
    const int nin = in_red ? r : s, nout = out_red ? r : s;
    const int* ig = in_red ? rg : sg;
    const int* og = out_red ? rg : sg;
    extern __shared__ double xs[];
    __shared__ double part[4][64 * RW];
    const int tid = threadIdx.x;
    for (int q = tid; q < nin; q += 256) xs[q] = u_in[ig[q]];
    __syncthreads();
    const int o0 = tile * 64 * RW;
    const int cnt = min(64 * RW, nout - o0);
    double acc = 0.0;
    int o = -1;
    if (!trans) {
        // y_i = sum_j M(i,j) x_j ; M has nout rows.  Four column groups.
        // Each lane forms rows i + 64 q (q < RW); a wave's loads of one
        // column cover 64 RW consecutive rows (256 RW bytes in FP32).
        const int i = tid & 63, g = tid >> 6;
        constexpr int U = 16 / RW;  // columns in flight per thread
        double pacc[RW];
#pragma unroll
        for (int q = 0; q < RW; ++q) pacc[q] = 0.0;
        if (i < cnt) {
            // U independent sums per row keep 16 loads per thread in
            // flight; past the last column the loads repeat it (cached) and
            // are weighted by zero, and rows past the tile's last repeat
            // that row (their sums are not used), so the loop has no branches
            const T* Mi = M + o0 + i;
            int rof[RW];
#pragma unroll
            for (int q = 0; q < RW; ++q) rof[q] = min(64 * q, cnt - 1 - i);
            double a[U][RW];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int q = 0; q < RW; ++q) a[u][q] = 0.0;
            for (int j = g; j < nin; j += 4 * U) {
                T mv[U][RW];
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int q = 0; q < RW; ++q)
                        mv[u][q] = Mi[(size_t)min(j + 4 * u, nin - 1) * nout + rof[q]];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int jj = j + 4 * u;
                    const double xv = jj < nin ? xs[min(jj, nin - 1)] : 0.0;
#pragma unroll
                    for (int q = 0; q < RW; ++q) a[u][q] += (double)mv[u][q] * xv;
                }
            }
#pragma unroll
            for (int w = U / 2; w > 0; w >>= 1)
#pragma unroll
                for (int u = 0; u < w; ++u)
#pragma unroll
                    for (int q = 0; q < RW; ++q) a[u][q] += a[u + w][q];
#pragma unroll
            for (int q = 0; q < RW; ++q) pacc[q] = a[0][q];
        }
#pragma unroll
        for (int q = 0; q < RW; ++q) part[g][i + 64 * q] = pacc[q];
        __syncthreads();
        if (tid < cnt) {
            o = o0 + tid;
            acc = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        }
    } else if constexpr (RW != 1) {
        return;  // the host launches RW = 2 for ops 1, 3 and 4 only
    } else if (nin <= 64) {
        // y_j = sum_i M(i,j) x_i with short columns: one thread per output
        // (outputs of the other pass are not formed)
        if (tid < cnt && !(sp && pass >= 0 && (int)sp[o0 + tid] != pass)) {
            o = o0 + tid;
            const T* Mj = M + (size_t)o * nin;
            for (int i = 0; i < nin; ++i) acc += (double)Mj[i] * xs[i];
        }
    } else {
        // long columns: 64 x 64 blocks of M are read along the columns
        // (coalesced), transposed through LDS, and each thread sums a
        // quarter of every block for one output; outputs of the other pass
        // are not read at all
        __shared__ T blk[64][65];  // blk[output][row]
        const int oo = tid & 63, g = tid >> 6;
        bool keep = oo < cnt;
        if (keep && sp && pass >= 0 && (int)sp[o0 + oo] != pass) keep = false;
        // this thread loads rows ii = tid & 63 of the columns g + 4q, q < 16:
        // which of them it reads is fixed for the whole tile
        unsigned cmask = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int jj = g + 4 * q;
            bool c = jj < cnt;
            if (c && sp && pass >= 0 && (int)sp[o0 + jj] != pass) c = false;
            cmask |= (c ? 1u : 0u) << q;
        }
        double pacc = 0.0;
        // the next 64-row block is loaded while the current one is summed
        T v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int jj = g + 4 * q;
            v[q] = (oo < nin && ((cmask >> q) & 1u)) ? M[(size_t)(o0 + jj) * nin + oo] : (T)0;
        }
        for (int i0 = 0; i0 < nin; i0 += 64) {
            const int ib = min(64, nin - i0);
            const int i1 = i0 + 64;
            const bool next_ok = oo < nin - i1;
            T vn[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int jj = g + 4 * q;
                vn[q] = (next_ok && ((cmask >> q) & 1u)) ? M[(size_t)(o0 + jj) * nin + i1 + oo]
                                                         : (T)0;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) blk[g + 4 * q][oo] = v[q];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int ii = g * 16 + q;
                if (ii < ib) pacc += (double)blk[oo][ii] * xs[i0 + ii];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = vn[q];
        }
        part[g][oo] = pacc;
        __syncthreads();
        if (tid < cnt && keep) {
            o = o0 + tid;
            acc = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        }
    }
    if (o < 0) return;
    if (op == 4) {
        u_out[og[o]] = acc;
    } else if (side_n > 0) {
        // side buffers: the contribution goes to slot [pass flag][DOF]
        u_out[(long long)sp[o] * side_n + og[o]] = alpha * acc;
    } else {
        if (sp && pass >= 0 && (int)sp[o] != pass) return;
        u_out[og[o]] += alpha * acc;
    }
}

__global__ void k_zero_boundary(double* u, int NC) {
    long long n = (long long)NC * NC * NC;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) {
        int i = (int)(t % NC), j = (int)((t / NC) % NC), k = (int)(t / ((long long)NC * NC));
        if (i == 0 || j == 0 || k == 0) u[t] = 0.0;
    }
}

__global__ void k_csr_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ v, const double* __restrict__ x,
                           double* __restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) acc += v[e] * x[ci[e]];
    y[i] = acc;
}

// Row i's entries q < len read from the slot-major copy; the same products,
// summed in the same order, as k_csr_spmv.
__global__ void k_ell_spmv(int n, int width, const int* __restrict__ rp,
                           const double* __restrict__ ev, const int* __restrict__ ec,
                           const double* __restrict__ x, double* __restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int len = rp[i + 1] - rp[i];
    double acc = 0.0;
    for (int q = 0; q < len; ++q) {
        const long long e = (long long)q * n + i;
        acc += ev[e] * x[ec[e]];
    }
    y[i] = acc;
}

__global__ void k_csr_to_ell(int n, int width, const int* __restrict__ rp,
                             const int* __restrict__ ci, const double* __restrict__ v,
                             double* __restrict__ ev, int* __restrict__ ec) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = rp[i], len = rp[i + 1] - b;
    for (int q = 0; q < width; ++q) {
        const long long e = (long long)q * n + i;
        ev[e] = q < len ? v[b + q] : 0.0;
        ec[e] = q < len ? ci[b + q] : i;
    }
}

__global__ void k_fill_zero_int(int* p) {
    if (threadIdx.x == 0) p[threadIdx.x] = 0;
}

__global__ void k_max_row(int n, const int* __restrict__ rp, int* out) {
    __shared__ int m[256];
    int best = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        best = max(best, rp[i + 1] - rp[i]);
    m[threadIdx.x] = best;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) m[threadIdx.x] = max(m[threadIdx.x], m[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, m[0]);
}

constexpr int kDotThreads = 256;

__global__ void __launch_bounds__(256) k_dot_partial(const double* __restrict__ x,
                                                     const double* __restrict__ y, long long n,
                                                     double* part) {
    __shared__ double red[kDotThreads / kWave];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        acc += x[i] * y[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kDotThreads / kWave; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(256) k_reduce(const double* __restrict__ part, int nblk,
                                                double* out) {
    __shared__ double red[kDotThreads / kWave];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) acc += part[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kDotThreads / kWave; ++w) t += red[w];
        *out = t;
    }
}

__global__ void k_axpy(double a, const double* __restrict__ x, double* __restrict__ y,
                       long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}

__global__ void k_xpby(const double* __restrict__ x, double b, double* __restrict__ y,
                       long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = x[i] + b * y[i];
}

__global__ void k_copy(const double* __restrict__ x, double* __restrict__ y, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = x[i];
}

int grid_for(long long n, int block = 256, int cap = 4096) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace

// Boxes whose inputs and outputs are at most 64 long (the level-0 boxes):
// one wave per box, four boxes per workgroup.  The wave stages the box's
// inputs in LDS; lane o forms output o.  Same ops and pass rule as
// k_apply_tiles.
template <class T>
__global__ void __launch_bounds__(256) k_apply_small(
    int nbox, const int* __restrict__ dr, const int* __restrict__ ds,
    const int* __restrict__ red_off, const int* __restrict__ skel_off,
    const int* __restrict__ red_gid, const int* __restrict__ skel_gid,
    const unsigned char* __restrict__ skel_pass, const long long* __restrict__ off,
    const T* __restrict__ Mb, int op, int pass, double alpha, const double* u_in,
    double* u_out, long long side_n, int msz) {
    __shared__ double xs[4][64];
    extern __shared__ __attribute__((aligned(16))) unsigned char mstage[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int b = blockIdx.x * 4 + wv;
    if (b >= nbox) return;
    const int r = dr[b], s = ds[b];
    if (r == 0 || (op != 4 && s == 0)) return;
    const bool in_red = (op == 0 || op == 3 || op == 4);
    const bool out_red = (op == 1 || op == 2 || op == 4);
    const int nin = in_red ? r : s, nout = out_red ? r : s;
    const int* rg = red_gid + red_off[b];
    const int* sg = skel_gid + skel_off[b];
    if (lane < nin) xs[wv][lane] = u_in[(in_red ? rg : sg)[lane]];
    const T* M = Mb + off[b];
    if (msz > 0) {
        // The box's block is contiguous: the wave copies it into its LDS
        // region with coalesced loads (eight in flight per lane), and the
        // lanes then read their columns or rows from LDS.
        T* Mw = reinterpret_cast<T*>(mstage) + (size_t)wv * msz;
        const int cnt = nin * nout;
        int e = lane;
        for (; e + 7 * 64 < cnt; e += 8 * 64) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = M[e + 64 * u];
#pragma unroll
            for (int u = 0; u < 8; ++u) Mw[e + 64 * u] = v[u];
        }
        for (; e < cnt; e += 64) Mw[e] = M[e];
        M = Mw;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane >= nout) return;
    if (op == 0 && skel_pass && pass >= 0 && (int)skel_pass[skel_off[b] + lane] != pass) return;
    const double* x = xs[wv];
    double acc = 0.0;
    if (op == 0 || op == 2) {  // column lane of M (nin x nout)
        const T* Mj = M + (size_t)lane * nin;
        for (int i = 0; i < nin; ++i) acc += (double)Mj[i] * x[i];
    } else {  // row lane of M (nout x nin)
        for (int j = 0; j < nin; ++j) acc += (double)M[lane + (size_t)j * nout] * x[j];
    }
    const int g = out_red ? rg[lane] : sg[lane];
    if (op == 4) u_out[g] = acc;
    else if (side_n > 0) u_out[(long long)skel_pass[skel_off[b] + lane] * side_n + g] = alpha * acc;
    else u_out[g] += alpha * acc;
}

// u[g] += side[p][g] for the skeleton entries with pass flag p: the two
// contributions a shared face DOF receives, added in pass order.
__global__ void k_side_add(int n, const int* __restrict__ gid, const unsigned char* __restrict__ sp,
                           const double* __restrict__ side, long long side_n, int p, double* u) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n || (int)sp[t] != p) return;
    const int g = gid[t];
    u[g] += side[(long long)p * side_n + g];
}

// The same when every DOF has one pass-0 and one pass-1 entry: the pass-1
// entry adds both, u[g] = (u[g] + side[0][g]) + side[1][g].
__global__ void k_side_add_pair(int n, const int* __restrict__ gid,
                                const unsigned char* __restrict__ sp,
                                const double* __restrict__ side, long long side_n, double* u) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n || sp[t] != 1) return;
    const int g = gid[t];
    u[g] = (u[g] + side[g]) + side[side_n + g];
}

template <class T>
void apply_small(int nbox, int nout_total, const int* d_r, const int* d_s, const int* d_red_off,
                 const int* d_skel_off, const int* d_red_gid, const int* d_skel_gid,
                 const unsigned char* d_skel_pass, const long long* d_off, const T* d_M, int op,
                 int pass, double alpha, const double* d_in, double* d_out, double flops,
                 double bytes, hipStream_t s, long long side_n, int max_block) {
    if (nbox <= 0 || nout_total <= 0) return;
    // stage each box's block in LDS where four of the largest fit in 64 KB
    const int msz = (max_block > 0 && 4LL * max_block * sizeof(T) <= 65536) ? max_block : 0;
    const size_t shm = (size_t)4 * msz * sizeof(T);
    KernelTimer* kt = active_timer();
    if (kt)
        kt->begin(s, sizeof(T) == sizeof(float) ? "k_apply_small<float>" : "k_apply_small<double>");
    k_apply_small<T><<<(nbox + 3) / 4, 256, shm, s>>>(nbox, d_r, d_s, d_red_off, d_skel_off,
                                                      d_red_gid, d_skel_gid, d_skel_pass, d_off,
                                                      d_M, op, pass, alpha, d_in, d_out, side_n,
                                                      msz);
    HIFDE_LAUNCH_CHECK();
    if (kt) kt->end(s, flops, bytes);
}

template <class T>
void apply_tiles(int nbox, int ntiles, const int* d_tile_off, const int* d_r, const int* d_s,
                 const int* d_red_off, const int* d_skel_off, const int* d_red_gid,
                 const int* d_skel_gid, const unsigned char* d_skel_pass, const long long* d_off,
                 const T* d_M, int op, int pass, double alpha, const double* d_in,
                 double* d_out, int max_in, double flops, double bytes, hipStream_t s,
                 long long side_n, int rows_per_lane) {
    if (nbox <= 0 || ntiles <= 0) return;
    if (rows_per_lane != 1 && (rows_per_lane != 2 || op == 0 || op == 2))
        throw Error(-3, "apply: two rows per lane are for ops 1, 3 and 4 only");
    size_t shm = (size_t)std::max(max_in, 1) * sizeof(double);
    if (shm > 150 * 1024) throw Error(-3, "apply: box too large for LDS");
    KernelTimer* kt = active_timer();
    const bool f32 = sizeof(T) == sizeof(float);
    if (kt)
        kt->begin(s, rows_per_lane == 2 ? (f32 ? "k_apply_tiles<float, 2>" : "k_apply_tiles<double, 2>")
                                        : (f32 ? "k_apply_tiles<float, 1>" : "k_apply_tiles<double, 1>"));
    if (rows_per_lane == 2)
        k_apply_tiles<T, 2><<<ntiles, 256, shm, s>>>(nbox, d_tile_off, d_r, d_s, d_red_off,
                                                     d_skel_off, d_red_gid, d_skel_gid,
                                                     d_skel_pass, d_off, d_M, op, pass, alpha,
                                                     d_in, d_out, side_n);
    else
        k_apply_tiles<T, 1><<<ntiles, 256, shm, s>>>(nbox, d_tile_off, d_r, d_s, d_red_off,
                                                     d_skel_off, d_red_gid, d_skel_gid,
                                                     d_skel_pass, d_off, d_M, op, pass, alpha,
                                                     d_in, d_out, side_n);
    HIFDE_LAUNCH_CHECK();
    if (kt) kt->end(s, flops, bytes);
}

void side_add(int n, const int* d_gid, const unsigned char* d_sp, const double* d_side,
              long long side_n, double* d_u, hipStream_t s, bool paired) {
    if (n <= 0) return;
    if (paired) {
        k_side_add_pair<<<(n + 255) / 256, 256, 0, s>>>(n, d_gid, d_sp, d_side, side_n, d_u);
        HIFDE_LAUNCH_CHECK();
        return;
    }
    for (int p = 0; p < 2; ++p) {
        k_side_add<<<(n + 255) / 256, 256, 0, s>>>(n, d_gid, d_sp, d_side, side_n, p, d_u);
        HIFDE_LAUNCH_CHECK();
    }
}

void zero_boundary(double* d_u, int NC, hipStream_t s) {
    k_zero_boundary<<<grid_for((long long)NC * NC * NC), 256, 0, s>>>(d_u, NC);
    HIFDE_LAUNCH_CHECK();
}

void csr_spmv(int n, const int* d_rowptr, const int* d_colind, const double* d_val,
              const double* d_x, double* d_y, hipStream_t s) {
    k_csr_spmv<<<(n + 255) / 256, 256, 0, s>>>(n, d_rowptr, d_colind, d_val, d_x, d_y);
    HIFDE_LAUNCH_CHECK();
}

int csr_max_row(int n, const int* d_rowptr, hipStream_t s) {
    if (n <= 0) return 0;
    DevBuf<int> d;
    d.reserve(1);
    k_fill_zero_int<<<1, 1, 0, s>>>(d.p);
    HIFDE_LAUNCH_CHECK();
    k_max_row<<<std::min((n + 255) / 256, 1024), 256, 0, s>>>(n, d_rowptr, d.p);
    HIFDE_LAUNCH_CHECK();
    int h = 0;
    HIFDE_CHECK(hipMemcpyAsync(&h, d.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIFDE_CHECK(hipStreamSynchronize(s));
    return h;
}

void csr_to_ell(int n, int width, const int* d_rowptr, const int* d_colind, const double* d_val,
                double* d_ell_v, int* d_ell_c, hipStream_t s) {
    if (n <= 0 || width <= 0) return;
    k_csr_to_ell<<<(n + 255) / 256, 256, 0, s>>>(n, width, d_rowptr, d_colind, d_val, d_ell_v,
                                                 d_ell_c);
    HIFDE_LAUNCH_CHECK();
}

void ell_spmv(int n, int width, const int* d_rowptr, const double* d_ell_v, const int* d_ell_c,
              const double* d_x, double* d_y, hipStream_t s) {
    k_ell_spmv<<<(n + 255) / 256, 256, 0, s>>>(n, width, d_rowptr, d_ell_v, d_ell_c, d_x, d_y);
    HIFDE_LAUNCH_CHECK();
}

int dot_blocks(long long n) { return grid_for(n, kDotThreads, 1024); }

void dot_partial(const double* x, const double* y, long long n, double* d_part, int nblk,
                 hipStream_t s) {
    k_dot_partial<<<nblk, kDotThreads, 0, s>>>(x, y, n, d_part);
    HIFDE_LAUNCH_CHECK();
}

void reduce_partials(const double* d_part, int nblk, double* d_out, hipStream_t s) {
    k_reduce<<<1, kDotThreads, 0, s>>>(d_part, nblk, d_out);
    HIFDE_LAUNCH_CHECK();
}

void axpy(double a, const double* x, double* y, long long n, hipStream_t s) {
    k_axpy<<<grid_for(n), 256, 0, s>>>(a, x, y, n);
    HIFDE_LAUNCH_CHECK();
}

void xpby(const double* x, double b, double* y, long long n, hipStream_t s) {
    k_xpby<<<grid_for(n), 256, 0, s>>>(x, b, y, n);
    HIFDE_LAUNCH_CHECK();
}

void vcopy(const double* x, double* y, long long n, hipStream_t s) {
    k_copy<<<grid_for(n), 256, 0, s>>>(x, y, n);
    HIFDE_LAUNCH_CHECK();
}

template void apply_small<double>(int, int, const int*, const int*, const int*, const int*,
                                  const int*, const int*, const unsigned char*, const long long*,
                                  const double*, int, int, double, const double*, double*, double,
                                  double, hipStream_t, long long, int);
template void apply_small<float>(int, int, const int*, const int*, const int*, const int*,
                                 const int*, const int*, const unsigned char*, const long long*,
                                 const float*, int, int, double, const double*, double*, double,
                                 double, hipStream_t, long long, int);
template void apply_tiles<double>(int, int, const int*, const int*, const int*, const int*,
                                  const int*, const int*, const int*, const unsigned char*,
                                  const long long*, const double*, int, int, double,
                                  const double*, double*, int, double, double, hipStream_t,
                                  long long, int);
template void apply_tiles<float>(int, int, const int*, const int*, const int*, const int*,
                                 const int*, const int*, const int*, const unsigned char*,
                                 const long long*, const float*, int, int, double, const double*,
                                 double*, int, double, double, hipStream_t, long long, int);

}  // namespace hifde
