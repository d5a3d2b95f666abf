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

// One 64-output tile of one box per workgroup.
// op 0: u_s[j] += alpha * sum_i M(i,j) u_r[i]   M: r x s   (X^H u_red)
// op 1: u_r[i] += alpha * sum_j M(i,j) u_s[j]   M: r x s   (X u_skel)
// op 2: u_r[j] += alpha * sum_i M(i,j) u_s[i]   M: s x r   (W^H u_skel)
// op 3: u_s[i] += alpha * sum_j M(i,j) u_r[j]   M: s x r   (W u_red)
// op 4: out[r_i] = sum_j M(i,j) u_r[j]          M: r x r   (A22 or A22^{-1})
// Ops 0-3 read and write disjoint DOF classes, so they update u in place;
// op 4 reads u_in and writes u_out (a different vector).  pass >= 0 restricts
// op 0/3 writes to skeleton DOFs whose pass flag equals pass.
template <class T>
__global__ void __launch_bounds__(256) k_apply_tiles(
    int nbox, const int* __restrict__ tile_off, const int* __restrict__ dr,
    const int* __restrict__ ds, const int* __restrict__ red_off,
    const int* __restrict__ skel_off, const int* __restrict__ red_gid,
    const int* __restrict__ skel_gid, const unsigned char* __restrict__ skel_pass,
    const long long* __restrict__ off, const T* __restrict__ Mb, int op, int pass, double alpha,
    const double* u_in, double* u_out) {
    const int t = blockIdx.x;
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
    __shared__ double part[4][64];
    const int tid = threadIdx.x;
    for (int q = tid; q < nin; q += 256) xs[q] = u_in[ig[q]];
    __syncthreads();
    const int o0 = tile * 64;
    const int cnt = min(64, nout - o0);
    double acc = 0.0;
    int o = -1;
    if (!trans) {
        // y_i = sum_j M(i,j) x_j ; M has nout rows.  Four column groups.
        const int i = tid & 63, g = tid >> 6;
        double pacc = 0.0;
        if (i < cnt) {
            // eight independent sums keep eight loads per thread in flight
            const T* Mi = M + o0 + i;
            double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            int j = g;
            for (; j + 28 < nin; j += 32) {
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    a[u] += (double)Mi[(size_t)(j + 4 * u) * nout] * xs[j + 4 * u];
            }
            for (; j < nin; j += 4) a[0] += (double)Mi[(size_t)j * nout] * xs[j];
            pacc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
        }
        part[g][i] = pacc;
        __syncthreads();
        if (tid < cnt) {
            o = o0 + tid;
            acc = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        }
    } else if (nin <= 64) {
        // y_j = sum_i M(i,j) x_i with short columns: one thread per output
        // (outputs of the other pass are not formed)
        if (tid < cnt && !(sp && pass >= 0 && (int)sp[o0 + tid] != pass)) {
            o = o0 + tid;
            const T* Mj = M + (size_t)o * nin;
            for (int i = 0; i < nin; ++i) acc += (double)Mj[i] * xs[i];
        }
    } else {
        // long columns: one wave per output, lanes over the column
        const int lane = tid & (kWave - 1), w = tid / kWave;
        for (int jj = w; jj < cnt; jj += 4) {
            if (sp && pass >= 0 && (int)sp[o0 + jj] != pass) continue;
            const T* Mj = M + (size_t)(o0 + jj) * nin;
            double p4[4] = {0, 0, 0, 0};
            int i = lane;
            for (; i + 3 * kWave < nin; i += 4 * kWave) {
#pragma unroll
                for (int u = 0; u < 4; ++u) p4[u] += (double)Mj[i + u * kWave] * xs[i + u * kWave];
            }
            for (; i < nin; i += kWave) p4[0] += (double)Mj[i] * xs[i];
            double a = wave_sum((p4[0] + p4[1]) + (p4[2] + p4[3]));
            if (lane == 0) part[0][jj] = a;
        }
        __syncthreads();
        if (tid < cnt) {
            o = o0 + tid;
            acc = part[0][tid];
        }
    }
    if (o < 0) return;
    if (op == 4) {
        u_out[og[o]] = acc;
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
// one thread per output over the flattened outputs of the batch, same ops and
// pass rule as k_apply_tiles, inputs read from the vector directly.
template <class T>
__global__ void __launch_bounds__(256) k_apply_small(
    int nbox, int nout_total, const int* __restrict__ dr, const int* __restrict__ ds,
    const int* __restrict__ red_off, const int* __restrict__ skel_off,
    const int* __restrict__ red_gid, const int* __restrict__ skel_gid,
    const unsigned char* __restrict__ skel_pass, const long long* __restrict__ off,
    const T* __restrict__ Mb, int op, int pass, double alpha, const double* u_in,
    double* u_out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nout_total) return;
    const bool out_red = (op == 1 || op == 2 || op == 4);
    const int* oof = out_red ? red_off : skel_off;
    const int b = find_box(oof, nbox, t);
    const int o = t - oof[b];
    const int r = dr[b], s = ds[b];
    if (r == 0 || (op != 4 && s == 0)) return;
    if (op == 0 && skel_pass && pass >= 0 && (int)skel_pass[skel_off[b] + o] != pass) return;
    const T* M = Mb + off[b];
    const int* rg = red_gid + red_off[b];
    const int* sg = skel_gid + skel_off[b];
    double acc = 0.0;
    switch (op) {
    case 0:  // M r x s, column o, inputs red
        for (int i = 0; i < r; ++i) acc += (double)M[i + (size_t)o * r] * u_in[rg[i]];
        break;
    case 1:  // M r x s, row o, inputs skel
        for (int j = 0; j < s; ++j) acc += (double)M[o + (size_t)j * r] * u_in[sg[j]];
        break;
    case 2:  // M s x r, column o, inputs skel
        for (int i = 0; i < s; ++i) acc += (double)M[i + (size_t)o * s] * u_in[sg[i]];
        break;
    case 3:  // M s x r, row o, inputs red
        for (int j = 0; j < r; ++j) acc += (double)M[o + (size_t)j * s] * u_in[rg[j]];
        break;
    default:  // 4: M r x r, row o
        for (int j = 0; j < r; ++j) acc += (double)M[o + (size_t)j * r] * u_in[rg[j]];
        break;
    }
    const int g = out_red ? rg[o] : sg[o];
    if (op == 4) u_out[g] = acc;
    else u_out[g] += alpha * acc;
}

template <class T>
void apply_small(int nbox, int nout_total, const int* d_r, const int* d_s, const int* d_red_off,
                 const int* d_skel_off, const int* d_red_gid, const int* d_skel_gid,
                 const unsigned char* d_skel_pass, const long long* d_off, const T* d_M, int op,
                 int pass, double alpha, const double* d_in, double* d_out, double flops,
                 double bytes, hipStream_t s) {
    if (nbox <= 0 || nout_total <= 0) return;
    KernelTimer* kt = active_timer();
    if (kt) kt->begin(s);
    k_apply_small<T><<<(nout_total + 255) / 256, 256, 0, s>>>(
        nbox, nout_total, d_r, d_s, d_red_off, d_skel_off, d_red_gid, d_skel_gid, d_skel_pass,
        d_off, d_M, op, pass, alpha, d_in, d_out);
    HIFDE_LAUNCH_CHECK();
    if (kt)
        kt->end(s, sizeof(T) == sizeof(float) ? "k_apply_small<float>" : "k_apply_small<double>",
                flops, bytes);
}

template <class T>
void apply_tiles(int nbox, int ntiles, const int* d_tile_off, const int* d_r, const int* d_s,
                 const int* d_red_off, const int* d_skel_off, const int* d_red_gid,
                 const int* d_skel_gid, const unsigned char* d_skel_pass, const long long* d_off,
                 const T* d_M, int op, int pass, double alpha, const double* d_in,
                 double* d_out, int max_in, double flops, double bytes, hipStream_t s) {
    if (nbox <= 0 || ntiles <= 0) return;
    size_t shm = (size_t)std::max(max_in, 1) * sizeof(double);
    if (shm > 150 * 1024) throw Error(-3, "apply: box too large for LDS");
    KernelTimer* kt = active_timer();
    if (kt) kt->begin(s);
    k_apply_tiles<T><<<ntiles, 256, shm, s>>>(nbox, d_tile_off, d_r, d_s, d_red_off, d_skel_off,
                                              d_red_gid, d_skel_gid, d_skel_pass, d_off, d_M, op,
                                              pass, alpha, d_in, d_out);
    HIFDE_LAUNCH_CHECK();
    if (kt)
        kt->end(s, sizeof(T) == sizeof(float) ? "k_apply_tiles<float>" : "k_apply_tiles<double>",
                flops, bytes);
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
                                  double, hipStream_t);
template void apply_small<float>(int, int, const int*, const int*, const int*, const int*,
                                 const int*, const int*, const unsigned char*, const long long*,
                                 const float*, int, int, double, const double*, double*, double,
                                 double, hipStream_t);
template void apply_tiles<double>(int, int, const int*, const int*, const int*, const int*,
                                  const int*, const int*, const int*, const unsigned char*,
                                  const long long*, const double*, int, int, double,
                                  const double*, double*, int, double, double, hipStream_t);
template void apply_tiles<float>(int, int, const int*, const int*, const int*, const int*,
                                 const int*, const int*, const int*, const unsigned char*,
                                 const long long*, const float*, int, int, double, const double*,
                                 double*, int, double, double, hipStream_t);

}  // namespace hifde
