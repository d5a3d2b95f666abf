// Device level matrix: an open-addressing hash table of (row, col) -> value
// for the DOFs still alive.  It carries the semantics of the reference's
// Sparse<Scalar> (include/hifde3d/core/sparse.hpp): gathers read absent
// entries as 0 (Find, :273-304), scatter-adds sum into existing entries or
// create them (Add, :103-115), and deleted rows/columns (DeleteRow/DeleteCol,
// :173-206) are dropped by a compaction that keeps only live pairs.
//
// Determinism: every scatter pass has at most one contributor per key, so the
// per-key sum order is fixed by the pass order (see table_scatter).

#include "hifde_internal.hpp"

namespace hifde {

namespace {

__device__ __forceinline__ unsigned long long hmix(unsigned long long x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ unsigned long long make_key(int r, int c) {
    return ((unsigned long long)(unsigned)r << 32) | (unsigned)c;
}

// Symmetric storage keeps one entry per unordered pair under (min, max).
__device__ __forceinline__ unsigned long long make_key_s(int r, int c, bool sym) {
    return (sym && r > c) ? make_key(c, r) : make_key(r, c);
}

__device__ __forceinline__ double table_find(const Slot* __restrict__ t, unsigned long long mask,
                                             unsigned long long key) {
    unsigned long long h = hmix(key) & mask;
    while (true) {
        unsigned long long k = t[h].key;
        if (k == key) return t[h].val;
        if (k == kEmptyKey) return 0.0;
        h = (h + 1) & mask;
    }
}

__device__ __forceinline__ void table_add(Slot* t, unsigned long long mask, unsigned long long key,
                                          double v) {
    unsigned long long h = hmix(key) & mask;
    while (true) {
        unsigned long long k = t[h].key;
        if (k == kEmptyKey) {
            unsigned long long prev = atomicCAS(&t[h].key, kEmptyKey, key);
            if (prev == kEmptyKey || prev == key) {
                atomicAdd(&t[h].val, v);
                return;
            }
            k = prev;
        }
        if (k == key) {
            atomicAdd(&t[h].val, v);
            return;
        }
        h = (h + 1) & mask;
    }
}

// Largest b in [0, nbox) with off[b] <= e (off is nondecreasing, off[0] = 0).
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

__global__ void k_table_clear(Slot* t, size_t cap) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < cap; i += stride) {
        t[i].key = kEmptyKey;
        t[i].val = 0.0;
    }
}

__global__ void k_table_insert_coo(Slot* t, unsigned long long mask, const int* __restrict__ rows,
                                   const int* __restrict__ cols, const double* __restrict__ vals,
                                   long long nnz, bool sym) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < nnz; i += stride) {
        if (sym && rows[i] > cols[i]) continue;
        table_add(t, mask, make_key(rows[i], cols[i]), vals[i]);
    }
}

template <class T>
__global__ void k_table_gather(const Slot* __restrict__ t, unsigned long long mask, int nbox,
                               const int* __restrict__ rows, const int* __restrict__ row_off,
                               const int* __restrict__ cols, const int* __restrict__ col_off,
                               const long long* __restrict__ out_off, long long total,
                               T* __restrict__ out, bool sym) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; e < total; e += stride) {
        int b = find_box(out_off, nbox, e);
        long long l = e - out_off[b];
        int m = row_off[b + 1] - row_off[b];
        int i = (int)(l % m), j = (int)(l / m);
        int r = rows[row_off[b] + i], c = cols[col_off[b] + j];
        out[e] = (T)table_find(t, mask, make_key_s(r, c, sym));
    }
}

template <class T>
__global__ void k_table_scatter(Slot* t, unsigned long long mask, int nbox,
                                const int* __restrict__ s_dim, const int* __restrict__ skel_off,
                                const int* __restrict__ skel_gid,
                                const unsigned char* __restrict__ skel_pass,
                                const long long* __restrict__ off_ss, long long total,
                                const T* __restrict__ S, int pass, bool sym) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; e < total; e += stride) {
        int b = find_box(off_ss, nbox, e);
        long long l = e - off_ss[b];
        int s = s_dim[b];
        int i = (int)(l % s), j = (int)(l / s);
        int so = skel_off[b];
        if (sym && i < j) continue;  // one entry per unordered pair, from the lower triangle
        if (pass >= 0) {
            int p2 = (skel_pass[so + i] && skel_pass[so + j]) ? 1 : 0;
            if (p2 != pass) continue;
        }
        table_add(t, mask, make_key_s(skel_gid[so + i], skel_gid[so + j], sym), (double)S[e]);
    }
}

__global__ void k_table_count_live(const Slot* __restrict__ t, size_t cap,
                                   const unsigned char* __restrict__ alive,
                                   unsigned long long* count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    unsigned long long c = 0;
    for (; i < cap; i += stride) {
        unsigned long long k = t[i].key;
        if (k != kEmptyKey && (!alive || (alive[k >> 32] && alive[k & 0xffffffffu]))) ++c;
    }
    // wave reduction then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

__global__ void k_table_rehash(const Slot* __restrict__ t, size_t cap,
                               const unsigned char* __restrict__ alive, Slot* nt,
                               unsigned long long nmask) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < cap; i += stride) {
        unsigned long long k = t[i].key;
        if (k == kEmptyKey) continue;
        if (alive && !(alive[k >> 32] && alive[k & 0xffffffffu])) continue;
        unsigned long long h = hmix(k) & nmask;
        while (true) {
            unsigned long long prev = atomicCAS(&nt[h].key, kEmptyKey, k);
            if (prev == kEmptyKey) {
                nt[h].val = t[i].val;
                break;
            }
            h = (h + 1) & nmask;
        }
    }
}

int grid_for(long long n, int block = 256, int cap = 1 << 16) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

size_t pow2_at_least(size_t x) {
    size_t c = 1024;
    while (c < x) c <<= 1;
    return c;
}

}  // namespace

void table_init(LevelTable& t, size_t min_entries, hipStream_t s) {
    size_t cap = pow2_at_least(min_entries * 2 + 16);
    t.slots.reserve(cap);  // grow-only: the buffer is kept across factorizations
    t.cap = cap;
    k_table_clear<<<grid_for((long long)cap), 256, 0, s>>>(t.slots.p, cap);
    HIFDE_LAUNCH_CHECK();
}

void table_insert_coo(LevelTable& t, const int* d_rows, const int* d_cols, const double* d_vals,
                      long long nnz, hipStream_t s) {
    const bool sym = t.sym;
    if (nnz <= 0) return;
    k_table_insert_coo<<<grid_for(nnz), 256, 0, s>>>(t.slots.p, t.cap - 1, d_rows, d_cols, d_vals,
                                                     nnz, sym);
    HIFDE_LAUNCH_CHECK();
}

template <class T>
void table_gather(const LevelTable& t, int nbox, const int* d_rows, const int* d_row_off,
                  const int* d_cols, const int* d_col_off, const long long* d_out_off,
                  long long total, T* d_out, hipStream_t s) {
    if (total <= 0 || nbox <= 0) return;
    k_table_gather<T><<<grid_for(total, 256, 1 << 20), 256, 0, s>>>(
        t.slots.p, t.cap - 1, nbox, d_rows, d_row_off, d_cols, d_col_off, d_out_off, total, d_out,
        t.sym);
    HIFDE_LAUNCH_CHECK();
}

template <class T>
void table_scatter(LevelTable& t, int nbox, const int* d_s, const int* d_skel_off,
                   const int* d_skel_gid, const unsigned char* d_skel_pass,
                   const long long* d_off_ss, long long total, const T* d_S, int pass,
                   hipStream_t s) {
    if (total <= 0 || nbox <= 0) return;
    k_table_scatter<T><<<grid_for(total, 256, 1 << 20), 256, 0, s>>>(
        t.slots.p, t.cap - 1, nbox, d_s, d_skel_off, d_skel_gid, d_skel_pass, d_off_ss, total, d_S,
        pass, t.sym);
    HIFDE_LAUNCH_CHECK();
}

size_t table_count(const LevelTable& t, hipStream_t s) {
    DevBuf<unsigned long long> cnt;
    cnt.reserve(1);
    HIFDE_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), s));
    k_table_count_live<<<grid_for((long long)t.cap), 256, 0, s>>>(t.slots.p, t.cap, nullptr,
                                                                   cnt.p);
    HIFDE_LAUNCH_CHECK();
    unsigned long long h = 0;
    HIFDE_CHECK(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
    HIFDE_CHECK(hipStreamSynchronize(s));
    return (size_t)h;
}

size_t table_compact(const LevelTable& src, LevelTable& dst, const unsigned char* d_alive,
                     size_t extra, DevBuf<unsigned long long>& counter, hipStream_t s) {
    counter.reserve(1);
    HIFDE_CHECK(hipMemsetAsync(counter.p, 0, sizeof(unsigned long long), s));
    k_table_count_live<<<grid_for((long long)src.cap), 256, 0, s>>>(src.slots.p, src.cap, d_alive,
                                                                     counter.p);
    HIFDE_LAUNCH_CHECK();
    unsigned long long live = 0;
    HIFDE_CHECK(hipMemcpyAsync(&live, counter.p, sizeof(live), hipMemcpyDeviceToHost, s));
    HIFDE_CHECK(hipStreamSynchronize(s));
    dst.sym = src.sym;
    table_init(dst, (size_t)(live + extra), s);
    k_table_rehash<<<grid_for((long long)src.cap), 256, 0, s>>>(src.slots.p, src.cap, d_alive,
                                                                 dst.slots.p, dst.cap - 1);
    HIFDE_LAUNCH_CHECK();
    return (size_t)live;
}

template void table_gather<double>(const LevelTable&, int, const int*, const int*, const int*,
                                   const int*, const long long*, long long, double*, hipStream_t);
template void table_gather<float>(const LevelTable&, int, const int*, const int*, const int*,
                                  const int*, const long long*, long long, float*, hipStream_t);
template void table_scatter<double>(LevelTable&, int, const int*, const int*, const int*,
                                    const unsigned char*, const long long*, long long,
                                    const double*, int, hipStream_t);
template void table_scatter<float>(LevelTable&, int, const int*, const int*, const int*,
                                   const unsigned char*, const long long*, long long, const float*,
                                   int, hipStream_t);

}  // namespace hifde
