// Element storage of the level matrix.
//
// The level matrix of HIFFactor (sp_matrix_, updated by UpdateMatrixAndDOFs,
// src/Factor.cpp:262-311) is held unassembled: A = sum over the cells c of
// the current level of an element block E_c on the live DOFs of the cell's
// closure that are not interior to it ([face-interior | edges and corners]).
// This is exact for operators whose couplings stay inside level-0 cell
// closures (every stencil of include/setup_stencil.hpp and its variable
// coefficient variant): cell-interior DOFs only couple inside their own cell,
// Schur complements only add couplings among one cell's face DOFs, and the
// skeletonization of a face only adds couplings among its own skeleton.  So
//   * a cell's Schur blocks A22, A21, A12 come from one frontal block;
//   * S = -A12 A22^{-1} A21 is added to the cell's own element, by the GEMM
//     that computes it (beta = 1), instead of a scatter into a global table;
//   * the ID matrix A(N, F) of a face is the sum of the two adjacent cells'
//     elements, and its Schur complement goes into one of them;
//   * the next level's frontal block of a cell is the extend-add of its
//     eight children's elements.
// Every entry is written by exactly one thread per pass, so the summation
// order is fixed and the result deterministic.

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

inline unsigned elem_grid(long long n, int per = 256) {
    long long g = (n + per - 1) / per;
    if (g < 1) g = 1;
    if (g > (1LL << 20)) g = 1LL << 20;
    return (unsigned)g;
}

template <class T>
__global__ void k_fill(T* p, long long n) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x)
        p[e] = (T)0;
}

// Level-0 frontal blocks from the assembled operator (CSR).  One workgroup
// per cell.  An entry A(a, b) is given to the lowest-numbered level-0 cell
// whose closure holds both DOFs (per axis the candidates of a coordinate x
// are x/w - 1 and x/w when x lies on a cell plane, else x/w).
__device__ __forceinline__ void axis_cells(int x, int w, int c, int& lo, int& hi) {
    if (x % w == 0) {
        lo = max(x / w - 1, 0);
        hi = min(x / w, c - 1);
    } else {
        lo = hi = x / w;
    }
}

template <class T, class TE>
__global__ void __launch_bounds__(256) k_frontal_init(const InitDesc* __restrict__ desc,
                                                      const int* __restrict__ tab,
                                                      const int* __restrict__ gids,
                                                      const int* __restrict__ rp,
                                                      const int* __restrict__ ci,
                                                      const double* __restrict__ v, int NC,
                                                      int w, int c, T* A22, T* A21, T* A12,
                                                      TE* E) {
    const InitDesc d = desc[blockIdx.x];
    const int r = d.r, s = d.s, e = d.s + d.o, t = d.r + d.s + d.o;
    for (long long q = threadIdx.x; q < (long long)r * r; q += blockDim.x) A22[d.rr + q] = (T)0;
    for (long long q = threadIdx.x; q < (long long)r * s; q += blockDim.x) {
        A21[d.rs + q] = (T)0;
        if (A12) A12[d.rs + q] = (T)0;
    }
    for (long long q = threadIdx.x; q < (long long)e * e; q += blockDim.x) E[d.ee + q] = (TE)0;
    __syncthreads();
    const int NC2 = NC * NC;
    for (int pa = threadIdx.x; pa < t; pa += blockDim.x) {
        const int a = gids[d.gid + pa];
        const int ax = a % NC, ay = (a / NC) % NC, az = a / NC2;
        int alo[3], ahi[3];
        axis_cells(ax, w, c, alo[0], ahi[0]);
        axis_cells(ay, w, c, alo[1], ahi[1]);
        axis_cells(az, w, c, alo[2], ahi[2]);
        for (int q = rp[a]; q < rp[a + 1]; ++q) {
            const int b = ci[q];
            const int bx = b % NC, by = (b / NC) % NC, bz = b / NC2;
            const int lx = bx - d.lo[0], ly = by - d.lo[1], lz = bz - d.lo[2];
            if (lx < 0 || ly < 0 || lz < 0 || lx >= d.ext[0] || ly >= d.ext[1] || lz >= d.ext[2])
                continue;
            const int pb = tab[d.tab + lx + d.ext[0] * (ly + d.ext[1] * lz)];
            if (pb < 0) continue;
            // owner: per axis the lowest cell holding both coordinates
            int blo, bhi, own = 1;
            const int bc[3] = {bx, by, bz};
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                axis_cells(bc[x], w, c, blo, bhi);
                const int lo = max(alo[x], blo);
                own &= (lo == d.cc[x]);
            }
            if (!own) continue;
            const T val = (T)v[q];
            if (pa < r) {
                if (pb < r) A22[d.rr + pa + (long long)pb * r] = val;
                else if (pb < r + s) A21[d.rs + pa + (long long)(pb - r) * r] = val;
            } else if (pb >= r) {
                E[d.ee + (pa - r) + (long long)(pb - r) * e] = (TE)v[q];
            } else if (pa < r + s && A12) {
                A12[d.rs + (pa - r) + (long long)pb * s] = val;
            }
        }
    }
}

// One pass of the extend-add: for every parent, the element of one child is
// added into the parent's frontal blocks.  Workgroup = one child column.
template <class Td, class TE>
__global__ void __launch_bounds__(128) k_extend_add(int npar, const long long* __restrict__ coff,
                                                    const ExtendDesc* __restrict__ desc,
                                                    const int* __restrict__ maps,
                                                    const TE* __restrict__ src, Td* A22,
                                                    Td* A21, Td* A12, TE* E) {
    const long long blk = blockIdx.x;
    const int p = find_slot(coff, npar, blk);
    const ExtendDesc d = desc[p];
    const int j = (int)(blk - coff[p]);
    const int* mp = maps + d.map;
    const int pj = mp[j];
    if (pj < 0) return;
    const int r = d.r, s = d.s;
    const TE* col = src + d.src + (long long)j * d.e;
    if (pj < r) {
        for (int i = threadIdx.x; i < d.e; i += blockDim.x) {
            const int pi = mp[i];
            if (pi < 0) continue;
            const Td x = (Td)col[i];
            if (pi < r) A22[d.rr + pi + (long long)pj * r] += x;
            else if (pi < r + s && A12) A12[d.rs + (pi - r) + (long long)pj * s] += x;
        }
    } else if (pj < r + s) {
        const int sj = pj - r;
        for (int i = threadIdx.x; i < d.e; i += blockDim.x) {
            const int pi = mp[i];
            if (pi < 0) continue;
            if (pi < r) A21[d.rs + pi + (long long)sj * r] += (Td)col[i];
            else E[d.ee + (pi - r) + (long long)sj * d.ep] += col[i];
        }
    } else {
        const int ej = pj - r;
        for (int i = threadIdx.x; i < d.e; i += blockDim.x) {
            const int pi = mp[i];
            if (pi < r) continue;  // dead, or red x other (identically zero)
            E[d.ee + (pi - r) + (long long)ej * d.ep] += col[i];
        }
    }
}

// M(i, j) = A(rows_i, cols_j) and C(i, j) = A(cols_i, cols_j) of every face,
// in FP64, summed from the two adjacent cells' elements (FP64 or FP32).
// Workgroup = one face column.
// Positions are packed two to a word: the low half is the position in the
// first cell's element, the high half in the second's, kNoPos when absent.
template <class TE>
__global__ void __launch_bounds__(128) k_face_gather(int nface, const long long* __restrict__ coff,
                                                     const FaceGatherDesc* __restrict__ desc,
                                                     const unsigned* __restrict__ rpos,
                                                     const unsigned* __restrict__ cpos,
                                                     const TE* __restrict__ E, double* M,
                                                     double* C) {
    const long long blk = blockIdx.x;
    const int f = find_slot(coff, nface, blk);
    const FaceGatherDesc d = desc[f];
    const int j = (int)(blk - coff[f]);
    const unsigned* rp = rpos + d.rows;
    const unsigned* cp = cpos + d.cols;
    const unsigned cj = cp[j];
    const TE* e1 = E + d.e1 + (long long)(cj & 0xFFFFu) * d.ld1;
    const TE* e2 = E + d.e2 + (long long)(cj >> 16) * d.ld2;
    for (int i = threadIdx.x; i < d.m; i += blockDim.x) {
        const unsigned x = rp[i];
        const unsigned p1 = x & 0xFFFFu, p2 = x >> 16;
        double v = 0.0;
        if (p1 != kNoPos) v += (double)e1[p1];
        if (p2 != kNoPos) v += (double)e2[p2];
        M[d.offM + i + (long long)j * d.m] = v;
    }
    for (int i = threadIdx.x; i < d.n; i += blockDim.x) {
        const unsigned x = cp[i];
        C[d.offC + i + (long long)j * d.n] = (double)e1[x & 0xFFFFu] + (double)e2[x >> 16];
    }
}

// E_owner(skel_f, skel_f) += S_f.  Workgroup = one face column.  With sym,
// only the lower triangle of S_f is valid and is read for both halves.
template <class T, class TE>
__global__ void __launch_bounds__(128) k_face_add(int nface, const long long* __restrict__ coff,
                                                  const FaceAddDesc* __restrict__ desc,
                                                  const int* __restrict__ spos,
                                                  const T* __restrict__ S, TE* E, int sym) {
    const long long blk = blockIdx.x;
    const int f = find_slot(coff, nface, blk);
    const FaceAddDesc d = desc[f];
    const int j = (int)(blk - coff[f]);
    const int* sp = spos + d.sp;
    TE* col = E + d.e1 + (long long)sp[j] * d.ld1;
    const T* Sf = S + d.offS;
    for (int i = threadIdx.x; i < d.k; i += blockDim.x) {
        const int a = (sym && i < j) ? j : i, b = (sym && i < j) ? i : j;
        col[sp[i]] += (TE)Sf[a + (long long)b * d.k];
    }
}

// E(j, i) = E(i, j) for i > j inside the leading s x s block of each element
// (the symmetric Schur product fills the lower triangle).  Workgroup = one
// 64 x 64 tile (I, J), I >= J, of the lower triangle: it is read along
// columns, transposed through LDS and written along the columns of tile
// (J, I), so both sides are coalesced.
template <class TE>
__global__ void __launch_bounds__(256) k_mirror_lower(int nbox, const long long* __restrict__ toff,
                                                      const MirrorDesc* __restrict__ desc,
                                                      TE* E) {
    const long long blk = blockIdx.x;
    const int b = find_slot(toff, nbox, blk);
    const MirrorDesc d = desc[b];
    // tile t -> (I, J) with I >= J, column-major over the lower triangle
    int t = (int)(blk - toff[b]);
    const int nt = (d.s + 63) / 64;
    int J = 0;
    while (t >= nt - J) {
        t -= nt - J;
        ++J;
    }
    const int I = J + t;
    __shared__ TE tile[64][65];
    TE* base = E + d.ee;
    const int i0 = I * 64, j0 = J * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int c = ty; c < 64; c += 4) {
        const int gi = i0 + tx, gj = j0 + c;
        tile[c][tx] = (gi < d.s && gj < d.s) ? base[gi + (long long)gj * d.ld] : (TE)0;
    }
    __syncthreads();
    // write E(j0 + tx, i0 + c) = E(i0 + c, j0 + tx) where i0 + c > j0 + tx
    for (int c = ty; c < 64; c += 4) {
        const int gi = i0 + c, gj = j0 + tx;
        if (gi < d.s && gj < d.s && gi > gj) base[gj + (long long)gi * d.ld] = tile[tx][c];
    }
}

template <class D>
std::vector<long long> col_prefix(const std::vector<D>& desc, int (*cols)(const D&)) {
    std::vector<long long> off(desc.size() + 1, 0);
    for (size_t i = 0; i < desc.size(); ++i) off[i + 1] = off[i] + cols(desc[i]);
    return off;
}

}  // namespace

template <class T>
void fill_zero(T* p, long long n, hipStream_t s) {
    if (n <= 0) return;
    k_fill<T><<<elem_grid(n), 256, 0, s>>>(p, n);
    HIFDE_LAUNCH_CHECK();
}

template <class T, class TE>
void frontal_init(int ncell, const InitDesc* d_desc, const int* d_tab, const int* d_gids,
                  const int* d_rp, const int* d_ci, const double* d_v, int NC, int w, int c,
                  T* A22, T* A21, T* A12, TE* E, hipStream_t s) {
    if (ncell <= 0) return;
    k_frontal_init<T, TE><<<(unsigned)ncell, 256, 0, s>>>(d_desc, d_tab, d_gids, d_rp, d_ci, d_v, NC,
                                                      w, c, A22, A21, A12, E);
    HIFDE_LAUNCH_CHECK();
}

template <class Td, class TE>
void extend_add(const std::vector<std::vector<ExtendDesc>>& passes, DevBuf<ExtendDesc>& dbuf,
                DevBuf<long long>& obuf, const int* d_maps, const TE* src, Td* A22, Td* A21,
                Td* A12, TE* E, hipStream_t s) {
    // all passes' descriptors and column prefixes in one upload each
    std::vector<ExtendDesc> all;
    std::vector<long long> off;
    std::vector<size_t> d0, o0;
    std::vector<long long> tot;
    for (const std::vector<ExtendDesc>& p : passes) {
        d0.push_back(all.size());
        o0.push_back(off.size());
        all.insert(all.end(), p.begin(), p.end());
        const std::vector<long long> po =
            col_prefix<ExtendDesc>(p, [](const ExtendDesc& d) { return d.e; });
        off.insert(off.end(), po.begin(), po.end());
        tot.push_back(po.back());
    }
    if (all.empty()) return;
    dbuf.upload(all, s);
    obuf.upload(off, s);
    for (size_t k = 0; k < passes.size(); ++k) {
        if (passes[k].empty() || tot[k] == 0) continue;
        k_extend_add<Td, TE><<<(unsigned)tot[k], 128, 0, s>>>((int)passes[k].size(), obuf.p + o0[k],
                                                          dbuf.p + d0[k], d_maps, src, A22, A21,
                                                          A12, E);
        HIFDE_LAUNCH_CHECK();
    }
}

long long face_gather_prepare(const std::vector<FaceGatherDesc>& desc,
                              DevBuf<FaceGatherDesc>& dbuf, DevBuf<long long>& obuf,
                              hipStream_t s) {
    if (desc.empty()) return 0;
    const std::vector<long long> off =
        col_prefix<FaceGatherDesc>(desc, [](const FaceGatherDesc& d) { return d.n; });
    dbuf.upload(desc, s);
    obuf.upload(off, s);
    return off.back();
}

template <class TE>
void face_gather(int nface, long long total_cols, const FaceGatherDesc* d_desc,
                 const long long* d_off, const unsigned* d_rpos, const unsigned* d_cpos,
                 const TE* E, double* M, double* C, hipStream_t s) {
    if (nface <= 0 || total_cols <= 0) return;
    k_face_gather<TE><<<(unsigned)total_cols, 128, 0, s>>>(nface, d_off, d_desc, d_rpos, d_cpos, E,
                                                       M, C);
    HIFDE_LAUNCH_CHECK();
}

template <class T, class TE>
void face_add(const std::vector<FaceAddDesc>& desc, DevBuf<FaceAddDesc>& dbuf,
              DevBuf<long long>& obuf, const int* d_spos, const T* S, TE* E, int sym,
              hipStream_t s) {
    if (desc.empty()) return;
    const std::vector<long long> off =
        col_prefix<FaceAddDesc>(desc, [](const FaceAddDesc& d) { return d.k; });
    if (off.back() == 0) return;
    dbuf.upload(desc, s);
    obuf.upload(off, s);
    k_face_add<T, TE><<<(unsigned)off.back(), 128, 0, s>>>((int)desc.size(), obuf.p, dbuf.p, d_spos,
                                                       S, E, sym);
    HIFDE_LAUNCH_CHECK();
}

template <class TE>
void mirror_lower(const std::vector<MirrorDesc>& desc, DevBuf<MirrorDesc>& dbuf,
                  DevBuf<long long>& obuf, TE* E, hipStream_t s) {
    if (desc.empty()) return;
    const std::vector<long long> off = col_prefix<MirrorDesc>(desc, [](const MirrorDesc& d) {
        const int nt = (d.s + 63) / 64;
        return nt * (nt + 1) / 2;
    });
    if (off.back() == 0) return;
    dbuf.upload(desc, s);
    obuf.upload(off, s);
    k_mirror_lower<TE><<<(unsigned)off.back(), 256, 0, s>>>((int)desc.size(), obuf.p, dbuf.p, E);
    HIFDE_LAUNCH_CHECK();
}

#define HIFDE_INST1(T, TE)                                                                      \
    template void frontal_init<T, TE>(int, const InitDesc*, const int*, const int*, const int*, \
                                      const int*, const double*, int, int, int, T*, T*, T*,     \
                                      TE*, hipStream_t);                                        \
    template void extend_add<T, TE>(const std::vector<std::vector<ExtendDesc>>&,                \
                                    DevBuf<ExtendDesc>&, DevBuf<long long>&, const int*,        \
                                    const TE*, T*, T*, T*, TE*, hipStream_t);                   \
    template void face_add<T, TE>(const std::vector<FaceAddDesc>&, DevBuf<FaceAddDesc>&,        \
                                  DevBuf<long long>&, const int*, const T*, TE*, int,           \
                                  hipStream_t);
// elements are FP64, or FP32 when every level of the factorization is FP32
HIFDE_INST1(float, double)
HIFDE_INST1(double, double)
HIFDE_INST1(float, float)
#undef HIFDE_INST1
#define HIFDE_INST2(TE)                                                                         \
    template void fill_zero<TE>(TE*, long long, hipStream_t);                                   \
    template void face_gather<TE>(int, long long, const FaceGatherDesc*, const long long*,      \
                                  const unsigned*, const unsigned*, const TE*, double*,         \
                                  double*, hipStream_t);                                        \
    template void mirror_lower<TE>(const std::vector<MirrorDesc>&, DevBuf<MirrorDesc>&,         \
                                   DevBuf<long long>&, TE*, hipStream_t);
HIFDE_INST2(float)
HIFDE_INST2(double)
#undef HIFDE_INST2

}  // namespace hifde
