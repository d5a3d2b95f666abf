// HIFDE3D CPU restatement -- TEST INFRASTRUCTURE ONLY.
//
// This file is the oracle: a plain, single-box-at-a-time FP64 restatement of
// the reference hierarchical interpolative factorization (HIF-DE) that the
// GPU library is checked against.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it, and only as the checker (or as the
// timed CPU baseline).  The product library never links or calls it.
//
// Every routine cites the reference file:line it restates (paths relative to
// the reference repository root).  Deviations from the reference, all of
// which change rounding only or complexity only, are:
//   * the level matrix is an O(nnz) open-addressing hash map instead of
//     std::map<col, std::map<row,val>> (include/hifde3d/core/sparse.hpp:21);
//     DeleteRow/DeleteCol (sparse.hpp:173-206) become an "alive" mask plus a
//     compaction after each half level.  Lookups only ever touch live DOFs,
//     so the values seen by every gather are identical.
//   * dense kernels are this file's own: unblocked column-pivoted Householder
//     QR with LAPACK's partial-norm downdating (the dgeqp3/dlaqp2 method),
//     right-looking LU with partial pivoting plus the getri inverse method,
//     triple-loop GEMM and an upper-triangular solve.  The reference calls an
//     unpinned external BLAS/LAPACK (CMakeLists.txt:38-75).
//
// Parity pins for this file live in tests/test_oracle.py: QRCP pivots and R
// against SciPy's LAPACK dgeqp3, LU inverse against numpy, the Schur closed
// form, the ID reconstruction bound, eps -> 0 exactness (F == A), brute-force
// dense solves at tiny grids, and the exact integer level sizes.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Dense kernels (column-major, leading dimension lda)
// ---------------------------------------------------------------------------

// First index of max |x| (BLAS idamax semantics, 0-based).
int idamax(int n, const double* x, int inc) {
    int best = 0;
    double bv = -1.0;
    for (int i = 0; i < n; ++i) {
        double v = std::fabs(x[i * inc]);
        if (v > bv) { bv = v; best = i; }
    }
    return best;
}

double nrm2(int n, const double* x) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i] * x[i];
    return std::sqrt(s);
}

// Householder generator (LAPACK dlarfg method).  On exit x holds v(2:n),
// *alpha holds beta, returns tau.
double larfg(int n, double* alpha, double* x) {
    if (n <= 1) return 0.0;
    double xnorm = nrm2(n - 1, x);
    if (xnorm == 0.0) return 0.0;
    double beta = -std::copysign(std::hypot(*alpha, xnorm), *alpha);
    double tau = (beta - *alpha) / beta;
    double scal = 1.0 / (*alpha - beta);
    for (int i = 0; i < n - 1; ++i) x[i] *= scal;
    *alpha = beta;
    return tau;
}

// Column-pivoted Householder QR with all columns free.  Restates the
// reference's PivotedQRWrapper -> dgeqp3 call (src/InterpDecomp.cpp:156-166,
// 198-211; include/hifde3d/core/lapack.hpp:569-591).  jpvt is returned
// 0-based (src/InterpDecomp.cpp:164-165).
void qrcp(int m, int n, double* A, int lda, int* jpvt, double* tau) {
    std::vector<double> vn1(n), vn2(n), w(n);
    for (int j = 0; j < n; ++j) {
        jpvt[j] = j;
        vn1[j] = vn2[j] = nrm2(m, A + (size_t)j * lda);
    }
    const double tol3z = std::sqrt(std::ldexp(1.0, -53));
    int mn = std::min(m, n);
    for (int i = 0; i < mn; ++i) {
        int pvt = i + idamax(n - i, vn1.data() + i, 1);
        if (pvt != i) {
            for (int r = 0; r < m; ++r)
                std::swap(A[r + (size_t)pvt * lda], A[r + (size_t)i * lda]);
            std::swap(jpvt[pvt], jpvt[i]);
            vn1[pvt] = vn1[i];
            vn2[pvt] = vn2[i];
        }
        double* col = A + (size_t)i * lda;
        double t;
        if (i < m - 1) t = larfg(m - i, col + i, col + i + 1);
        else t = 0.0;
        tau[i] = t;
        if (i < n - 1 && t != 0.0) {
            double aii = col[i];
            col[i] = 1.0;
            for (int j = i + 1; j < n; ++j) {
                double* cj = A + (size_t)j * lda;
                double s = 0.0;
                for (int r = i; r < m; ++r) s += cj[r] * col[r];
                w[j] = s;
            }
            for (int j = i + 1; j < n; ++j) {
                double* cj = A + (size_t)j * lda;
                double f = t * w[j];
                for (int r = i; r < m; ++r) cj[r] -= f * col[r];
            }
            col[i] = aii;
        }
        for (int j = i + 1; j < n; ++j) {
            if (vn1[j] != 0.0) {
                double* cj = A + (size_t)j * lda;
                double temp = std::fabs(cj[i]) / vn1[j];
                temp = 1.0 - temp * temp;
                temp = std::max(temp, 0.0);
                double q = vn1[j] / vn2[j];
                double temp2 = temp * q * q;
                if (temp2 <= tol3z) {
                    if (i < m - 1) {
                        vn1[j] = nrm2(m - i - 1, cj + i + 1);
                        vn2[j] = vn1[j];
                    } else {
                        vn1[j] = 0.0;
                        vn2[j] = 0.0;
                    }
                } else {
                    vn1[j] *= std::sqrt(temp);
                }
            }
        }
    }
    for (int i = mn; i < n; ++i) tau[i] = 0.0;
}

// LU with partial pivoting followed by the explicit inverse (getrf + getri
// method).  Restates hmat_tools::Invert for GENERAL matrices
// (src/hmat_tools/Invert.cpp:35-42).  Returns 0, or k+1 if U(k,k) == 0.
int lu_inverse(int n, double* A, int lda) {
    std::vector<int> ipiv(n);
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int p = j + idamax(n - j, A + j + (size_t)j * lda, 1);
        ipiv[j] = p;
        if (A[p + (size_t)j * lda] != 0.0) {
            if (p != j)
                for (int c = 0; c < n; ++c)
                    std::swap(A[j + (size_t)c * lda], A[p + (size_t)c * lda]);
            double r = 1.0 / A[j + (size_t)j * lda];
            for (int i = j + 1; i < n; ++i) A[i + (size_t)j * lda] *= r;
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; ++c) {
            double f = A[j + (size_t)c * lda];
            if (f != 0.0)
                for (int i = j + 1; i < n; ++i)
                    A[i + (size_t)c * lda] -= A[i + (size_t)j * lda] * f;
        }
    }
    if (info) return info;
    // inv(U) in place (unblocked triangular inverse, upper, non-unit).
    for (int j = 0; j < n; ++j) {
        double* cj = A + (size_t)j * lda;
        cj[j] = 1.0 / cj[j];
        double ajj = -cj[j];
        for (int jj = 0; jj < j; ++jj) {
            double t = cj[jj];
            if (t != 0.0) {
                const double* cjj = A + (size_t)jj * lda;
                for (int i = 0; i < jj; ++i) cj[i] += t * cjj[i];
                cj[jj] = t * cjj[jj];
            }
        }
        for (int i = 0; i < j; ++i) cj[i] *= ajj;
    }
    // Solve inv(A) * L = inv(U) for inv(A), last column first.
    std::vector<double> work(n);
    for (int j = n - 1; j >= 0; --j) {
        double* cj = A + (size_t)j * lda;
        for (int i = j + 1; i < n; ++i) { work[i] = cj[i]; cj[i] = 0.0; }
        for (int l = j + 1; l < n; ++l) {
            double f = work[l];
            if (f != 0.0) {
                const double* cl = A + (size_t)l * lda;
                for (int i = 0; i < n; ++i) cj[i] -= f * cl[i];
            }
        }
    }
    for (int j = n - 2; j >= 0; --j) {
        int jp = ipiv[j];
        if (jp != j)
            for (int i = 0; i < n; ++i)
                std::swap(A[i + (size_t)j * lda], A[i + (size_t)jp * lda]);
    }
    return 0;
}

// C(m x n) = alpha * op(A) * B, op(A) = A (m x k) or A^T (A is k x m).
void gemm(bool transA, int m, int n, int k, double alpha, const double* A, int lda,
          const double* B, int ldb, double* C, int ldc) {
    for (int j = 0; j < n; ++j) {
        double* cj = C + (size_t)j * ldc;
        for (int i = 0; i < m; ++i) cj[i] = 0.0;
        if (k == 0) continue;
        if (!transA) {
            for (int l = 0; l < k; ++l) {
                double t = alpha * B[l + (size_t)j * ldb];
                const double* al = A + (size_t)l * lda;
                for (int i = 0; i < m; ++i) cj[i] += t * al[i];
            }
        } else {
            for (int i = 0; i < m; ++i) {
                const double* ai = A + (size_t)i * lda;
                double s = 0.0;
                for (int l = 0; l < k; ++l) s += ai[l] * B[l + (size_t)j * ldb];
                cj[i] = alpha * s;
            }
        }
    }
}

// Solve R X = B in place, R upper triangular k x k (non-unit), B k x n.
// Restates TriangularSolveWrapper -> dtrsm('L','U','N','N')
// (src/InterpDecomp.cpp:213-231).
void trsm_upper(int k, int n, const double* R, int ldr, double* B, int ldb) {
    for (int j = 0; j < n; ++j) {
        double* b = B + (size_t)j * ldb;
        for (int i = k - 1; i >= 0; --i) {
            if (b[i] != 0.0) {
                b[i] /= R[i + (size_t)i * ldr];
                double t = b[i];
                for (int r = 0; r < i; ++r) b[r] -= t * R[r + (size_t)i * ldr];
            }
        }
    }
}

// Interpolative decomposition M ~= M(:, skel) W.  Restates InterpDecomp
// (src/InterpDecomp.cpp:86-138).  M (m x n, lda = m) is overwritten.
// skel/red receive jpvt values (column indices of the original M); W is
// k x q column-major.
void interp_decomp(int m, int n, double* M, double eps, std::vector<int>& skel_cols,
                   std::vector<int>& red_cols, std::vector<double>& W) {
    std::vector<int> jpvt(n);
    std::vector<double> tau(n);
    int lda = std::max(m, 1);
    qrcp(m, n, M, lda, jpvt.data(), tau.data());
    // The reference reads M.Get(i, i) even when i >= m; in its column-major
    // buffer that is element i*(lda+1).  Reads past the buffer are undefined
    // behaviour in the reference; they are taken as 0 here.
    size_t total = (size_t)lda * n;
    auto diag = [&](int i) -> double {
        size_t idx = (size_t)i * (lda + 1);
        return idx < total ? M[idx] : 0.0;
    };
    double tol = std::fabs(diag(0)) * eps;  // src/InterpDecomp.cpp:101
    std::vector<int> skel, red;
    for (int i = 0; i < n; ++i) {
        if (std::fabs(diag(i)) > tol) {  // IsSkel, src/InterpDecomp.cpp:78-84
            skel_cols.push_back(jpvt[i]);
            skel.push_back(i);
        } else {
            red_cols.push_back(jpvt[i]);
            red.push_back(i);
        }
    }
    int k = (int)skel.size(), q = (int)red.size();
    // W_full(i, j) = R(skel[i], j) for j >= i, else 0 (src/InterpDecomp.cpp:115-124)
    std::vector<double> Wf((size_t)k * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < k && i <= j; ++i) {
            size_t idx = (size_t)skel[i] + (size_t)j * lda;
            Wf[i + (size_t)j * k] = idx < total ? M[idx] : 0.0;
        }
    std::vector<double> Rs((size_t)k * k);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) Rs[i + (size_t)j * k] = Wf[i + (size_t)j * k];
    if (k > 0) trsm_upper(k, n, Rs.data(), k, Wf.data(), k);
    W.assign((size_t)k * q, 0.0);
    for (int j = 0; j < q; ++j)
        for (int i = 0; i < k; ++i) W[i + (size_t)j * k] = Wf[i + (size_t)red[j] * k];
}

struct Box {
    std::vector<int> g, red, skel;  // red/skel: positions into g
    std::vector<int> rows;          // faces: rows of the ID matrix (global DOF ids)
    std::vector<double> A22, A22inv, X, W, S;
    int face = -1;
};

// Schur complement of the red block.  Restates Schur (include/Schur.hpp:72-100):
// A12 = M(skel, red), A21 = M(red, skel), A22 = M(red, red),
// X = A22^{-1} A21, S = -A12 X.  M is n x n column-major.
int schur(const double* M, int n, Box& b) {
    int r = (int)b.red.size(), s = (int)b.skel.size();
    std::vector<double> A12((size_t)s * r), A21((size_t)r * s);
    b.A22.assign((size_t)r * r, 0.0);
    for (int j = 0; j < r; ++j)
        for (int i = 0; i < s; ++i) A12[i + (size_t)j * s] = M[b.skel[i] + (size_t)b.red[j] * n];
    for (int j = 0; j < s; ++j)
        for (int i = 0; i < r; ++i) A21[i + (size_t)j * r] = M[b.red[i] + (size_t)b.skel[j] * n];
    for (int j = 0; j < r; ++j)
        for (int i = 0; i < r; ++i) b.A22[i + (size_t)j * r] = M[b.red[i] + (size_t)b.red[j] * n];
    b.A22inv = b.A22;
    int info = lu_inverse(r, b.A22inv.data(), std::max(r, 1));
    b.X.assign((size_t)r * s, 0.0);
    gemm(false, r, s, r, 1.0, b.A22inv.data(), std::max(r, 1), A21.data(), std::max(r, 1),
         b.X.data(), std::max(r, 1));
    b.S.assign((size_t)s * s, 0.0);
    gemm(false, s, s, r, -1.0, A12.data(), std::max(s, 1), b.X.data(), std::max(r, 1),
         b.S.data(), std::max(s, 1));
    return info;
}

// ---------------------------------------------------------------------------
// Level matrix: A(i, j) keyed by (i << 32 | j); absent entries read as 0.
// ---------------------------------------------------------------------------
struct LevelMatrix {
    static constexpr uint64_t EMPTY = ~0ull;
    std::vector<uint64_t> keys;
    std::vector<double> vals;
    size_t count = 0;
    uint64_t mask = 0;

    static uint64_t mix(uint64_t x) {
        x ^= x >> 33; x *= 0xff51afd7ed558ccdULL;
        x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
        x ^= x >> 33;
        return x;
    }
    void init(size_t cap) {
        size_t c = 16;
        while (c < cap) c <<= 1;
        keys.assign(c, EMPTY);
        vals.assign(c, 0.0);
        mask = c - 1;
        count = 0;
    }
    double find(int i, int j) const {
        uint64_t key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
        uint64_t h = mix(key) & mask;
        while (true) {
            uint64_t k = keys[h];
            if (k == key) return vals[h];
            if (k == EMPTY) return 0.0;
            h = (h + 1) & mask;
        }
    }
    void grow() {
        std::vector<uint64_t> ok;
        std::vector<double> ov;
        ok.swap(keys);
        ov.swap(vals);
        init(ok.size() * 2);
        for (size_t t = 0; t < ok.size(); ++t)
            if (ok[t] != EMPTY) insert_new(ok[t], ov[t]);
    }
    void insert_new(uint64_t key, double v) {
        uint64_t h = mix(key) & mask;
        while (keys[h] != EMPTY) h = (h + 1) & mask;
        keys[h] = key;
        vals[h] = v;
        ++count;
    }
    // sparsemat_[j][i] += v  (include/hifde3d/core/sparse.hpp:80-86)
    void add(int i, int j, double v) {
        if ((count + 1) * 2 > keys.size()) grow();
        uint64_t key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
        uint64_t h = mix(key) & mask;
        while (true) {
            uint64_t k = keys[h];
            if (k == key) { vals[h] += v; return; }
            if (k == EMPTY) { keys[h] = key; vals[h] = v; ++count; return; }
            h = (h + 1) & mask;
        }
    }
    // DeleteRow + DeleteCol of all dead DOFs (sparse.hpp:173-206), as a
    // compaction that keeps only entries whose row and column are alive.
    void compact(const std::vector<uint8_t>& alive) {
        std::vector<uint64_t> ok;
        std::vector<double> ov;
        ok.swap(keys);
        ov.swap(vals);
        size_t live = 0;
        for (size_t t = 0; t < ok.size(); ++t)
            if (ok[t] != EMPTY && alive[ok[t] >> 32] && alive[ok[t] & 0xffffffffu]) ++live;
        init(live * 2 + 16);
        for (size_t t = 0; t < ok.size(); ++t)
            if (ok[t] != EMPTY && alive[ok[t] >> 32] && alive[ok[t] & 0xffffffffu])
                insert_new(ok[t], ov[t]);
    }
};

enum Face { TOP = 0, BOTTOM, RIGHT, LEFT, FRONT, BACK };  // include/data.hpp:7

struct Oracle {
    int N, P, NC, L;
    double eps;
    bool quirk;
    LevelMatrix A;
    std::vector<uint8_t> remaining;
    std::vector<std::vector<Box>> schur_data, skel_data;
    std::vector<double> t_schur, t_skel, t_update;
    int lu_info = 0;

    Oracle(int N_, int P_, double eps_, bool quirk_)
        : N(N_), P(P_), NC(N_ + 1), eps(eps_), quirk(quirk_) {
        // src/Factor.cpp:22-38
        remaining.assign((size_t)NC * NC * NC, 0);
        for (int k = 1; k < NC; ++k)
            for (int j = 1; j < NC; ++j)
                for (int i = 1; i < NC; ++i) remaining[lin(i, j, k)] = 1;
        L = (int)std::lround(std::log2((double)(NC / P))) + 1;
        schur_data.resize(L);
        skel_data.resize(L);
        t_schur.assign(L, 0.0);
        t_skel.assign(L, 0.0);
        t_update.assign(L, 0.0);
        A.init(16);
    }

    int lin(int i, int j, int k) const { return i + NC * j + NC * NC * k; }  // :78-89
    int width(int level) const { return (1 << level) * P; }
    bool interior(int level, int a) const {  // IsInterior :314-320
        return a > 0 && a <= N && (a % width(level)) != 0;
    }
    int n_interior(int level, int i, int j, int k) const {
        return (int)interior(level, i) + (int)interior(level, j) + (int)interior(level, k);
    }
    bool rem(int i, int j, int k) const { return remaining[lin(i, j, k)] != 0; }

    // InteriorCellIndexData (src/Factor.cpp:359-392)
    void cell_index(int ci, int cj, int ck, int level, Box& b) const {
        int w = width(level);
        int lo[3] = {std::max(w * ci, 1), std::max(w * cj, 1), std::max(w * ck, 1)};
        int hi[3] = {std::min(w * (ci + 1), N), std::min(w * (cj + 1), N), std::min(w * (ck + 1), N)};
        int pos = 0;
        for (int i = lo[0]; i <= hi[0]; ++i)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int k = lo[2]; k <= hi[2]; ++k) {
                    if (!rem(i, j, k)) continue;
                    int ni = n_interior(level, i, j, k);
                    if (ni == 3) { b.g.push_back(lin(i, j, k)); b.red.push_back(pos++); }
                    else if (ni == 2) { b.g.push_back(lin(i, j, k)); b.skel.push_back(pos++); }
                }
    }

    // InteriorFaceDOFs (want == 2) and InteriorEdgeDOFs (want == 1)
    // (src/Factor.cpp:402-643), including the FRONT guard at :486/:610 that
    // tests k == 0 instead of k == 1 when `quirk` is set.
    void face_dofs(int ci, int cj, int ck, int face, int level, int want,
                   std::vector<int>& out) const {
        int w = width(level);
        int lo[3] = {std::max(w * ci, 1), std::max(w * cj, 1), std::max(w * ck, 1)};
        int hi[3] = {std::min(w * (ci + 1), N), std::min(w * (cj + 1), N), std::min(w * (ck + 1), N)};
        auto take = [&](int i, int j, int k) {
            if (rem(i, j, k) && n_interior(level, i, j, k) == want) out.push_back(lin(i, j, k));
        };
        switch (face) {
        case TOP: {
            int i = lo[0];
            if (i == 1) return;
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int k = lo[2]; k <= hi[2]; ++k) take(i, j, k);
            return;
        }
        case BOTTOM: {
            int i = hi[0];
            if (i == N) return;
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int k = lo[2]; k <= hi[2]; ++k) take(i, j, k);
            return;
        }
        case LEFT: {
            int j = lo[1];
            if (j == 1) return;
            for (int i = lo[0]; i <= hi[0]; ++i)
                for (int k = lo[2]; k <= hi[2]; ++k) take(i, j, k);
            return;
        }
        case RIGHT: {
            int j = hi[1];
            if (j == N) return;
            for (int i = lo[0]; i <= hi[0]; ++i)
                for (int k = lo[2]; k <= hi[2]; ++k) take(i, j, k);
            return;
        }
        case FRONT: {
            int k = lo[2];
            if (k == (quirk ? 0 : 1)) return;
            for (int i = lo[0]; i <= hi[0]; ++i)
                for (int j = lo[1]; j <= hi[1]; ++j) take(i, j, k);
            return;
        }
        case BACK: {
            int k = hi[2];
            if (k == N) return;
            for (int i = lo[0]; i <= hi[0]; ++i)
                for (int j = lo[1]; j <= hi[1]; ++j) take(i, j, k);
            return;
        }
        }
    }

    // SkelInteractionIndexData (src/Factor.cpp:645-755).  Returns false when
    // the face does not exist.
    bool skel_index(int ci, int cj, int ck, int face, int level, std::vector<int>& rows,
                    std::vector<int>& cols) const {
        int w = width(level);
        int lo[3] = {std::max(w * ci, 1), std::max(w * cj, 1), std::max(w * ck, 1)};
        static const int others_top[5] = {BOTTOM, LEFT, RIGHT, FRONT, BACK};
        static const int nbr_top[5] = {TOP, LEFT, RIGHT, FRONT, BACK};
        static const int others_left[5] = {TOP, BOTTOM, RIGHT, FRONT, BACK};
        static const int nbr_left[5] = {TOP, BOTTOM, LEFT, FRONT, BACK};
        static const int others_front[5] = {TOP, BOTTOM, LEFT, RIGHT, BACK};
        static const int nbr_front[5] = {TOP, BOTTOM, LEFT, RIGHT, FRONT};
        const int *own, *nbr;
        int di = 0, dj = 0, dk = 0;
        if (face == TOP) {
            if (lo[0] == 1) return false;
            own = others_top; nbr = nbr_top; di = 1;
        } else if (face == LEFT) {
            if (lo[1] == 1) return false;
            own = others_left; nbr = nbr_left; dj = 1;
        } else {
            if (lo[2] == 1) return false;
            own = others_front; nbr = nbr_front; dk = 1;
        }
        face_dofs(ci, cj, ck, face, level, 2, cols);
        face_dofs(ci, cj, ck, face, level, 1, rows);
        for (int t = 0; t < 5; ++t) face_dofs(ci, cj, ck, own[t], level, 2, rows);
        for (int t = 0; t < 5; ++t) face_dofs(ci - di, cj - dj, ck - dk, nbr[t], level, 2, rows);
        return true;
    }

    void gather(const std::vector<int>& rows, const std::vector<int>& cols,
                std::vector<double>& M) const {  // Sparse::Find (sparse.hpp:273-304)
        size_t m = rows.size(), n = cols.size();
        M.assign(std::max<size_t>(m, 1) * n, 0.0);
        for (size_t j = 0; j < n; ++j)
            for (size_t i = 0; i < m; ++i) M[i + j * std::max<size_t>(m, 1)] = A.find(rows[i], cols[j]);
    }

    // SchurAfterID (src/Factor.cpp:113-153): Rot = I, Rot(skel, red) = -W,
    // result = Rot^T A(F, F) Rot, then Schur(result).
    void schur_after_id(Box& b) {
        int n = (int)b.g.size();
        std::vector<double> Rot((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) Rot[i + (size_t)i * n] = 1.0;
        int k = (int)b.skel.size();
        for (size_t j = 0; j < b.red.size(); ++j)
            for (int i = 0; i < k; ++i)
                Rot[b.skel[i] + (size_t)b.red[j] * n] = -b.W[i + j * k];
        std::vector<double> sub, tmp((size_t)n * n), res((size_t)n * n);
        gather(b.g, b.g, sub);
        gemm(false, n, n, n, 1.0, sub.data(), n, Rot.data(), n, tmp.data(), n);
        gemm(true, n, n, n, 1.0, Rot.data(), n, tmp.data(), n, res.data(), n);
        int info = schur(res.data(), n, b);
        if (info && !lu_info) lu_info = info;
    }

    // UpdateMatrixAndDOFs (src/Factor.cpp:262-311) + UpdateRemainingDOFs (:91-111).
    void update(std::vector<Box>& level) {
        for (Box& b : level)
            for (int r : b.red) remaining[b.g[r]] = 0;
        for (Box& b : level) {
            int s = (int)b.skel.size();
            for (int i = 0; i < s; ++i)
                for (int j = 0; j < s; ++j)
                    A.add(b.g[b.skel[i]], b.g[b.skel[j]], b.S[i + (size_t)j * s]);
            std::vector<double>().swap(b.S);  // S.Clear() (:293)
        }
        A.compact(remaining);
    }

    // LevelFactorSchur (src/Factor.cpp:184-213)
    void level_schur(int c, int level, int nthreads) {
        std::vector<Box>& data = schur_data[level];
        data.assign((size_t)c * c * c, Box());
        for (int i = 0; i < c; ++i)
            for (int j = 0; j < c; ++j)
                for (int k = 0; k < c; ++k) cell_index(i, j, k, level, data[(i * c + j) * c + k]);
        long nb = (long)data.size();
        int info_any = 0;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) reduction(max : info_any)
        for (long t = 0; t < nb; ++t) {
            Box& b = data[t];
            std::vector<double> M;
            gather(b.g, b.g, M);
            int info = schur(M.data(), (int)b.g.size(), b);
            if (info) info_any = std::max(info_any, info);
        }
        if (info_any && !lu_info) lu_info = info_any;
    }

    // LevelFactorSkel (src/Factor.cpp:232-260) with Skeletonize (:155-182)
    void level_skel(int c, int level, int nthreads) {
        std::vector<Box>& data = skel_data[level];
        data.clear();
        std::vector<std::vector<int>> rows_list;
        for (int i = 0; i < c; ++i)
            for (int j = 0; j < c; ++j)
                for (int k = 0; k < c; ++k) {
                    const int faces[3] = {TOP, FRONT, LEFT};
                    for (int f : faces) {
                        std::vector<int> rows, cols;
                        if (!skel_index(i, j, k, f, level, rows, cols)) continue;
                        if (cols.empty()) continue;
                        Box b;
                        b.face = f;
                        b.g = cols;
                        data.push_back(std::move(b));
                        rows_list.push_back(std::move(rows));
                    }
                }
        long nb = (long)data.size();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
        for (long t = 0; t < nb; ++t) {
            Box& b = data[t];
            std::vector<double> M;
            gather(rows_list[t], b.g, M);
            interp_decomp((int)rows_list[t].size(), (int)b.g.size(), M.data(), eps, b.skel, b.red,
                          b.W);
            schur_after_id(b);
            b.rows = std::move(rows_list[t]);
        }
    }

    // Factor (src/Factor.cpp:41-63)
    void factor(int nthreads) {
        using clk = std::chrono::steady_clock;
        for (int level = 0; level < L; ++level) {
            int c = NC / width(level);
            auto t0 = clk::now();
            level_schur(c, level, nthreads);
            auto t1 = clk::now();
            update(schur_data[level]);
            auto t2 = clk::now();
            t_schur[level] = std::chrono::duration<double>(t1 - t0).count();
            t_update[level] += std::chrono::duration<double>(t2 - t1).count();
            if (level < L - 1) {
                level_skel(c, level, nthreads);
                auto t3 = clk::now();
                update(skel_data[level]);
                auto t4 = clk::now();
                t_skel[level] = std::chrono::duration<double>(t3 - t2).count();
                t_update[level] += std::chrono::duration<double>(t4 - t3).count();
            }
        }
    }

    // Apply (src/Factor.cpp:992-1081) with UpdateSkeleton/UpdateRedundant
    // (:891-950) and ApplyA22 (:957-975).
    void apply(double* u, bool inverse) const {
        for (int k = 0; k < NC; ++k)  // ZeroOutBoundary (:977-990)
            for (int j = 0; j < NC; ++j)
                for (int i = 0; i < NC; ++i)
                    if (i == 0 || j == 0 || k == 0) u[lin(i, j, k)] = 0.0;
        std::vector<double> us, ur;
        auto load = [&](const Box& b) {
            us.resize(b.skel.size());
            ur.resize(b.red.size());
            for (size_t i = 0; i < b.skel.size(); ++i) us[i] = u[b.g[b.skel[i]]];
            for (size_t i = 0; i < b.red.size(); ++i) ur[i] = u[b.g[b.red[i]]];
        };
        // u_skel += alpha op(Mat) u_red ; Mat is (rows x cols) col-major
        auto upd_skel = [&](const Box& b, const std::vector<double>& Mt, int mr, int mc, bool adj,
                            double alpha) {
            if (us.empty() || ur.empty()) return;
            std::vector<double> y(us);
            if (adj) {  // Mt is mr x mc with mr == |red|, mc == |skel|
                for (int j = 0; j < mc; ++j) {
                    double s = 0.0;
                    for (int i = 0; i < mr; ++i) s += Mt[i + (size_t)j * mr] * ur[i];
                    y[j] += alpha * s;
                }
            } else {  // Mt is mr x mc with mr == |skel|, mc == |red|
                for (int j = 0; j < mc; ++j) {
                    double t = alpha * ur[j];
                    for (int i = 0; i < mr; ++i) y[i] += t * Mt[i + (size_t)j * mr];
                }
            }
            us = y;
            for (size_t i = 0; i < b.skel.size(); ++i) u[b.g[b.skel[i]]] = us[i];
        };
        auto upd_red = [&](const Box& b, const std::vector<double>& Mt, int mr, int mc, bool adj,
                           double alpha) {
            if (us.empty() || ur.empty()) return;
            std::vector<double> y(ur);
            if (adj) {  // Mt is mr x mc with mr == |skel|, mc == |red|
                for (int j = 0; j < mc; ++j) {
                    double s = 0.0;
                    for (int i = 0; i < mr; ++i) s += Mt[i + (size_t)j * mr] * us[i];
                    y[j] += alpha * s;
                }
            } else {  // Mt is mr x mc with mr == |red|, mc == |skel|
                for (int j = 0; j < mc; ++j) {
                    double t = alpha * us[j];
                    for (int i = 0; i < mr; ++i) y[i] += t * Mt[i + (size_t)j * mr];
                }
            }
            ur = y;
            for (size_t i = 0; i < b.red.size(); ++i) u[b.g[b.red[i]]] = ur[i];
        };
        for (int level = 0; level < L - 1; ++level) {
            for (const Box& b : schur_data[level]) {
                load(b);
                int r = (int)b.red.size(), s = (int)b.skel.size();
                if (inverse) upd_skel(b, b.X, r, s, true, -1.0);
                else upd_red(b, b.X, r, s, false, 1.0);
            }
            for (const Box& b : skel_data[level]) {
                load(b);
                int q = (int)b.red.size(), k = (int)b.skel.size();
                if (inverse) {
                    upd_red(b, b.W, k, q, true, -1.0);
                    upd_skel(b, b.X, q, k, true, -1.0);
                } else {
                    upd_skel(b, b.W, k, q, false, 1.0);
                    upd_red(b, b.X, q, k, false, 1.0);
                }
            }
        }
        for (int level = 0; level < L; ++level) {
            for (int pass = 0; pass < 2; ++pass) {
                const std::vector<Box>& data = pass == 0 ? schur_data[level] : skel_data[level];
                for (const Box& b : data) {
                    int r = (int)b.red.size();
                    if (r == 0) continue;
                    std::vector<double> x(r), y(r, 0.0);
                    for (int i = 0; i < r; ++i) x[i] = u[b.g[b.red[i]]];
                    const std::vector<double>& Mt = inverse ? b.A22inv : b.A22;
                    for (int j = 0; j < r; ++j)
                        for (int i = 0; i < r; ++i) y[i] += Mt[i + (size_t)j * r] * x[j];
                    for (int i = 0; i < r; ++i) u[b.g[b.red[i]]] = y[i];
                }
            }
        }
        for (int level = L - 2; level >= 0; --level) {
            for (const Box& b : skel_data[level]) {
                load(b);
                int q = (int)b.red.size(), k = (int)b.skel.size();
                if (inverse) {
                    upd_red(b, b.X, q, k, false, -1.0);
                    upd_skel(b, b.W, k, q, false, -1.0);
                } else {
                    upd_skel(b, b.X, q, k, true, 1.0);
                    upd_red(b, b.W, k, q, true, 1.0);
                }
            }
            for (const Box& b : schur_data[level]) {
                load(b);
                int r = (int)b.red.size(), s = (int)b.skel.size();
                if (inverse) upd_red(b, b.X, r, s, false, -1.0);
                else upd_skel(b, b.X, r, s, true, 1.0);
            }
        }
    }
};

// SetupStencil (include/setup_stencil.hpp:6-76): nodal coefficient tensors a
// and V of size (NC+1)^3 indexed i + j*m + k*m*n (include/numtns.hpp:99).
// symmetric == 0 reproduces the reference exactly (off-diagonal -a(q)/h^2);
// symmetric == 1 uses -(a(p)+a(q))/(2h^2) and the matching row-sum diagonal,
// which equals the reference stencil when a is constant.
long stencil_coo(int N, double h, const double* a, const double* V, int symmetric, int* rows,
                 int* cols, double* vals) {
    int NC = N + 1, m = NC + 1;
    auto T = [&](const double* t, int i, int j, int k) { return t[i + (size_t)m * (j + (size_t)m * k)]; };
    auto L = [&](int i, int j, int k) { return i + NC * j + NC * NC * k; };
    double hh = h * h;
    long nnz = 0;
    for (int i = 1; i <= N; ++i)
        for (int j = 1; j <= N; ++j)
            for (int k = 1; k <= N; ++k) {
                const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                                      {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                double ap = T(a, i, j, k);
                double d = 0.0;
                if (!symmetric) {
                    d += (T(a, i - 1, j, k) + T(a, i + 1, j, k)) / hh;
                    d += (T(a, i, j - 1, k) + T(a, i, j + 1, k)) / hh;
                    d += (T(a, i, j, k - 1) + T(a, i, j, k + 1)) / hh;
                } else {
                    for (int t = 0; t < 6; ++t)
                        d += 0.5 * (ap + T(a, nb[t][0], nb[t][1], nb[t][2])) / hh;
                }
                d += T(V, i, j, k);
                int p = L(i, j, k);
                rows[nnz] = p; cols[nnz] = p; vals[nnz] = d; ++nnz;
                for (int t = 0; t < 6; ++t) {
                    int x = nb[t][0], y = nb[t][1], z = nb[t][2];
                    if (x < 1 || x > N || y < 1 || y > N || z < 1 || z > N) continue;
                    double aq = T(a, x, y, z);
                    double v = symmetric ? -0.5 * (ap + aq) / hh : -aq / hh;
                    rows[nnz] = p; cols[nnz] = L(x, y, z); vals[nnz] = v; ++nnz;
                }
            }
    return nnz;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI used by the tests (ctypes) and the CPU baseline leg of bench.py.
// ---------------------------------------------------------------------------
extern "C" {

void* hifo_create(int N, int P, double eps, int front_quirk) {
    return new Oracle(N, P, eps, front_quirk != 0);
}
void hifo_destroy(void* h) { delete static_cast<Oracle*>(h); }
int hifo_num_levels(void* h) { return static_cast<Oracle*>(h)->L; }

// Sparse::Add semantics: duplicates are summed in the order given.
int hifo_set_matrix_coo(void* h, long nnz, const int* rows, const int* cols, const double* vals) {
    Oracle* o = static_cast<Oracle*>(h);
    o->A.init((size_t)nnz * 2 + 16);
    for (long t = 0; t < nnz; ++t) o->A.add(rows[t], cols[t], vals[t]);
    return 0;
}

int hifo_factor(void* h, int nthreads) {
    Oracle* o = static_cast<Oracle*>(h);
    o->factor(nthreads > 0 ? nthreads : 1);
    return o->lu_info;
}

int hifo_apply(void* h, double* u, int inverse) {
    static_cast<Oracle*>(h)->apply(u, inverse != 0);
    return 0;
}

// kind 0 = Schur (cells), 1 = Skel (faces)
int hifo_num_boxes(void* h, int level, int kind) {
    Oracle* o = static_cast<Oracle*>(h);
    if (level < 0 || level >= o->L) return -1;
    return (int)(kind == 0 ? o->schur_data[level].size() : o->skel_data[level].size());
}

// sizes[0..2] = |global|, |red|, |skel|; sizes[3] = face id (-1 for cells)
int hifo_box_sizes(void* h, int level, int kind, int box, int* sizes) {
    Oracle* o = static_cast<Oracle*>(h);
    const std::vector<Box>& d = kind == 0 ? o->schur_data[level] : o->skel_data[level];
    if (box < 0 || box >= (int)d.size()) return -1;
    const Box& b = d[box];
    sizes[0] = (int)b.g.size();
    sizes[1] = (int)b.red.size();
    sizes[2] = (int)b.skel.size();
    sizes[3] = b.face;
    return 0;
}

// Any output pointer may be null.
int hifo_box_data(void* h, int level, int kind, int box, int* g, int* red, int* skel, double* A22,
                  double* A22inv, double* X, double* W) {
    Oracle* o = static_cast<Oracle*>(h);
    const std::vector<Box>& d = kind == 0 ? o->schur_data[level] : o->skel_data[level];
    if (box < 0 || box >= (int)d.size()) return -1;
    const Box& b = d[box];
    if (g) std::copy(b.g.begin(), b.g.end(), g);
    if (red) std::copy(b.red.begin(), b.red.end(), red);
    if (skel) std::copy(b.skel.begin(), b.skel.end(), skel);
    if (A22) std::copy(b.A22.begin(), b.A22.end(), A22);
    if (A22inv) std::copy(b.A22inv.begin(), b.A22inv.end(), A22inv);
    if (X) std::copy(b.X.begin(), b.X.end(), X);
    if (W) std::copy(b.W.begin(), b.W.end(), W);
    return 0;
}

// Rows of a face's ID matrix (SkelInteractionIndexData's global_rows, in its
// order, duplicates kept).  Returns the row count; out may be null.
int hifo_face_rows(void* h, int level, int box, int* out) {
    Oracle* o = static_cast<Oracle*>(h);
    if (level < 0 || level >= o->L) return -1;
    const std::vector<Box>& d = o->skel_data[level];
    if (box < 0 || box >= (int)d.size()) return -1;
    if (out) std::copy(d[box].rows.begin(), d[box].rows.end(), out);
    return (int)d[box].rows.size();
}

// Per-level wall time of the last factor (seconds): out[0..2] = schur, skel, update.
int hifo_level_times(void* h, int level, double* out) {
    Oracle* o = static_cast<Oracle*>(h);
    if (level < 0 || level >= o->L) return -1;
    out[0] = o->t_schur[level];
    out[1] = o->t_skel[level];
    out[2] = o->t_update[level];
    return 0;
}

long hifo_stencil_coo(int N, double h, const double* a, const double* V, int symmetric, int* rows,
                      int* cols, double* vals) {
    return stencil_coo(N, h, a, V, symmetric, rows, cols, vals);
}

// Dense kernels exposed for pinning.
int hifo_qrcp(int m, int n, double* A, int lda, int* jpvt, double* tau) {
    qrcp(m, n, A, lda, jpvt, tau);
    return 0;
}

int hifo_lu_inverse(int n, double* A) { return lu_inverse(n, A, std::max(n, 1)); }

// ID of M (m x n, column-major, overwritten).  skel/red receive column
// indices; W (k x q) must hold n*n doubles.  counts[0] = k, counts[1] = q.
int hifo_interp_decomp(int m, int n, double* M, double eps, int* skel, int* red, double* W,
                       int* counts) {
    std::vector<int> s, r;
    std::vector<double> w;
    interp_decomp(m, n, M, eps, s, r, w);
    std::copy(s.begin(), s.end(), skel);
    std::copy(r.begin(), r.end(), red);
    std::copy(w.begin(), w.end(), W);
    counts[0] = (int)s.size();
    counts[1] = (int)r.size();
    return 0;
}

// Schur on a dense n x n M given red/skel positions.  Outputs A22inv (r x r),
// X (r x s), S (s x s).
int hifo_schur(const double* M, int n, const int* red, int r, const int* skel, int s,
               double* A22inv, double* X, double* S) {
    Box b;
    b.red.assign(red, red + r);
    b.skel.assign(skel, skel + s);
    int info = schur(M, n, b);
    std::copy(b.A22inv.begin(), b.A22inv.end(), A22inv);
    std::copy(b.X.begin(), b.X.end(), X);
    std::copy(b.S.begin(), b.S.end(), S);
    return info;
}

}  // extern "C"
