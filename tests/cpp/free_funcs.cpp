// The drop-in's free functions -- DenseSubmatrix (x3), PivotedQRWrapper,
// TriangularSolveWrapper, InterpDecomp and Schur -- called through
// include/hifde3d_amd.hpp exactly as reference code calls them
// (include/Schur.hpp, include/InterpDecomp.hpp).  Reads its inputs from a
// binary file and writes every output to another, for the GPU test to check
// against the oracle and NumPy.  Usage: free_funcs IN OUT
//
// IN : int32 count, then per matrix int32 m, n and m*n doubles (column-major)
//      -- the QR / ID inputs; double eps;
//      int32 count, then per pair int32 m, n, R (m*m doubles), B (m*n doubles);
//      int32 m, n, dense (m*n doubles), int32 nr, rows, int32 nc, cols;
//      int64 nnz, per entry int32 row, int32 col, double val;
//      int32 r, s, Schur matrix ((r+s)^2 doubles; red = first r, skel = rest).
// OUT: per QR input: QR factors (m*n), jpvt (n int32);
//      per ID input: overwritten M (m*n), int32 k, skel (k), red (n-k), W (k*(n-k));
//      per solve: R^{-1} B (m*n);
//      dense(rows, cols), dense(:, cols), sparse(rows, cols);
//      A_22^{-1} (r*r), X (r*s), S (s*s).
#include <cstdio>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "hifde3d_amd.hpp"

using namespace hifde3d;

namespace {
struct In {
    FILE* f;
    template <class T>
    T get() {
        T v;
        if (fread(&v, sizeof v, 1, f) != 1) throw std::runtime_error("short input");
        return v;
    }
    Dense<double> dense(int m, int n) {
        Dense<double> D(m, n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) D.Set(i, j, get<double>());
        return D;
    }
    std::vector<int> ints() {
        int n = get<int32_t>();
        std::vector<int> v(n);
        for (int& x : v) x = get<int32_t>();
        return v;
    }
};
struct Out {
    FILE* f;
    void put(const Dense<double>& D) {
        for (int j = 0; j < D.Width(); ++j)
            for (int i = 0; i < D.Height(); ++i) {
                double v = D.Get(i, j);
                fwrite(&v, sizeof v, 1, f);
            }
    }
    void put(const std::vector<int>& v) {
        for (int x : v) {
            int32_t y = x;
            fwrite(&y, sizeof y, 1, f);
        }
    }
    void put(int x) {
        int32_t y = x;
        fwrite(&y, sizeof y, 1, f);
    }
};
}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: free_funcs IN OUT\n");
        return 2;
    }
    In in{fopen(argv[1], "rb")};
    Out out{fopen(argv[2], "wb")};
    if (!in.f || !out.f) {
        fprintf(stderr, "cannot open files\n");
        return 2;
    }
    const int nq = in.get<int32_t>();
    std::vector<Dense<double>> mats;
    for (int t = 0; t < nq; ++t) {
        int m = in.get<int32_t>(), n = in.get<int32_t>();
        mats.push_back(in.dense(m, n));
    }
    const double eps = in.get<double>();
    for (const Dense<double>& M0 : mats) {
        Dense<double> A = M0;
        std::vector<int> jpvt;
        PivotedQRWrapper(A, jpvt);
        out.put(A);
        out.put(jpvt);
    }
    for (const Dense<double>& M0 : mats) {
        Dense<double> M = M0, W;
        std::vector<int> skel, red;
        InterpDecomp(M, W, skel, red, eps);
        out.put(M);
        out.put((int)skel.size());
        out.put(skel);
        out.put(red);
        out.put(W);
    }
    const int ns = in.get<int32_t>();
    for (int t = 0; t < ns; ++t) {
        int m = in.get<int32_t>(), n = in.get<int32_t>();
        Dense<double> R = in.dense(m, m), B = in.dense(m, n);
        TriangularSolveWrapper(R, B);
        out.put(B);
    }
    {
        int m = in.get<int32_t>(), n = in.get<int32_t>();
        Dense<double> D = in.dense(m, n), S1, S2, S3;
        std::vector<int> rows = in.ints(), cols = in.ints();
        DenseSubmatrix(D, rows, cols, S1);
        DenseSubmatrix(D, cols, S2);
        long long nnz = in.get<int64_t>();
        Sparse<double> sp(m, n);
        for (long long e = 0; e < nnz; ++e) {
            int i = in.get<int32_t>(), j = in.get<int32_t>();
            sp.Add(i, j, in.get<double>());
        }
        DenseSubmatrix(sp, rows, cols, S3);
        out.put(S1);
        out.put(S2);
        out.put(S3);
    }
    {
        int r = in.get<int32_t>(), s = in.get<int32_t>();
        Dense<double> M = in.dense(r + s, r + s);
        FactorData<double> data;
        for (int i = 0; i < r; ++i) data.ind_data().redundant_inds().push_back(i);
        for (int i = 0; i < s; ++i) data.ind_data().skeleton_inds().push_back(r + i);
        Schur(M, data);
        out.put(data.A_22_inv());
        out.put(data.X_mat());
        out.put(data.Schur_comp());
    }
    fclose(out.f);
    fclose(in.f);
    return 0;
}
