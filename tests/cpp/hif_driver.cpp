// Reference driver protocol (tests/HIFDE3D/main.cpp) written against the
// drop-in header: a = 1, V = 0, v = 1 on the interior, w = SpMV(A, v) taken
// before Factor(), then e_a = ||w - F v|| / ||w|| and
// e_s = ||v - F^{-1} w|| / ||v||.  Usage: hif_driver -tol T -N N -P P
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <map>
#include <string>

#include "hifde3d_amd.hpp"

using namespace hifde3d;

int main(int argc, char** argv) {
    std::map<std::string, std::string> opts;
    for (int k = 1; k + 1 < argc; k += 2) opts[argv[k]] = argv[k + 1];
    double epsilon = opts.count("-tol") ? atof(opts["-tol"].c_str()) : 1e-3;
    int N = opts.count("-N") ? atoi(opts["-N"].c_str()) : 31;
    int P = opts.count("-P") ? atoi(opts["-P"].c_str()) : 4;

    HIFFactor<double> factor(N, P, epsilon);
    int NC = factor.N() + 1;
    NumTns<double> A(NC + 1, NC + 1, NC + 1), V(NC + 1, NC + 1, NC + 1);
    for (int i = 0; i < NC + 1; ++i)
        for (int j = 0; j < NC + 1; ++j)
            for (int k = 0; k < NC + 1; ++k) {
                A(i, j, k) = 1;
                V(i, j, k) = 0;
            }
    SetupStencil(factor, NC - 1, 1.0 / NC, A, V);
    Vector<double> v(NC * NC * NC);
    for (int k = 0; k < NC; ++k)
        for (int j = 0; j < NC; ++j)
            for (int i = 0; i < NC; ++i)
                v.Set(factor.Tensor2LinearInd(Index3(i, j, k)), (i && j && k) ? 1.0 : 0.0);
    Vector<double> w(NC * NC * NC);
    SpMV(factor.sp_matrix(), v, w);
    factor.Initialize();
    factor.Factor();
    Vector<double> u(v.Size());
    for (int i = 0; i < u.Size(); ++i) u.Set(i, v.Get(i));
    factor.Apply(u, false);
    double e_a = RelativeErrorNorm2(w, u);
    factor.Apply(w, true);
    double e_s = RelativeErrorNorm2(v, w);
    std::vector<std::vector<int>> lc = factor.LevelCounts();
    for (size_t l = 0; l < lc.size(); ++l)
        printf("level %zu: cells %d eliminated %d faces %d eliminated %d\n", l, lc[l][0], lc[l][1],
               lc[l][2], lc[l][3]);
    printf("e_a %.17g\ne_s %.17g\n", e_a, e_s);
    // FactorData of every box, materialised through the drop-in
    std::vector<std::vector<FactorData<double>>>& sd = factor.schur_level_data();
    std::vector<std::vector<FactorData<double>>>& kd = factor.skel_level_data();
    double inv_err = 0.0;
    for (size_t l = 0; l < sd.size(); ++l) {
        int rc = 0, rf = 0;
        for (FactorData<double>& d : sd[l]) {
            rc += d.NumDOFsEliminated();
            // A_22 A_22^{-1} = I on every cell box
            const int r = d.A_22().Height();
            for (int i = 0; i < r; ++i)
                for (int j = 0; j < r; ++j) {
                    double acc = 0.0;
                    for (int t = 0; t < r; ++t) acc += d.A_22().Get(i, t) * d.A_22_inv().Get(t, j);
                    inv_err = std::max(inv_err, std::fabs(acc - (i == j ? 1.0 : 0.0)));
                }
        }
        for (FactorData<double>& d : kd[l]) rf += d.NumDOFsEliminated();
        printf("data level %zu: cells %zu eliminated %d faces %zu eliminated %d\n", l, sd[l].size(), rc,
               kd[l].size(), rf);
    }
    printf("inv_err %.3g\n", inv_err);
    return 0;
}
