// Host driver for the initial-guess helpers of the demag solvers
// (win_differences, win_accumulate, win_coefficients in tm_internal.hpp):
// feeds them a smooth sequence of right-hand sides and writes, per step,
// the sequence vector and the coefficients over the kept vectors, for a
// least-squares check in numpy.  Usage: win_ls_check OUT DT
#include "tm_internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using namespace tmh;

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const int n = 2000, modes = 7, steps = 14;
    const double dt = std::atof(argv[2]);
    std::mt19937_64 g(11);
    std::normal_distribution<double> N(0.0, 1.0);
    std::vector<double> V((size_t)modes * n), om(modes), ph(modes);
    for (auto& v : V) v = N(g);
    for (int k = 0; k < modes; ++k) {
        om[k] = 1.0 + 0.7 * k;
        ph[k] = N(g);
    }
    FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    const int hdr[3] = {n, steps, kSolveWindow};
    std::fwrite(hdr, sizeof(int), 3, f);
    std::vector<std::vector<double>> kept;  // newest first
    for (int st = 0; st < steps; ++st) {
        std::vector<double> b(n, 0.0);
        for (int k = 0; k < modes; ++k)
            for (int e = 0; e < n; ++e) b[e] += std::sin(om[k] * st * dt + ph[k]) * V[(size_t)k * n + e];
        const int cnt = (int)kept.size();
        double a[kSolveWindow] = {};
        if (cnt > 0) {
            double acc[kWinDots] = {};
            for (int e = 0; e < n; ++e) {
                double v[kSolveWindow] = {};
                for (int i = 0; i < cnt; ++i) v[i] = kept[i][e];
                const double r = b[e] - v[0];
                win_differences(v, cnt);
                win_accumulate(acc, v, r);
            }
            win_coefficients(cnt, acc, a);
        }
        std::fwrite(&cnt, sizeof(int), 1, f);
        std::fwrite(a, sizeof(double), kSolveWindow, f);
        std::fwrite(b.data(), sizeof(double), n, f);
        kept.insert(kept.begin(), b);
        if ((int)kept.size() > kSolveWindow) kept.pop_back();
    }
    std::fclose(f);
    return 0;
}
