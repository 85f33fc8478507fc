// Host check of the blocked natural-spline line solve (lpt_device.hpp,
// line_solve_blocked) used by every x/y prefilter pass:
//  * the definition of the natural spline: at every knot
//    (c_i + 4 c_{i+1} + c_{i+2}) / 6 = f_i, and c_0 - 2 c_1 + c_2 = 0,
//    c_{M-1} - 2 c_M + c_{M+1} = 0 (zero second derivative at both ends);
//  * agreement with the plain two-sweep Thomas solve of the same system.
// Lines of 4 .. 1030 knots (fewer than one register block, exact multiples of
// it and ragged ends), strides 1 and 7, random and smooth data.
//   spline_line_check <residual tol> <sweep tol>    (relative to max |f|)
// Built with -DLINE_BLOCK=8 the warm-up spans 16 rows instead of 32 (its error
// 0.268^16 = 7e-10 of the line's values), which the check must reject.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "lpt_device.hpp"

#ifndef LINE_BLOCK
#define LINE_BLOCK lpt::kLineBlock
#endif

int main(int argc, char** argv)
{
    const double tol_res = argc > 1 ? atof(argv[1]) : 1e-14;
    const double tol_seq = argc > 2 ? atof(argv[2]) : 1e-15;
    const int Ms[] = {4, 5, 22, 33, 34, 35, 65, 66, 67, 98, 130, 131, 200, 518, 1030};
    unsigned long long s = 0x9e3779b97f4a7c15ull;
    auto rnd = [&]() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (s >> 11) * (1.0 / 9007199254740992.0) - 0.5; };
    double worst_res = 0.0, worst_seq = 0.0;
    int cases = 0;
    for (int M : Ms)
        for (int st : {1, 7})
            for (int kind = 0; kind < 2; kind++) {
                const int n = M - 2;
                std::vector<double> piv(n), f(M), line((size_t)(M + 2) * st, NAN);
                for (int r = 0; r < n; r++) piv[r] = 1.0 / (4.0 - (r > 0 ? piv[r - 1] : 0.0));
                double fmax = 0.0;
                for (int p = 0; p < M; p++) {
                    f[p] = kind == 0 ? rnd() : std::sin(0.05 * p) + 0.3 * std::cos(0.11 * p + 1.0);
                    line[(size_t)p * st] = f[p];
                    fmax = std::fmax(fmax, std::fabs(f[p]));
                }
                lpt::line_solve_blocked<LINE_BLOCK>(M, line.data(), st, piv.data());
                std::vector<double> c(M + 2);
                for (int p = 0; p < M + 2; p++) c[p] = line[(size_t)p * st];
                // definition of the natural spline
                double res = std::fabs(c[0] - 2 * c[1] + c[2]) + std::fabs(c[M - 1] - 2 * c[M] + c[M + 1]);
                for (int i = 0; i < M; i++) res = std::fmax(res, std::fabs((c[i] + 4 * c[i + 1] + c[i + 2]) / 6 - f[i]));
                // the plain two-sweep solve
                std::vector<double> dp(n), x(n), q(M + 2);
                double d = 0.0;
                for (int r = 0; r < n; r++) {
                    double rhs = 6.0 * f[r + 1];
                    if (r == 0) rhs -= f[0];
                    if (r == n - 1) rhs -= f[M - 1];
                    d = (rhs - d) * piv[r];
                    dp[r] = d;
                }
                x[n - 1] = dp[n - 1];
                for (int r = n - 2; r >= 0; r--) x[r] = dp[r] - piv[r] * x[r + 1];
                for (int r = 0; r < n; r++) q[r + 2] = x[r];
                q[1] = f[0];
                q[M] = f[M - 1];
                q[0] = 2 * f[0] - x[0];
                q[M + 1] = 2 * f[M - 1] - x[n - 1];
                double seq = 0.0;
                for (int p = 0; p < M + 2; p++) seq = std::fmax(seq, std::fabs(c[p] - q[p]));
                worst_res = std::fmax(worst_res, res / fmax);
                worst_seq = std::fmax(worst_seq, seq / fmax);
                cases++;
            }
    printf("cases %d  natural-spline residual %.3e  vs two-sweep solve %.3e\n", cases, worst_res, worst_seq);
    return (worst_res <= tol_res && worst_seq <= tol_seq) ? 0 : 1;
}
