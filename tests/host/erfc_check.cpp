// half_erfc (csrc/gkm_math.hpp) -- 0.5 erfc(x) from exp(-x^2) and a fitted
// erfcx -- against the C library's erfc: every x of a dense grid on [-12, 12]
// plus seeded random points, and the values the flux feeds it (z = sqrt(lambda) u
// of air at 200-400 K up to Mach 3).  Prints the largest absolute and relative
// differences; exit status 1 if the absolute one is above argv[1] (default 1e-15).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "gkm_math.hpp"

int main(int argc, char** argv)
{
    const double bound = argc > 1 ? std::atof(argv[1]) : 1e-15;
    double worst_abs = 0.0, worst_rel = 0.0, at = 0.0;
    long n = 0;
    auto check = [&](double x) {
        const double v = gkm::half_erfc<double>(x, std::exp(-x * x));
        const double r = 0.5 * std::erfc(x);
        const double a = std::fabs(v - r);
        if (a > worst_abs) {
            worst_abs = a;
            at = x;
        }
        if (r > 1e-300) worst_rel = std::fmax(worst_rel, a / r);
        n++;
    };
    for (int i = 0; i <= 2400000; i++) check(-12.0 + 1e-5 * i);
    std::mt19937_64 rng(4242);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int i = 0; i < 1000000; i++) check(6.0 * U(rng));
    // z = sqrt(lambda) u, lambda = 1/(2 R T), |u| up to Mach 3
    const double R = 287.0, gam = 1.4;
    for (int i = 0; i < 1000000; i++) {
        const double T = 300.0 + 100.0 * U(rng);
        const double u = 3.0 * std::sqrt(gam * R * T) * U(rng);
        check(std::sqrt(1.0 / (2.0 * R * T)) * u);
    }
    std::printf("half_erfc vs 0.5 erfc: %ld points, largest absolute difference %.3e (at x = %.6f), "
                "largest relative %.3e\n",
                n, worst_abs, at, worst_rel);
    return worst_abs <= bound ? 0 : 1;
}
