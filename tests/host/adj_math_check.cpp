// weno_side_grad (reverse mode) against forward-mode dual numbers on seeded
// stencils: smooth, kinked, discontinuous and constant-plus-noise data, both
// branches of the indicator maximum.  Prints the largest relative difference
// of the six partial derivatives; exit status 1 above the bound in argv[1].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "adj_math.hpp"
#include "gkm_dual.hpp"

using gkm::Dual;

int main(int argc, char** argv)
{
    const double bound = argc > 1 ? std::atof(argv[1]) : 1e-12;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double eps = 1e-6;
    double worst = 0.0, worst_kind[4] = {0, 0, 0, 0};
    long n = 0;
    for (int kind = 0; kind < 4; kind++)
        for (int it = 0; it < 20000; it++) {
            double x[6];
            for (int k = 0; k < 6; k++) {
                const double s = 0.3 * k;
                switch (kind) {
                case 0: x[k] = 1.0 + std::sin(s + U(rng)) + 0.01 * U(rng); break;  // smooth
                case 1: x[k] = std::fabs(s - 0.75 + 0.2 * U(rng)) + 0.1 * U(rng); break;  // kink
                case 2: x[k] = (k < 3 + (it % 3) - 1 ? 1.0 : 0.2) + 0.05 * U(rng); break;  // jump
                default: x[k] = 2.0 + 1e-4 * U(rng); break;  // nearly constant
                }
            }
            const double wbar = 1.0 + U(rng);
            double g[6] = {0, 0, 0, 0, 0, 0};
            gkm::weno_side_grad(x[0], x[1], x[2], x[3], x[4], x[5], eps, wbar, g);
            double scale = 0.0, dref[6];
            for (int k = 0; k < 6; k++) {
                Dual d[6];
                for (int j = 0; j < 6; j++) d[j] = Dual(x[j], j == k ? 1.0 : 0.0);
                dref[k] = wbar * gkm::weno_side<Dual>(d[0], d[1], d[2], d[3], d[4], d[5], Dual(eps)).d;
                scale = std::fmax(scale, std::fabs(dref[k]));
            }
            for (int k = 0; k < 6; k++) {
                const double r = std::fabs(g[k] - dref[k]) / std::fmax(scale, 1e-300);
                worst = std::fmax(worst, r);
                worst_kind[kind] = std::fmax(worst_kind[kind], r);
            }
            n++;
        }
    std::printf("weno_side_grad vs dual numbers: %ld stencils, largest relative difference %.3e "
                "(smooth %.1e, kink %.1e, jump %.1e, nearly constant %.1e)\n",
                n, worst, worst_kind[0], worst_kind[1], worst_kind[2], worst_kind[3]);
    return worst <= bound ? 0 : 1;
}
