// Reverse-mode pieces of the adjoint against forward-mode dual numbers:
//  1. weno_side_grad on seeded stencils: smooth, kinked, discontinuous and
//     constant-plus-noise data, both branches of the indicator maximum;
//  2. face_flux_vjp (one reverse pass through the whole face flux) on seeded
//     faces (air at 300 K, velocities up to Mach 0.6, density and pressure
//     jumps across the face, slopes of either sign): all 50 partials of
//     face_flux (cell states, WENO states, six slopes; the collision time from
//     the left cell) contracted with a random adjoint flux, against one dual
//     pass per input.
// Prints the largest relative differences; exit status 1 if either is above its
// bound (argv[1] for the WENO part, argv[2] for the flux part).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "adj_flux.hpp"
#include "gkm_dual.hpp"

using gkm::Dual;

static double check_face_flux_vjp(int nfaces)
{
    std::mt19937_64 rng(777);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double gam = 1.4, R = 287.0, T0 = 300.0, c0 = std::sqrt(gam * R * T0);
    const gkm::Gas<double> g{gam, R, 1.0 / 0.7, 1.8e-3, T0, 1e-6, 2};
    const gkm::Gas<Dual> gd = gkm::dual_gas(g);
    const double dx = 2.0 * 3.141592653589793 / 64.0;
    double worst = 0.0;
    auto state = [&](double rho, double T, double Ma, double W[5]) {
        const double u = Ma * c0 * U(rng), v = Ma * c0 * U(rng), w = Ma * c0 * U(rng);
        W[0] = rho;
        W[1] = rho * u;
        W[2] = rho * v;
        W[3] = rho * w;
        W[4] = rho * R * T / (gam - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
    };
    for (int it = 0; it < nfaces; it++) {
        const double Ma = 0.6 * std::fabs(U(rng));
        const double jump = (it % 4 == 0) ? 0.3 : 0.01;  // every fourth face: a strong jump
        double WL[5], WR[5], Wl[5], Wr[5], D[6][5];
        state(1.0 + 0.1 * U(rng), T0 * (1.0 + 0.1 * U(rng)), Ma, WL);
        state(WL[0] * (1.0 + jump * U(rng)), T0 * (1.0 + jump * U(rng)), Ma, WR);
        for (int q = 0; q < 5; q++) {
            Wl[q] = WL[q] * (1.0 + 0.02 * U(rng));
            Wr[q] = WR[q] * (1.0 + 0.02 * U(rng));
        }
        for (int d = 0; d < 6; d++)
            for (int q = 0; q < 5; q++) D[d][q] = 0.2 * (std::fabs(WL[q]) + std::fabs(WR[q])) * U(rng) / dx;
        const double dt = 0.5 * dx / ((1.0 + Ma * 1.8) * c0);
        double x[50], Fbar[5], xbar[50];
        for (int q = 0; q < 5; q++) {
            x[q] = WL[q];
            x[5 + q] = WR[q];
            x[10 + q] = Wl[q];
            x[15 + q] = Wr[q];
            for (int d = 0; d < 6; d++) x[20 + 5 * d + q] = D[d][q];
        }
        for (int q = 0; q < 5; q++) Fbar[q] = U(rng);
        gkm::face_flux_vjp(x, dt, dx, g, Fbar, xbar);
        double xref[50], scale = 0.0;
        for (int k = 0; k < 50; k++) {
            Dual xd[50], F[5];
            for (int i = 0; i < 50; i++) xd[i] = Dual(x[i], i == k ? 1.0 : 0.0);
            gkm::face_flux<Dual>(xd, Dual(dt), Dual(dx), gd, F);
            double s = 0.0;
            for (int q = 0; q < 5; q++) s += Fbar[q] * F[q].d;
            xref[k] = s;
            // each partial times its input's size, so that inputs of different
            // magnitude (energy against density) weigh alike
            scale = std::fmax(scale, std::fabs(s * x[k]));
        }
        for (int k = 0; k < 50; k++)
            worst = std::fmax(worst, std::fabs((xbar[k] - xref[k]) * x[k]) / std::fmax(scale, 1e-300));
    }
    return worst;
}

int main(int argc, char** argv)
{
    const double bound = argc > 1 ? std::atof(argv[1]) : 1e-12;
    const double fbound = argc > 2 ? std::atof(argv[2]) : 1e-12;
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
    const int nfaces = 4000;
    const double fworst = check_face_flux_vjp(nfaces);
    std::printf("face_flux_vjp vs dual numbers: %d faces x 50 inputs, largest relative difference %.3e\n",
                nfaces, fworst);
    return worst <= bound && fworst <= fbound ? 0 : 1;
}
