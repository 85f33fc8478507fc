// The drop-in flux() (dns_gkm_lpt_amd/host/GKMflux.cpp) over a batch of faces,
// for the CPU test against the oracle: in holds n x 50 doubles (WL, WR, Wl,
// Wr, DWxl, DWxr, DWyl, DWyr, DWzl, DWzr), tau n doubles, F n x 5.
#include "GKMflux.h"

extern "C" void host_flux_batch(long n, const double *in, const double *tau, double dt, double gam_, double mu0_,
                                double T0_, int nc_, double *F)
{
    gam = gam_;
    K = int((5.0 - 3.0 * gam) / (gam - 1.0));
    mu0 = mu0_;
    T0 = T0_;
    nc = nc_;
    dx = 2.0 * pi / nc;
    for (long f = 0; f < n; f++) {
        double v[10][5];
        for (int k = 0; k < 10; k++)
            for (int q = 0; q < 5; q++) v[k][q] = in[50 * f + 5 * k + q];
        flux(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], tau[f], dt, dx, F + 5 * f);
    }
}
