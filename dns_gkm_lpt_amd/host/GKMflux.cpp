// GKMflux.cpp -- flux() of the reference's GKMflux.cpp:11-229 evaluated by the
// gfx950 device flux through gkm_flux_batch; c2p of GKMflux.cpp:302-309.
#include "GKMflux.h"

#include "dnsgkm_host.h"

void flux(ptype WL[5], ptype WR[5], ptype Wl[5], ptype Wr[5], ptype DWxl[5], ptype DWxr[5], ptype DWyl[5],
          ptype DWyr[5], ptype DWzl[5], ptype DWzr[5], ptype tau, ptype dt, ptype dx, ptype F[5])
{
    double in[50];
    const ptype *parts[10] = {WL, WR, Wl, Wr, DWxl, DWxr, DWyl, DWyr, DWzl, DWzr};
    for (int v = 0; v < 10; v++)
        for (int q = 0; q < 5; q++) in[5 * v + q] = parts[v][q];
    dnsgkm_check(gkm_flux_batch(dnsgkm_context(), 1, in, &tau, dt, dx, GKM_FP64, F), "flux");
}

void c2p(ptype W[5], ptype &den, ptype U[3], ptype &P)
{
    den = W[0];
    U[0] = W[1] / den;
    U[1] = W[2] / den;
    U[2] = W[3] / den;
    P = (den * (gam - 1)) * (W[4] / den - 0.5 * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]));
}
