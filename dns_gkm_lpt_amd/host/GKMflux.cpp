// GKMflux.cpp -- flux() of the reference's GKMflux.cpp:11-229, evaluated on the
// host by the same rest-frame GKS flux template the gfx950 sweeps run
// (csrc/gkm_flux.hpp, FP64 instance), with the gas of the param.h globals
// (K and gam as the reference reads them); c2p of GKMflux.cpp:302-309.
//
// gkm_flux.hpp comes first: param.h defines pi, R and eps as macros, which
// must not reach the template's identifiers.
#include "../csrc/gkm_flux.hpp"

#include "GKMflux.h"

void flux(ptype WL[5], ptype WR[5], ptype Wl[5], ptype Wr[5], ptype DWxl[5], ptype DWxr[5], ptype DWyl[5],
          ptype DWyr[5], ptype DWzl[5], ptype DWzr[5], ptype tau, ptype dt, ptype dx, ptype F[5])
{
    const double D[6][5] = {
        {DWxl[0], DWxl[1], DWxl[2], DWxl[3], DWxl[4]}, {DWxr[0], DWxr[1], DWxr[2], DWxr[3], DWxr[4]},
        {DWyl[0], DWyl[1], DWyl[2], DWyl[3], DWyl[4]}, {DWyr[0], DWyr[1], DWyr[2], DWyr[3], DWyr[4]},
        {DWzl[0], DWzl[1], DWzl[2], DWzl[3], DWzl[4]}, {DWzr[0], DWzr[1], DWzr[2], DWzr[3], DWzr[4]}};
    // the flux reads K, gam and 1/Pr; mu0, T0 and eps enter only the collision time and WENO
    const gkm::Gas<double> g{gam, R, 1.0 / PRN, mu0, T0, eps, K};
    gkm::gks_flux_tf<double, double>(WL, WR, Wl, Wr, D, tau, dt, dx, g, F);
}

void c2p(ptype W[5], ptype &den, ptype U[3], ptype &P)
{
    den = W[0];
    U[0] = W[1] / den;
    U[1] = W[2] / den;
    U[2] = W[3] / den;
    P = (den * (gam - 1)) * (W[4] / den - 0.5 * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]));
}
