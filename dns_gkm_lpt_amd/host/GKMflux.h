// GKMflux.h -- drop-in for the reference's src/GKMflux.h.  flux() evaluates the
// FP64 GKS face flux the gfx950 sweeps use, on the host, for one face; c2p is
// the conversion of GKMflux.cpp:302-309.
#ifndef GKMFLUXHEADERDEF
#define GKMFLUXHEADERDEF

#include "param.h"

void flux(ptype WL[5], ptype WR[5], ptype Wl[5], ptype Wr[5], ptype DWxl[5], ptype DWxr[5], ptype DWyl[5],
          ptype DWyr[5], ptype DWzl[5], ptype DWzr[5], ptype tau, ptype dt, ptype dx, ptype F[5]);
void c2p(ptype W[5], ptype &den, ptype U[3], ptype &P);

#endif
