// ref_shim.cpp -- C entry points into the reference's own CPU code, compiled by
// oracle/Makefile (target `ref`) from the sources under the reference's src/
// into oracle/_ref/libgkm_ref.so.
//
// TEST INFRASTRUCTURE ONLY: this binds the reference's GKMflux.cpp (flux,
// GKMflux.cpp:11-229), ApplyWENO (update.cpp:556-680) and evolveEulerian
// (update.cpp:206-516) so tests can pin the oracle's restatement against the
// reference itself.  It declares the reference functions it calls and sets the
// reference's globals (param.h:11-12, 27; update.cpp:6) before each call.
#include <cstdlib>
#include "param.h"

void flux(ptype WL[5], ptype WR[5], ptype Wl[5], ptype Wr[5], ptype DWxl[5], ptype DWxr[5],
          ptype DWyl[5], ptype DWyr[5], ptype DWzl[5], ptype DWzr[5], ptype tau, ptype dt, ptype dx, ptype F[5]);
void ApplyWENO(ptype WLLL[5], ptype WLL[5], ptype WL[5], ptype WR[5], ptype WRR[5], ptype WRRR[5],
               ptype Wl[5], ptype Wr[5]);
void evolveEulerian(ptype *W[5], ptype *Fx[5], ptype *Fy[5], ptype *Fz[5]);
extern ptype dt;  // update.cpp:6

extern "C" {

// GetParams (FlowField.cpp:9-58) with procDim 1x1x1
void ref_set_params(int n, double gam_, double mu0_, double p0_, double den0_, double dt_)
{
    procDim[0] = procDim[1] = procDim[2] = 1;
    nc = n;
    nt = nc + 6;
    nc_segx = nc_segy = nc_segz = nc;
    nt_segx = nt_segy = nt_segz = nc + 6;
    Nc = nc * nc * nc;
    Nt = nt * nt * nt;
    Nc_seg = Nc;
    Nt_seg = Nt;
    gam = gam_;
    K = int((5.0 - 3.0 * gam) / (gam - 1.0));
    den0 = den0_;
    p0 = p0_;
    dx = 2.0 * pi / nc;
    T0 = p0 / (R * den0);
    mu0 = mu0_;
    ::dt = dt_;
}

int ref_get_K(void) { return K; }
double ref_get_dx(void) { return dx; }
double ref_get_T0(void) { return T0; }

// in: n x 50 doubles (WL, WR, Wl, Wr, DWxl, DWxr, DWyl, DWyr, DWzl, DWzr)
void ref_flux_batch(long n, const double *in, const double *tau, double dt_, double dx_, double *F)
{
    for (long i = 0; i < n; i++) {
        double v[50];
        for (int q = 0; q < 50; q++) v[q] = in[50 * i + q];
        flux(v, v + 5, v + 10, v + 15, v + 20, v + 25, v + 30, v + 35, v + 40, v + 45, tau[i], dt_, dx_, F + 5 * i);
    }
}

// in: n x 30 doubles (cells i-2..i+3); out: n x 10 (Wl, Wr)
void ref_weno_batch(long n, const double *in, double *out)
{
    for (long i = 0; i < n; i++) {
        double s[30];
        for (int q = 0; q < 30; q++) s[q] = in[30 * i + q];
        ApplyWENO(s, s + 5, s + 10, s + 15, s + 20, s + 25, out + 10 * i, out + 10 * i + 5);
    }
}

// One evolveEulerian on a padded state (5 planes of nt^3, back to back), in place.
void ref_eulerian(double *Wflat)
{
    const long N = (long)Nt_seg;
    ptype *W[5], *Fx[5], *Fy[5], *Fz[5];
    for (int q = 0; q < 5; q++) {
        W[q] = Wflat + q * N;
        Fx[q] = (ptype *)calloc(N, sizeof(ptype));
        Fy[q] = (ptype *)calloc(N, sizeof(ptype));
        Fz[q] = (ptype *)calloc(N, sizeof(ptype));
    }
    evolveEulerian(W, Fx, Fy, Fz);
    for (int q = 0; q < 5; q++) {
        free(Fx[q]);
        free(Fy[q]);
        free(Fz[q]);
    }
}

}  // extern "C"
