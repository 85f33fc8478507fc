/*
 * gkm_oracle.h -- FP64 CPU restatement of the reference's GKM-WENO step and
 * Lagrangian particle tracker (nishantp9/DNS_GKM_LPT, src/).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (dns_gkm_lpt_amd/) never links, loads or calls it.
 *
 * Layout conventions (param.h:40-41 of the reference):
 *   padded edge nt = nc + 6, cell (j,i,k) = (y,x,z) at ((j*nt)+i)*nt+k, z fastest;
 *   a state is 5 SoA planes q = rho, rho*u, rho*v, rho*w, rho*E, each nt^3 doubles,
 *   stored back to back (plane q starts at q*nt^3).
 */
#ifndef GKM_ORACLE_H
#define GKM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nc;          /* interior cells per edge                                  */
    int K;           /* internal dof, int((5-3gam)/(gam-1))   FlowField.cpp:42   */
    double gam;      /* ratio of specific heats                                  */
    double R;        /* gas constant 287                      param.h:21         */
    double prn;      /* Prandtl number 0.7                    param.h:19         */
    double mu0;      /* reference viscosity                   FlowField.cpp:50   */
    double T0;       /* p0/(R rho0)                           FlowField.cpp:47   */
    double dx;       /* 2 pi / nc                             FlowField.cpp:46   */
    double eps;      /* WENO epsilon 1e-10                    param.h:18         */
} orc_gas;

/* --- pointwise pieces ---------------------------------------------------- */
void orc_c2p(const double W[5], double gam, double *den, double U[3], double *P);
void orc_weno(const double *S /* 6x5: cells i-2..i+3 */, const orc_gas *g,
              double Wl[5], double Wr[5]);
void orc_flux(const double WL[5], const double WR[5], const double Wl[5], const double Wr[5],
              const double DWxl[5], const double DWxr[5], const double DWyl[5],
              const double DWyr[5], const double DWzl[5], const double DWzr[5],
              double tau, double dt, double dx, const orc_gas *g, double F[5]);
void orc_flux_batch(long n, const double *in /* n x 50 */, const double *tau,
                    double dt, const orc_gas *g, double *F /* n x 5 */);
void orc_weno_batch(long n, const double *in /* n x 30 */, const orc_gas *g,
                    double *out /* n x 10 */);
double orc_tau(const double WL[5], const orc_gas *g);

/* --- whole-field pieces (padded layout) ---------------------------------- */
void orc_halo_fill(int nc, double *W);
void orc_fluxes(int nc, const double *W, double dt, const orc_gas *g, int nthreads,
                double *Fx, double *Fy, double *Fz);
void orc_update(int nc, double *W, const double *Fx, const double *Fy, const double *Fz,
                double dx);
void orc_eulerian_step(int nc, double *W, double dt, const orc_gas *g, int nthreads);
double orc_cfl_dt(int nc, const double *W, double cfl, const orc_gas *g);

/* --- natural tricubic B-spline on the padded cell-centre grid ------------ */
/* coefs: (nt+2)^3 doubles, index ((a*(nt+2))+b)*(nt+2)+c over (y,x,z).        */
void orc_spline_coefs(int nc, const double *f, double *coefs);
void orc_spline_eval(int nc, const double *coefs, double dx, double y, double x, double z,
                     double *val, double grad[3], double hess[9]);

/* --- particles ----------------------------------------------------------- */
void orc_velocity_coefs(int nc, const double *W, double *cu, double *cv, double *cw);
void orc_particle_sample(int nc, const double *cu, const double *cv, const double *cw,
                         double dx, long np, const double *pos, double *U, double *gU,
                         double *gV, double *gW);
void orc_particle_heun(int nc, const double *cu, const double *cv, const double *cw,
                       double dx, long np, double *pos, const double *U, double dt);
void orc_lap_field(int nc, const double *coefs, double dx, double *lap);
void orc_particle_props(int nc, const double *W, double gam, double dx, long np, const double *pos,
                        double *out);

/* --- diagnostics --------------------------------------------------------- */
typedef struct {
    double tke, diss, sijsij, skew[3], kurt[3], prms, trms, vrms;
    double pmean, tmean, vmean, ens;
} orc_stats_t;
void orc_field_stats(int nc, const double *W, const double *cu, const double *cv,
                     const double *cw, const orc_gas *g, double p0, double Mt, double den0,
                     orc_stats_t *out);
typedef struct { double tke, sijsij, skew_u, skew_w; } orc_pstats_t;
void orc_particle_stats(long np, long np_total, const double *U, const double *gU,
                        const double *gV, const double *gW, orc_pstats_t *out);

#ifdef __cplusplus
}
#endif
#endif
