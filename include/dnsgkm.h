/*
 * dnsgkm.h -- C ABI of the MI355X GKM-WENO step and Lagrangian particle tracker.
 *
 * Plain pointers and sizes only.  The caller owns every host buffer; the
 * library owns all device memory of a context.  The state stays resident on
 * the GPU between calls; data moves between host and device only in the
 * upload/download/sample calls.  Every function returns 0 on success and a
 * non-zero GKM_E* code on failure; gkm_last_error() then holds a message.
 * Calls are ordered on the context's stream; a context is not thread-safe.
 *
 * Host layout of a state (the reference's, param.h:40-41): 5 arrays (rho,
 * rho u, rho v, rho w, rho E) of nt^3 doubles, nt = nc + 6, cell (j,i,k) =
 * (y,x,z) at (j*nt + i)*nt + k with 3 ghost layers per face.
 * Particle positions use the reference's order pos = (y, x, z)
 * (ParticleTrack.cpp:52-55); velocities (u, v, w).
 */
#ifndef DNSGKM_H
#define DNSGKM_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GKM_OK = 0,
    GKM_EINVAL = 1,   /* bad argument or configuration                    */
    GKM_EHIP = 2,     /* a HIP runtime call failed                         */
    GKM_ENOMEM = 3,   /* device allocation failed                          */
    GKM_ESTATE = 4,   /* non-physical state (rho <= 0 or p <= 0 or NaN)    */
    GKM_ENOTREADY = 5 /* call order violated (e.g. step before upload)     */
};

enum { GKM_FP64 = 0, GKM_MIXED = 1 };

/* Run configuration.  gam, rho0, Mt, Re, mu0, p0, t0 are the reference's
 * params file (FlowField.cpp:16-40); cfl and prandtl are compile-time in the
 * reference (param.h:19-20) and run-time here. */
typedef struct {
    int nc;           /* interior cells per edge of the 2*pi periodic cube */
    int precision;    /* GKM_FP64 or GKM_MIXED                             */
    double gam, rho0, Mt, Re, mu0, p0, t0;
    double cfl;       /* CFL number of timestep()                          */
    double prandtl;   /* Pr of the slope solver (0.7 in the reference)     */
    int device;       /* HIP device ordinal                                */
    int check_state;  /* non-zero: gkm_cfl_dt reports GKM_ESTATE on a      */
                      /* non-physical state                                */
} gkm_config;

typedef struct gkm_ctx gkm_ctx;

/* Statistics of PrintStats (update.cpp:688-842) plus enstrophy <|curl u|^2>. */
typedef struct {
    double tke;       /* <rho |u|^2> (no 1/2, as the reference prints)     */
    double diss;      /* <mu(T) du_i/dx_j du_i/dx_j>                       */
    double sijsij;    /* <S_ij S_ij>                                       */
    double skew[3];   /* skewness of du/dx, dv/dy, dw/dz                   */
    double kurt[3];   /* kurtosis of du/dx, dv/dy, dw/dz                   */
    double prms, trms, vrms; /* normalised as update.cpp:829-833           */
    double pmean, tmean, vmean;
    double enstrophy; /* <|curl u|^2>                                      */
} gkm_stats;

/* Particle statistics of GetParticleStats (update.cpp:1111-1162). */
typedef struct {
    double tke, sijsij, skew_u, skew_w;
} gkm_pstats;

/* Replaces GetParams/segment/DevAlloc (FlowField.cpp:9-58, 275-284;
 * GPUAlloc.cu:3-132): validates cfg and allocates the resident state. */
int gkm_create(const gkm_config *cfg, gkm_ctx **out);
void gkm_destroy(gkm_ctx *ctx);
const char *gkm_last_error(const gkm_ctx *ctx);
/* internal degrees of freedom K = int((5-3gam)/(gam-1)) (FlowField.cpp:42) */
int gkm_get_K(const gkm_ctx *ctx);

/* Compensated accumulation of the update W^{n+1} = W^n - (dFx + dFy + dFz)/dx
 * (update.cpp:505-516; no reference counterpart): with it on, each cell keeps
 * the rounding error of its last FP64 update in a per-variable FP32 carry, adds
 * it to the next increment and stores the new error (TwoSum), so W + carry is
 * the exact running sum of the rounded increments.  On by default for
 * GKM_MIXED, off for GKM_FP64 (whose update is the reference's).  Uploading a
 * state clears the carry.  gkm_set_compensation fails on an ensemble member.
 * gkm_download_carry: the carry of the resident state, 5 x nc^3 floats,
 * unpadded, (y, x, z) with z fastest. */
int gkm_set_compensation(gkm_ctx *ctx, int on);
int gkm_get_compensation(const gkm_ctx *ctx);
int gkm_download_carry(gkm_ctx *ctx, float *out);

/* Host -> device state, padded layout, 5 separate arrays (Initialize_DHIT's
 * W[5], FlowField.cpp:60-114; HtoD of GPUupdate.cu:89-93). Ghosts are ignored. */
int gkm_upload_state(gkm_ctx *ctx, const double *const W[5]);
/* Device -> host state with periodic ghost layers filled on the device
 * (HaloTransfer, update.cpp:855-1029, single domain). */
int gkm_download_state(gkm_ctx *ctx, double *const W[5]);
/* Same, with the 5 arrays back to back in one buffer of 5*nt^3 doubles. */
int gkm_upload_state_packed(gkm_ctx *ctx, const double *W5);
int gkm_download_state_packed(gkm_ctx *ctx, double *W5);

/* One Eulerian step of size dt: derivsX/Y/Z + update + periodic halo
 * (evolveEulerian, update.cpp:206-212; GPU evolve, GPUupdate.cu:79-219). */
int gkm_step(gkm_ctx *ctx, double dt);
/* dt = cfl dx / max(|U| + c) (timestep, update.cpp:1031-1054). */
int gkm_cfl_dt(gkm_ctx *ctx, double *dt);
/* n steps with dt recomputed on the device from the CFL condition after each
 * step (the Evolve loop order, update.cpp:124-170, Eulerian part); no host
 * round trip inside.  If dt_hist is non-NULL it receives the n dts used. */
int gkm_run(gkm_ctx *ctx, int nsteps, double *dt_hist);
/* Flux differences of one sweep direction (dir 0 = x, 1 = y, 2 = z) of the
 * resident state with step dt: out (5 x nc^3, unpadded, (y,x,z) z fastest)
 * receives F(face i+1/2) - F(face i-1/2) per cell, i.e. one of the three
 * terms update() sums (update.cpp:505-516).  The state is not changed.
 * A per-direction parity hook for derivsX/Y/Z (update.cpp:214-503). */
int gkm_sweep_flux_diff(gkm_ctx *ctx, int dir, double dt, double *out);
/* Sweep tile shape per direction (x, y, z): variant 0..2 of
 * {faces along the sweep axis} x {tangential lines} per workgroup; results do
 * not depend on it.  No reference counterpart (the reference's CUDA path
 * fixes its block shapes in GPUupdate.cu:119-202). */
int gkm_set_sweep_tiles(gkm_ctx *ctx, const int tile[3]);
int gkm_get_sweep_tiles(const gkm_ctx *ctx, int tile[3]);
/* Time every tile variant of every sweep on the resident state (reps timed
 * launches each after one warm-up, the CFL dt of the state) and keep the
 * fastest per direction.  The resident state is not changed.  tile (may be
 * NULL) receives the choice, ms (may be NULL, 9 doubles) the average launch
 * time of variant v of direction d at [3 d + v]. */
int gkm_autotune(gkm_ctx *ctx, int reps, int tile[3], double ms[9]);
/* Wait for all work on the context's stream. */
int gkm_sync(gkm_ctx *ctx);
/* One Eulerian step of size dt on caller-owned device buffers, allocating no
 * device memory: W and Wout are device arrays of 5 x nc^3 doubles in the
 * library's device layout (rho, rho u, rho v, rho w, rho E planes, cell
 * (y, x, z) at (y nc + x) nc + z, no ghost layers), W != Wout; stream is a
 * hipStream_t (NULL: the null stream).  Results equal gkm_step's on a context
 * with the same configuration and default tiles (for GKM_MIXED, with
 * compensation off: this step does the plain FP64 update).  No reference
 * counterpart: for callers that cannot allocate (code under HIP graph
 * capture, serving loops).  Its scratch (the flux-difference accumulator and
 * the device dt) is one file-scope device array per GPU sized for nc <= 64;
 * a larger nc is refused with GKM_EINVAL.  Calls that overlap on one GPU race
 * on that array.  msg (may be NULL) receives an error message of at most
 * msglen bytes. */
int gkm_step_noalloc(const gkm_config *cfg, const double *W, double *Wout, double dt, void *stream, char *msg,
                     int msglen);
/* Flow statistics of the resident state (PrintStats, update.cpp:688-842). */
int gkm_stats_compute(gkm_ctx *ctx, gkm_stats *out);

/* Per-kernel timing with HIP events on the context's stream.  While enabled,
 * every sweep/update launch is bracketed by events; gkm_timing_read returns
 * accumulated milliseconds and launch counts for slots
 * 0 = x sweep, 1 = y sweep, 2 = z sweep (fused flux difference + update),
 * 3 = CFL reduction, 4 = spline prefilter, 5 = particle sample + Heun step,
 * 6 = particle cell sort. */
int gkm_timing_enable(gkm_ctx *ctx, int on);
int gkm_timing_read(gkm_ctx *ctx, double ms[8], long counts[8]);

/* Pointwise device functions on host arrays (parity checks of the kernels'
 * math).  gkm_flux_batch: flux() of GKMflux.cpp:11-229 on n faces; in holds
 * n x 50 doubles (WL, WR, Wl, Wr, DWxl, DWxr, DWyl, DWyr, DWzl, DWzr), tau
 * n doubles, F n x 5.  gkm_weno_batch: ApplyWENO of update.cpp:556-680 on n
 * stencils of 6 cells x 5 (cells i-2..i+3), out n x 10 (Wl, Wr).
 * precision: 0 = the FP64 device math, 1 = the mixed-precision device math
 * the GKM_MIXED sweeps use, 2 = all-FP32 device math (for error studies). */
int gkm_flux_batch(gkm_ctx *ctx, long n, const double *in, const double *tau, double dt, double dx,
                   int precision, double *F);
int gkm_weno_batch(gkm_ctx *ctx, long n, const double *in, int precision, double *out);

/* ---- Tangent-linear and discrete adjoint of the fixed-dt FP64 step --------
 * (no reference counterpart).  Both start from the resident state W_0, use the
 * fixed dt for nsteps steps of the same FP64 step gkm_step evaluates, and
 * leave the resident state at W_N.  Vectors are in the padded host layout
 * (5 x (nc+6)^3, like gkm_upload_state_packed): input ghost layers are
 * ignored, output ghost layers are filled periodically.
 * gkm_tl_run: dW_N = (dW_N/dW_0) dW_0.
 * gkm_adjoint_run: lam_0 = (dW_N/dW_0)^T lam_N, with the forward trajectory
 * checkpointed every `every` steps and each segment but the last recomputed on
 * the way back.  It is gkm_adjoint_forward followed by gkm_adjoint_backward.
 * gkm_adjoint_forward: the forward part alone; it leaves W_N resident, so a
 * caller whose lam_N depends on W_N (a gradient of J(W_N)) downloads W_N and
 * forms lam_N without a forward run of its own.
 * gkm_adjoint_backward: the backward part from the trajectory of the last
 * gkm_adjoint_forward, which it uses up (GKM_ENOTREADY without one); dt,
 * nsteps and every are the forward call's.
 * Contexts must be GKM_FP64. */
int gkm_tl_run(gkm_ctx *ctx, double dt, int nsteps, const double *dW0, double *dWN);
int gkm_adjoint_run(gkm_ctx *ctx, double dt, int nsteps, int every, const double *lamN, double *lam0);
int gkm_adjoint_forward(gkm_ctx *ctx, double dt, int nsteps, int every);
int gkm_adjoint_backward(gkm_ctx *ctx, const double *lamN, double *lam0);

/* ---- Ensemble of independent runs (no reference counterpart) -------------
 * n members, each a full context (own grid size, gas, initial condition,
 * CFL dt), stepped together: per step one launch computes every member's dt
 * and one launch per direction sweeps every member; with use_graph the step
 * is a captured HIP graph (one per buffer parity) replayed nsteps times.
 * Tracers uploaded to a member (lpt_upload on gkm_ensemble_member) advance
 * with it exactly as lpt_coupled_step advances them in a single context.
 * Members must share precision and device.  A member's state goes in and out
 * through gkm_upload_state / gkm_download_state / gkm_stats_compute on
 * gkm_ensemble_member(e, m) (owned by the ensemble; do not destroy it, and do
 * not step a member on its own between ensemble runs).  Each member's result
 * is its single run's to rounding. */
typedef struct gkm_ensemble gkm_ensemble;
int gkm_ensemble_create(int n, const gkm_config *cfgs, gkm_ensemble **out);
void gkm_ensemble_destroy(gkm_ensemble *e);
const char *gkm_ensemble_last_error(const gkm_ensemble *e);
gkm_ctx *gkm_ensemble_member(gkm_ensemble *e, int m);
/* one tile variant per direction for all members; gkm_ensemble_autotune times
 * every variant of every direction on the members' current states (which it
 * does not change) and keeps the fastest; ms (may be NULL) as gkm_autotune */
int gkm_ensemble_set_tiles(gkm_ensemble *e, const int tile[3]);
int gkm_ensemble_autotune(gkm_ensemble *e, int reps, int tile[3], double ms[9]);
/* nsteps steps of every member, each with dt from its own CFL condition */
int gkm_ensemble_run(gkm_ensemble *e, int nsteps, int use_graph);
/* on = 1: each step is one launch of a persistent kernel (as many workgroups
 * as stay resident; every phase -- CFL, x, y, z sweeps, tracer splines and
 * Heun -- takes its work items from the same tables, phases separated by grid
 * barriers with a bounded wait); results equal the per-phase launches' bit
 * for bit.  A barrier that timed out makes gkm_ensemble_run fail. */
int gkm_ensemble_set_persistent(gkm_ensemble *e, int on);
/* one persistent step (eager) with the wall-clock time of each phase, as
 * workgroup 0 saw the barriers: *nphases = 5 (CFL maxima + tracer velocities,
 * dt, x, y, z sweeps) or 9 (+ spline z, x, y passes, Heun) entries of ms */
int gkm_ensemble_persistent_profile(gkm_ensemble *e, double ms[9], int *nphases);
/* the dt of each member's last step (n doubles) */
int gkm_ensemble_dt(gkm_ensemble *e, double *dt);

/* ---- Lagrangian particle tracker (ParticleTrack.cpp, update.cpp:1056-1247) */
/* Upload np positions (np x 3, order y, x, z); replaces locateParticles
 * (update.cpp:1286-1322). Tags are 0..np-1 in upload order. */
int lpt_upload(gkm_ctx *ctx, long np, const double *pos);
/* Tracer ownership as the reference's single-rank run has it (call before
 * lpt_upload; default 0).  on = 1: a tracer whose block id int(pos / 2 pi)
 * (Particle::identifyBlock, ParticleTrack.cpp:21-40, procDim 1) is not 0 on
 * some axis -- a coordinate of exactly 2 pi at upload or after the wrap of a
 * Heun step -- is dropped, as locateParticles (update.cpp:1286-1322) and
 * evolveParticles (update.cpp:1056-1109) drop it: its position reads NaN, it
 * samples zero velocity and gradients (so it adds nothing to lpt_stats, which
 * still divides by the uploaded count, as GetParticleStats divides by
 * numparticles), lpt_properties gives it a row of NaN, and it is not
 * advanced.  on = 0: such a tracer is kept and sampled at the box face. */
int lpt_set_reference_drop(gkm_ctx *ctx, int on);
/* Tracers not dropped (the reference's per-rank `count`, update.cpp:1056-1109). */
int lpt_count(gkm_ctx *ctx, long *alive);
/* Download positions (np x 3) and, if non-NULL, velocities (np x 3) in tag order. */
int lpt_download(gkm_ctx *ctx, double *pos, double *vel);
/* Sample u, v, w and their gradients (over y, x, z) at the particles from the
 * natural tricubic spline of the resident state (getParticleProp velocity
 * part, ParticleTrack.cpp:99-113); any output pointer may be NULL. */
int lpt_sample(gkm_ctx *ctx, double *U, double *gU, double *gV, double *gW);
/* Particle statistics from the last sample (GetParticleStats). */
int lpt_stats(gkm_ctx *ctx, gkm_pstats *out);
/* Particle::getParticleProp (ParticleTrack.cpp:99-113) of every particle on
 * the resident state: splines of rho, u, v, w, p and of the Laplacians of u,
 * v, w (generateSplines, update.cpp:1200-1233), evaluated at the particle.
 * out: np x 41 doubles in upload order, the members of the reference's
 * Particle in declaration order (ParticleTrack.h:21-22): DEN, U[3], P,
 * gradD[3], gradU[3], gradV[3], gradW[3], gradP[3], lapU, lapV, lapW,
 * hessP[9], gradLapU[3], gradLapV[3], gradLapW[3]; vectors over (y, x, z),
 * hessP row-major over (y, x, z). */
#define LPT_PROP_WIDTH 41
int lpt_properties(gkm_ctx *ctx, double *out);
/* One coupled step of the Evolve loop (update.cpp:124-170) with the resident
 * state: sample U at x from the current field, advance the field by dt
 * (gkm_step), rebuild the velocity spline, Heun-advance the particles
 * (Particle::updatePosition, ParticleTrack.cpp:48-73). */
int lpt_coupled_step(gkm_ctx *ctx, double dt);
/* Particle-only Heun step against the current field, using the velocities
 * stored by the last lpt_sample (U at the start of the step). */
int lpt_heun(gkm_ctx *ctx, double dt);

#ifdef __cplusplus
}
#endif
#endif
