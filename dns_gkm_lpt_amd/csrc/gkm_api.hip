// gkm_api.hip -- C ABI over the gfx950 kernels (declared in include/dnsgkm.h).
//
// A context owns one HIP stream (default priority) and all device memory of a
// run: the resident state (double-buffered), the flux-difference accumulator,
// a padded staging buffer for host transfers, the device dt scalar and the
// reduction scratch.  Every HIP call is checked; failures set the context's
// last-error string and return a GKM_E* code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/dnsgkm.h"
#include "adj_kernels.hpp"
#include "gkm_config.hpp"
#include "ens_persist.hpp"
#include "gkm_kernels.hpp"
#include "lpt_kernels.hpp"

struct gkm_ctx {
    gkm_config cfg{};
    int nc = 0;
    long plane = 0;     // nc^3
    long padded = 0;    // (nc+6)^3
    double dx = 0.0;
    gkm::Gas<double> gas{};
    hipStream_t stream = nullptr;
    double* W[2] = {nullptr, nullptr};
    int cur = 0;
    double* dW = nullptr;
    // compensated accumulation of the update: carried rounding error per cell
    // and variable (5 x plane floats) of W[0] and W[1]; on by default for
    // GKM_MIXED (allocated at create), off for GKM_FP64
    float* cmp[2] = {nullptr, nullptr};
    bool compensate = false;
    bool member = false;           // owned by an ensemble (its tables hold the carry pointers)
    double* stage = nullptr;       // 5 * padded
    double* dscal = nullptr;       // [0] dt, [1] vmax
    int* dflag = nullptr;          // non-physical-state flag
    double* partial = nullptr;     // CFL block maxima
    bool have_state = false;
    int tile[3] = {0, 0, 0};       // sweep tile variant per direction
    int slots = 512;               // concurrent sweep workgroups: 2 per CU
    std::string err;
    // per-kernel timing
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double tms[8] = {0};
    long tcount[8] = {0};
    // particles
    lpt::State lpt;
    // tangent-linear / adjoint (allocated on first use)
    double* adj_scratch = nullptr;
    double* adj_vec[2] = {nullptr, nullptr};  // tangent or adjoint vectors, 5 x plane each
    std::vector<double*> adj_states;          // checkpoints and segment states
    // the trajectory recorded by the last gkm_adjoint_forward (adj_n < 0: none)
    int adj_n = -1, adj_every = 1;
    double adj_dt = 0.0;
};

namespace {

int fail(gkm_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                            \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return fail((ctx), GKM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

// run f() bracketed by timing events when enabled; accumulates into slot
template <typename Fn>
int timed(gkm_ctx* c, int slot, Fn f)
{
    if (!c->timing) {
        hipError_t e = f();
        if (e != hipSuccess) return fail(c, GKM_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
        return GKM_OK;
    }
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    hipError_t e = f();
    if (e != hipSuccess) return fail(c, GKM_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->tms[slot] += ms;
    c->tcount[slot] += 1;
    return GKM_OK;
}

gkm::SweepArgs sweep_args(const gkm_ctx* c, int dir, int mode)
{
    gkm::SweepArgs a{};
    a.W = c->W[c->cur];
    a.dW = c->dW;
    a.Wout = c->W[1 - c->cur];
    a.cmp_in = c->compensate ? c->cmp[c->cur] : nullptr;
    a.cmp_out = c->compensate ? c->cmp[1 - c->cur] : nullptr;
    a.dt = c->dscal;
    a.plane = c->plane;
    a.nc = c->nc;
    a.dx = c->dx;
    a.inv_dx = 1.0 / c->dx;
    a.g = c->gas;
    a.tile = c->tile[dir];
    a.seg = gkm::sweep_seg_len(dir, c->nc, a.tile, c->slots);
    a.nseg = gkm::sweep_nseg(c->nc, a.seg);
    a.mode = mode;
    return a;
}

int eulerian_step(gkm_ctx* c)
{
    for (int dir = 0; dir < 3; dir++) {
        // x: dW = dF, y: dW += dF, z: Wout = W - (dW + dF)/dx
        const gkm::SweepArgs a = sweep_args(c, dir, dir);
        int rc = timed(c, dir, [&] { return gkm::launch_sweep(dir, a, c->cfg.precision == GKM_MIXED, c->stream); });
        if (rc) return rc;
    }
    c->cur = 1 - c->cur;
    c->lpt.valid = false;  // the resident spline described the previous state
    return GKM_OK;
}

int cfl_on_device(gkm_ctx* c)
{
    return timed(c, 3, [&] {
        return gkm::launch_cfl(c->W[c->cur], c->plane, c->gas.gam, c->cfg.cfl, c->dx, c->partial, c->dscal,
                               c->dscal + 1, c->dflag, c->stream);
    });
}

}  // namespace

extern "C" {

int gkm_create(const gkm_config* cfg, gkm_ctx** out)
{
    if (!cfg || !out) return GKM_EINVAL;
    *out = nullptr;
    gkm_ctx* c = new (std::nothrow) gkm_ctx();
    if (!c) return GKM_ENOMEM;
    c->cfg = *cfg;
    const int nc = cfg->nc;
    auto bad = [&](const std::string& m) {
        // keep the message reachable: return the context so last_error works
        c->err = m;
        *out = c;
        return GKM_EINVAL;
    };
    {
        const std::string m = gkm::config_error(*cfg, gkm::sweep_grid_ok);
        if (!m.empty()) return bad(m);
    }
    c->nc = nc;
    c->plane = (long)nc * nc * nc;
    c->padded = (long)(nc + 6) * (nc + 6) * (nc + 6);
    c->dx = gkm::config_dx(*cfg);
    c->gas = gkm::config_gas(*cfg);
    for (int d = 0; d < 3; d++) c->tile[d] = gkm::sweep_first_tile(d, nc);

    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) return bad(std::string("hipSetDevice: ") + hipGetErrorString(e));
    *out = c;
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {
        int cus = 0;
        HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, cfg->device));
        c->slots = 2 * (cus > 0 ? cus : 1);  // the sweep kernels run two workgroups per CU
    }
    HIPCHK(c, hipEventCreate(&c->ev[0]));
    HIPCHK(c, hipEventCreate(&c->ev[1]));
    const size_t sbytes = sizeof(double) * 5 * (size_t)c->plane;
    for (int b = 0; b < 2; b++)
        if (hipMalloc(&c->W[b], sbytes) != hipSuccess) return fail(c, GKM_ENOMEM, "hipMalloc state");
    if (hipMalloc(&c->dW, sbytes) != hipSuccess) return fail(c, GKM_ENOMEM, "hipMalloc dW");
    if (cfg->precision == GKM_MIXED) {
        const size_t cbytes = sizeof(float) * 5 * (size_t)gkm::carry_plane(c->nc);
        for (int b = 0; b < 2; b++) {
            if (hipMalloc(&c->cmp[b], cbytes) != hipSuccess) return fail(c, GKM_ENOMEM, "hipMalloc carry");
            HIPCHK(c, hipMemsetAsync(c->cmp[b], 0, cbytes, c->stream));
        }
        c->compensate = true;
    }
    if (hipMalloc(&c->stage, sizeof(double) * 5 * (size_t)c->padded) != hipSuccess)
        return fail(c, GKM_ENOMEM, "hipMalloc staging");
    if (hipMalloc(&c->dscal, sizeof(double) * 4) != hipSuccess) return fail(c, GKM_ENOMEM, "hipMalloc scalars");
    if (hipMalloc(&c->dflag, sizeof(int) * 4) != hipSuccess) return fail(c, GKM_ENOMEM, "hipMalloc flag");
    if (hipMalloc(&c->partial, sizeof(double) * gkm::kCflBlocks) != hipSuccess)
        return fail(c, GKM_ENOMEM, "hipMalloc partial");
    return GKM_OK;
}

void gkm_destroy(gkm_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    lpt::release(c->lpt);
    (void)hipFree(c->adj_scratch);
    (void)hipFree(c->adj_vec[0]);
    (void)hipFree(c->adj_vec[1]);
    for (double* p : c->adj_states) (void)hipFree(p);
    (void)hipFree(c->W[0]);
    (void)hipFree(c->W[1]);
    (void)hipFree(c->dW);
    (void)hipFree(c->cmp[0]);
    (void)hipFree(c->cmp[1]);
    (void)hipFree(c->stage);
    (void)hipFree(c->dscal);
    (void)hipFree(c->dflag);
    (void)hipFree(c->partial);
    if (c->ev[0]) (void)hipEventDestroy(c->ev[0]);
    if (c->ev[1]) (void)hipEventDestroy(c->ev[1]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* gkm_last_error(const gkm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gkm_get_K(const gkm_ctx* c) { return c ? c->gas.K : -1; }

int gkm_set_compensation(gkm_ctx* c, int on)
{
    if (!c) return GKM_EINVAL;
    if (c->member) return fail(c, GKM_EINVAL, "an ensemble member keeps its precision's compensation setting");
    if (on && !c->cmp[0]) {
        const size_t cbytes = sizeof(float) * 5 * (size_t)gkm::carry_plane(c->nc);
        for (int b = 0; b < 2; b++)
            if (hipMalloc(&c->cmp[b], cbytes) != hipSuccess) return fail(c, GKM_ENOMEM, "hipMalloc carry");
    }
    if (on && !c->compensate) {
        // the resident state starts with no carried error
        HIPCHK(c, hipMemsetAsync(c->cmp[c->cur], 0, sizeof(float) * 5 * (size_t)gkm::carry_plane(c->nc), c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->compensate = on != 0;
    return GKM_OK;
}

int gkm_get_compensation(const gkm_ctx* c) { return c ? (int)c->compensate : -1; }

int gkm_download_carry(gkm_ctx* c, float* out)
{
    if (!c || !out) return fail(c, GKM_EINVAL, "null argument");
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    if (!c->compensate) return fail(c, GKM_EINVAL, "compensation is off");
    // the device keeps the z sweep's blocked layout (gkm::carry_index); out is
    // in the state's (y, x, z) order
    const int nc = c->nc;
    const long cp = gkm::carry_plane(nc);
    std::vector<float> blk((size_t)5 * cp);
    HIPCHK(c, hipMemcpyAsync(blk.data(), c->cmp[c->cur], sizeof(float) * blk.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int q = 0; q < 5; q++)
        for (int y = 0; y < nc; y++)
            for (int x = 0; x < nc; x++)
                for (int z = 0; z < nc; z++)
                    out[q * c->plane + ((long)y * nc + x) * nc + z] = blk[q * cp + gkm::carry_index(y, x, z, nc)];
    return GKM_OK;
}

int gkm_upload_state(gkm_ctx* c, const double* const W[5])
{
    if (!c || !W) return fail(c, GKM_EINVAL, "null argument");
    for (int q = 0; q < 5; q++) {
        if (!W[q]) return fail(c, GKM_EINVAL, "null state plane");
        HIPCHK(c, hipMemcpyAsync(c->stage + q * c->padded, W[q], sizeof(double) * c->padded, hipMemcpyHostToDevice,
                                 c->stream));
    }
    int rc = timed(c, 7, [&] { return gkm::launch_unpack(c->stage, c->W[c->cur], c->nc, c->stream); });
    if (rc) return rc;
    if (c->cmp[c->cur])  // a new state carries no rounding error
        HIPCHK(c, hipMemsetAsync(c->cmp[c->cur], 0, sizeof(float) * 5 * (size_t)gkm::carry_plane(c->nc), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_state = true;
    c->lpt.valid = false;  // a spline built earlier described the previous state
    return GKM_OK;
}

int gkm_upload_state_packed(gkm_ctx* c, const double* W5)
{
    if (!c || !W5) return fail(c, GKM_EINVAL, "null argument");
    const double* W[5];
    for (int q = 0; q < 5; q++) W[q] = W5 + q * c->padded;
    return gkm_upload_state(c, W);
}

int gkm_download_state(gkm_ctx* c, double* const W[5])
{
    if (!c || !W) return fail(c, GKM_EINVAL, "null argument");
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    int rc = timed(c, 7, [&] { return gkm::launch_pack(c->W[c->cur], c->stage, c->nc, c->stream); });
    if (rc) return rc;
    for (int q = 0; q < 5; q++) {
        if (!W[q]) return fail(c, GKM_EINVAL, "null state plane");
        HIPCHK(c, hipMemcpyAsync(W[q], c->stage + q * c->padded, sizeof(double) * c->padded, hipMemcpyDeviceToHost,
                                 c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GKM_OK;
}

int gkm_download_state_packed(gkm_ctx* c, double* W5)
{
    if (!c || !W5) return fail(c, GKM_EINVAL, "null argument");
    double* W[5];
    for (int q = 0; q < 5; q++) W[q] = W5 + q * c->padded;
    return gkm_download_state(c, W);
}

int gkm_step(gkm_ctx* c, double dt)
{
    if (!c) return GKM_EINVAL;
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, GKM_EINVAL, "dt must be positive and finite");
    HIPCHK(c, hipMemcpyAsync(c->dscal, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // dt lives on the host stack
    return eulerian_step(c);
}

int gkm_cfl_dt(gkm_ctx* c, double* dt)
{
    if (!c || !dt) return fail(c, GKM_EINVAL, "null argument");
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    int rc = cfl_on_device(c);
    if (rc) return rc;
    double s[2];
    int flag = 0;
    HIPCHK(c, hipMemcpyAsync(s, c->dscal, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&flag, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (flag && c->cfg.check_state) return fail(c, GKM_ESTATE, "non-physical state (rho <= 0, p <= 0 or NaN)");
    *dt = s[0];
    return GKM_OK;
}

int gkm_run(gkm_ctx* c, int nsteps, double* dt_hist)
{
    if (!c || nsteps < 0) return fail(c, GKM_EINVAL, "bad argument");
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    for (int n = 0; n < nsteps; n++) {
        int rc = cfl_on_device(c);  // dt of this step from W^n (update.cpp:115, 167)
        if (rc) return rc;
        if (dt_hist) HIPCHK(c, hipMemcpyAsync(dt_hist + n, c->dscal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        rc = eulerian_step(c);
        if (rc) return rc;
    }
    // the flag so far describes W^0 .. W^{n-1}; check the state the last step produced
    if (nsteps > 0) {
        int rc = cfl_on_device(c);
        if (rc) return rc;
    }
    int flag = 0;
    HIPCHK(c, hipMemcpyAsync(&flag, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (flag && c->cfg.check_state) return fail(c, GKM_ESTATE, "non-physical state (rho <= 0, p <= 0 or NaN)");
    return GKM_OK;
}

int gkm_sweep_flux_diff(gkm_ctx* c, int dir, double dt, double* out)
{
    if (!c || !out || dir < 0 || dir > 2) return fail(c, GKM_EINVAL, "bad argument");
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, GKM_EINVAL, "dt must be positive and finite");
    HIPCHK(c, hipMemcpyAsync(c->dscal, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const gkm::SweepArgs a = sweep_args(c, dir, 0);
    int rc = timed(c, dir, [&] { return gkm::launch_sweep(dir, a, c->cfg.precision == GKM_MIXED, c->stream); });
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->dW, sizeof(double) * 5 * c->plane, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GKM_OK;
}

// ---- tangent-linear and adjoint of the fixed-dt FP64 step (no reference counterpart)
namespace {
int adj_prepare(gkm_ctx* c, double dt, int nstates)
{
    if (c->cfg.precision != GKM_FP64) return fail(c, GKM_EINVAL, "tangent-linear / adjoint runs need GKM_FP64");
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, GKM_EINVAL, "dt must be positive and finite");
    const size_t sb = sizeof(double) * 5 * (size_t)c->plane;
    if (!c->adj_scratch &&
        hipMalloc(&c->adj_scratch, sizeof(double) * gkm::adj_scratch_doubles(c->plane)) != hipSuccess)
        return fail(c, GKM_ENOMEM, "hipMalloc adjoint scratch");
    for (auto& v : c->adj_vec)
        if (!v && hipMalloc(&v, sb) != hipSuccess) return fail(c, GKM_ENOMEM, "hipMalloc adjoint vectors");
    while ((int)c->adj_states.size() < nstates) {
        double* p = nullptr;
        if (hipMalloc(&p, sb) != hipSuccess) return fail(c, GKM_ENOMEM, "hipMalloc checkpoints");
        c->adj_states.push_back(p);
    }
    HIPCHK(c, hipMemcpyAsync(c->dscal, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // dt lives on the host stack
    return GKM_OK;
}

int vec_upload(gkm_ctx* c, const double* host_padded, double* dev)
{
    HIPCHK(c, hipMemcpyAsync(c->stage, host_padded, sizeof(double) * 5 * c->padded, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, gkm::launch_unpack(c->stage, dev, c->nc, c->stream));
    return GKM_OK;
}

int vec_download(gkm_ctx* c, const double* dev, double* host_padded)
{
    HIPCHK(c, gkm::launch_pack(dev, c->stage, c->nc, c->stream));
    HIPCHK(c, hipMemcpyAsync(host_padded, c->stage, sizeof(double) * 5 * c->padded, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GKM_OK;
}
}  // namespace

int gkm_tl_run(gkm_ctx* c, double dt, int nsteps, const double* dW0, double* dWN)
{
    if (!c || !dW0 || !dWN || nsteps < 0) return fail(c, GKM_EINVAL, "bad argument");
    int rc = adj_prepare(c, dt, 0);
    if (rc) return rc;
    gkm::AdjScratch s{};
    gkm::adj_scratch_carve(c->adj_scratch, c->plane, s);
    rc = vec_upload(c, dW0, c->adj_vec[0]);
    if (rc) return rc;
    int t = 0;
    for (int n = 0; n < nsteps; n++) {
        HIPCHK(c, gkm::launch_tl_step(c->W[c->cur], c->adj_vec[t], c->plane, c->nc, c->gas, dt, c->dx, s,
                                      c->adj_vec[1 - t], c->stream));
        t = 1 - t;
        rc = eulerian_step(c);
        if (rc) return rc;
    }
    return vec_download(c, c->adj_vec[t], dWN);
}

// The forward trajectory of the adjoint: checkpoints ckpt[k] = W_{k every}, the
// states W_{s0}..W_{s1-1} of one segment in seg[], and W_N in final_state.
// The forward pass stores the last segment's states as it goes, so the
// backward pass recomputes only the segments before it.
int gkm_adjoint_forward(gkm_ctx* c, double dt, int nsteps, int every)
{
    if (!c || nsteps < 0 || every < 1) return fail(c, GKM_EINVAL, "bad argument");
    c->adj_n = -1;
    const int nseg = (nsteps + every - 1) / every;
    // checkpoints, the states of one segment, and the final state
    int rc = adj_prepare(c, dt, nseg + every + 1);
    if (rc) return rc;
    const size_t sb = sizeof(double) * 5 * (size_t)c->plane;
    double** ckpt = c->adj_states.data();
    double** seg = ckpt + nseg;
    double* final_state = ckpt[nseg + every];
    const int last0 = (nseg - 1) * every;  // first step of the last segment
    for (int n = 0; n < nsteps; n++) {
        if (n % every == 0) HIPCHK(c, hipMemcpyAsync(ckpt[n / every], c->W[c->cur], sb, hipMemcpyDeviceToDevice, c->stream));
        if (n >= last0) HIPCHK(c, hipMemcpyAsync(seg[n - last0], c->W[c->cur], sb, hipMemcpyDeviceToDevice, c->stream));
        rc = eulerian_step(c);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemcpyAsync(final_state, c->W[c->cur], sb, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->adj_n = nsteps;
    c->adj_every = every;
    c->adj_dt = dt;
    return GKM_OK;
}

int gkm_adjoint_backward(gkm_ctx* c, const double* lamN, double* lam0)
{
    if (!c || !lamN || !lam0) return fail(c, GKM_EINVAL, "bad argument");
    if (c->adj_n < 0) return fail(c, GKM_ENOTREADY, "no forward trajectory (gkm_adjoint_forward)");
    const int nsteps = c->adj_n, every = c->adj_every;
    const double dt = c->adj_dt;
    const int nseg = (nsteps + every - 1) / every;
    int rc = adj_prepare(c, dt, nseg + every + 1);  // restores the fixed dt of the trajectory
    if (rc) return rc;
    gkm::AdjScratch s{};
    gkm::adj_scratch_carve(c->adj_scratch, c->plane, s);
    const size_t sb = sizeof(double) * 5 * (size_t)c->plane;
    double** ckpt = c->adj_states.data();
    double** seg = ckpt + nseg;
    double* final_state = ckpt[nseg + every];
    rc = vec_upload(c, lamN, c->adj_vec[0]);
    if (rc) return rc;
    int l = 0;
    for (int k = nseg - 1; k >= 0; k--) {
        const int s0 = k * every, s1 = (s0 + every < nsteps) ? s0 + every : nsteps;
        if (k < nseg - 1) {  // the last segment's states are still there from the forward pass
            HIPCHK(c, hipMemcpyAsync(c->W[c->cur], ckpt[k], sb, hipMemcpyDeviceToDevice, c->stream));
            for (int n = s0; n < s1; n++) {
                HIPCHK(c, hipMemcpyAsync(seg[n - s0], c->W[c->cur], sb, hipMemcpyDeviceToDevice, c->stream));
                if (n + 1 < s1) {
                    rc = eulerian_step(c);
                    if (rc) return rc;
                }
            }
        }
        for (int n = s1 - 1; n >= s0; n--) {
            HIPCHK(c, gkm::launch_adj_step(seg[n - s0], c->adj_vec[l], c->plane, c->nc, c->gas, dt, c->dx, s,
                                           c->adj_vec[1 - l], c->stream));
            l = 1 - l;
        }
    }
    HIPCHK(c, hipMemcpyAsync(c->W[c->cur], final_state, sb, hipMemcpyDeviceToDevice, c->stream));
    c->lpt.valid = false;
    c->adj_n = -1;  // the segment buffers now hold the first segment: the trajectory is used up
    return vec_download(c, c->adj_vec[l], lam0);
}

int gkm_adjoint_run(gkm_ctx* c, double dt, int nsteps, int every, const double* lamN, double* lam0)
{
    if (!c || !lamN || !lam0 || nsteps < 0 || every < 1) return fail(c, GKM_EINVAL, "bad argument");
    const int rc = gkm_adjoint_forward(c, dt, nsteps, every);
    if (rc) return rc;
    return gkm_adjoint_backward(c, lamN, lam0);
}

int gkm_set_sweep_tiles(gkm_ctx* c, const int tile[3])
{
    if (!c || !tile) return fail(c, GKM_EINVAL, "null argument");
    for (int d = 0; d < 3; d++)
        if (!gkm::sweep_tile_ok(d, tile[d], c->nc))
            return fail(c, GKM_EINVAL, "tile variant out of range or larger than the grid");
    for (int d = 0; d < 3; d++) c->tile[d] = tile[d];
    return GKM_OK;
}

int gkm_get_sweep_tiles(const gkm_ctx* c, int tile[3])
{
    if (!c || !tile) return GKM_EINVAL;
    for (int d = 0; d < 3; d++) tile[d] = c->tile[d];
    return GKM_OK;
}

int gkm_autotune(gkm_ctx* c, int reps, int tile[3], double ms[9])
{
    if (!c || reps < 1) return fail(c, GKM_EINVAL, "bad argument");
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    // the CFL dt of the resident state, so each candidate does the step's real work
    int rc = cfl_on_device(c);
    if (rc) return rc;
    const bool mixed = c->cfg.precision == GKM_MIXED;
    for (int dir = 0; dir < 3; dir++) {
        double best = 0.0;
        int best_v = c->tile[dir];
        for (int v = 0; v < gkm::kTileVariants; v++) {
            if (!gkm::sweep_tile_ok(dir, v, c->nc)) {
                if (ms) ms[3 * dir + v] = -1.0;
                continue;
            }
            gkm::SweepArgs a = sweep_args(c, dir, dir);
            a.tile = v;
            a.seg = gkm::sweep_seg_len(dir, c->nc, v, c->slots);
            a.nseg = gkm::sweep_nseg(c->nc, a.seg);
            // dW is scratch and Wout is the next state's buffer: the resident state is not changed
            hipError_t e = gkm::launch_sweep(dir, a, mixed, c->stream);  // warm-up
            if (e != hipSuccess) return fail(c, GKM_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
            HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
            for (int r = 0; r < reps; r++) {
                e = gkm::launch_sweep(dir, a, mixed, c->stream);
                if (e != hipSuccess) return fail(c, GKM_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
            }
            HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
            HIPCHK(c, hipEventSynchronize(c->ev[1]));
            float el = 0.f;
            HIPCHK(c, hipEventElapsedTime(&el, c->ev[0], c->ev[1]));
            const double avg = el / reps;
            if (ms) ms[3 * dir + v] = avg;
            if (best == 0.0 || avg < best) {
                best = avg;
                best_v = v;
            }
        }
        c->tile[dir] = best_v;
    }
    if (tile)
        for (int d = 0; d < 3; d++) tile[d] = c->tile[d];
    return GKM_OK;
}

int gkm_sync(gkm_ctx* c)
{
    if (!c) return GKM_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GKM_OK;
}

int gkm_timing_enable(gkm_ctx* c, int on)
{
    if (!c) return GKM_EINVAL;
    c->timing = on != 0;
    for (int i = 0; i < 8; i++) {
        c->tms[i] = 0.0;
        c->tcount[i] = 0;
    }
    return GKM_OK;
}

int gkm_timing_read(gkm_ctx* c, double ms[8], long counts[8])
{
    if (!c || !ms || !counts) return fail(c, GKM_EINVAL, "null argument");
    for (int i = 0; i < 8; i++) {
        ms[i] = c->tms[i];
        counts[i] = c->tcount[i];
    }
    return GKM_OK;
}

int gkm_flux_batch(gkm_ctx* c, long n, const double* in, const double* tau, double dt, double dx, int precision,
                   double* F)
{
    if (!c || n < 0 || (n > 0 && (!in || !tau || !F)) || !(dx > 0.0) || precision < 0 || precision > 2)
        return fail(c, GKM_EINVAL, "bad argument");
    if (n == 0) return GKM_OK;
    double *din = nullptr, *dtau = nullptr, *dF = nullptr;
    int rc = GKM_OK;
    if (hipMalloc(&din, sizeof(double) * 50 * n) != hipSuccess || hipMalloc(&dtau, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&dF, sizeof(double) * 5 * n) != hipSuccess) {
        rc = fail(c, GKM_ENOMEM, "hipMalloc flux batch");
    } else {
        hipError_t e = hipMemcpyAsync(din, in, sizeof(double) * 50 * n, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dtau, tau, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = gkm::launch_flux_batch(n, din, dtau, dt, dx, c->gas, precision, dF, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(F, dF, sizeof(double) * 5 * n, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(c, GKM_EHIP, std::string("flux batch: ") + hipGetErrorString(e));
    }
    (void)hipFree(din);
    (void)hipFree(dtau);
    (void)hipFree(dF);
    return rc;
}

int gkm_weno_batch(gkm_ctx* c, long n, const double* in, int precision, double* out)
{
    if (!c || n < 0 || (n > 0 && (!in || !out)) || precision < 0 || precision > 2)
        return fail(c, GKM_EINVAL, "bad argument");
    if (n == 0) return GKM_OK;
    double *din = nullptr, *dout = nullptr;
    int rc = GKM_OK;
    if (hipMalloc(&din, sizeof(double) * 30 * n) != hipSuccess || hipMalloc(&dout, sizeof(double) * 10 * n) != hipSuccess) {
        rc = fail(c, GKM_ENOMEM, "hipMalloc weno batch");
    } else {
        hipError_t e = hipMemcpyAsync(din, in, sizeof(double) * 30 * n, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = gkm::launch_weno_batch(n, din, c->gas, precision, dout, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(double) * 10 * n, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(c, GKM_EHIP, std::string("weno batch: ") + hipGetErrorString(e));
    }
    (void)hipFree(din);
    (void)hipFree(dout);
    return rc;
}

}  // extern "C"

// ===========================================================================
// Splines, statistics and particles
// ===========================================================================
namespace {

lpt::Geom geom_of(const gkm_ctx* c)
{
    lpt::Geom g;
    g.nc = c->nc;
    g.nt = c->nc + 6;
    g.N = c->nc + 8;
    // Interpolation.cpp:7-9 with sta = 0: start = -3 dx + dx/2, end = start + (nt-1) dx
    const double s = (0 - 3) * c->dx + 0.5 * c->dx;
    const double e = s + (g.nt - 1) * c->dx;
    g.start = s;
    g.dinv = 1.0 / ((e - s) / (double)(g.nt - 1));
    g.drop = c->lpt.drop;
    return g;
}

int ensure_spline_mem(gkm_ctx* c)
{
    lpt::State& L = c->lpt;
    if (L.pivots) return GKM_OK;
    const long N = c->nc + 8;
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < 3; q++)
            if (hipMalloc(&L.coef[b][q], sizeof(double) * N * N * N) != hipSuccess)
                return fail(c, GKM_ENOMEM, "hipMalloc spline coefficients");
    const int n = c->nc + 6 - 2;
    double* piv = new double[n];
    for (int r = 0; r < n; r++) piv[r] = 1.0 / (4.0 - (r > 0 ? piv[r - 1] : 0.0));
    if (hipMalloc(&L.pivots, sizeof(double) * n) != hipSuccess) {
        delete[] piv;
        return fail(c, GKM_ENOMEM, "hipMalloc pivots");
    }
    hipError_t e = hipMemcpy(L.pivots, piv, sizeof(double) * n, hipMemcpyHostToDevice);
    delete[] piv;
    if (e != hipSuccess) return fail(c, GKM_EHIP, std::string("hipMemcpy pivots: ") + hipGetErrorString(e));
    if (hipMalloc(&L.red, sizeof(double) * 32 * lpt::kStatBlocks) != hipSuccess)
        return fail(c, GKM_ENOMEM, "hipMalloc reduction scratch");
    return GKM_OK;
}

// velocity spline of the resident state into coefficient buffer b
int build_spline(gkm_ctx* c, int b)
{
    int rc = ensure_spline_mem(c);
    if (rc) return rc;
    const lpt::Geom g = geom_of(c);
    if (!lpt::ens_z_chunks(g.nt)) return fail(c, GKM_EINVAL, "splines need nc <= 1082");
    return timed(c, 4, [&] {
        return lpt::launch_prefilter_velocity(c->W[c->cur], c->plane, c->gas.gam, g, c->lpt.pivots, c->lpt.coef[b],
                                              c->stream);
    });
}

int ensure_spline(gkm_ctx* c)
{
    if (c->lpt.valid) return GKM_OK;
    int rc = build_spline(c, c->lpt.cur);
    if (rc) return rc;
    c->lpt.valid = true;
    return GKM_OK;
}

int sum_partials(gkm_ctx* c, int stride, int off, int K, double* out)
{
    std::vector<double> h((size_t)stride * lpt::kStatBlocks);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->lpt.red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int b = 0; b < lpt::kStatBlocks; b++) s += h[(size_t)b * stride + off + k];
        out[k] = s;
    }
    return GKM_OK;
}

}  // namespace

extern "C" {

int gkm_stats_compute(gkm_ctx* c, gkm_stats* out)
{
    if (!c || !out) return fail(c, GKM_EINVAL, "null argument");
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    int rc = ensure_spline(c);
    if (rc) return rc;
    const lpt::Geom g = geom_of(c);
    const lpt::StatParams sp{c->gas.gam, c->gas.R, c->gas.mu0, c->gas.T0};
    const double* cu[3] = {c->lpt.coef[c->lpt.cur][0], c->lpt.coef[c->lpt.cur][1], c->lpt.coef[c->lpt.cur][2]};
    hipError_t e = lpt::launch_stats(c->W[c->cur], c->plane, cu, g, sp, c->lpt.red, c->stream);
    if (e != hipSuccess) return fail(c, GKM_EHIP, std::string("stats: ") + hipGetErrorString(e));
    double s[16];
    rc = sum_partials(c, 32, 0, 16, s);
    if (rc) return rc;
    const double Nc = (double)c->plane;
    out->tke = s[0] / Nc;
    out->diss = s[1] / Nc;
    out->sijsij = s[2] / Nc;
    out->enstrophy = s[3] / Nc;
    for (int a = 0; a < 3; a++) {
        const double m2 = s[4 + a] / Nc, m3 = s[7 + a] / Nc, m4 = s[10 + a] / Nc;
        out->skew[a] = m3 / std::pow(m2, 1.5);
        out->kurt[a] = m4 / (m2 * m2);
    }
    const double pm = s[13] / Nc, tm = s[14] / Nc, vm = s[15] / Nc;
    e = lpt::launch_stats2(c->W[c->cur], c->plane, sp, pm, tm, vm, c->lpt.red, c->stream);
    if (e != hipSuccess) return fail(c, GKM_EHIP, std::string("stats2: ") + hipGetErrorString(e));
    double r[3];
    rc = sum_partials(c, 32, 16, 3, r);
    if (rc) return rc;
    const double Mt = c->cfg.Mt, gam = c->gas.gam;
    out->pmean = pm;
    out->tmean = tm;
    out->vmean = vm;
    out->prms = std::sqrt(r[0] / Nc) / (gam * c->cfg.p0 * Mt * Mt);         // update.cpp:825-833
    out->trms = std::sqrt(r[1] / Nc) / ((gam - 1) * c->gas.T0 * Mt * Mt);
    out->vrms = std::sqrt(r[2] / Nc) / (Mt * Mt / c->cfg.rho0);
    return GKM_OK;
}

int lpt_upload(gkm_ctx* c, long np, const double* pos)
{
    if (!c || np < 0 || np > 0x7fffffffL || (np > 0 && !pos)) return fail(c, GKM_EINVAL, "bad argument");
    lpt::State& L = c->lpt;
    const double twopi = 2.0 * (std::atan(1.0) * 4.0);
    for (long p = 0; p < 3 * np; p++)
        if (!(pos[p] >= 0.0 && pos[p] <= twopi)) return fail(c, GKM_EINVAL, "particle position outside [0, 2 pi]");
    for (void* ptr : {(void*)L.pos, (void*)L.U, (void*)L.G, (void*)L.tag, (void*)L.tag2, (void*)L.pos2, (void*)L.keys,
                      (void*)L.perm, L.sort_tmp, (void*)L.props})
        (void)hipFree(ptr);
    L.pos = L.U = L.G = L.pos2 = L.props = nullptr;
    L.tag = L.tag2 = L.perm = nullptr;
    L.keys = nullptr;
    L.sort_tmp = nullptr;
    L.sort_tmp_bytes = 0;
    L.np = np;
    L.sampled = false;
    if (np == 0) return GKM_OK;
    if (hipMalloc(&L.pos, sizeof(double) * 3 * np) != hipSuccess || hipMalloc(&L.U, sizeof(double) * 3 * np) != hipSuccess ||
        hipMalloc(&L.G, sizeof(double) * 9 * np) != hipSuccess || hipMalloc(&L.pos2, sizeof(double) * 9 * np) != hipSuccess ||
        hipMalloc(&L.tag, sizeof(int) * np) != hipSuccess || hipMalloc(&L.tag2, sizeof(int) * np) != hipSuccess ||
        hipMalloc(&L.keys, sizeof(unsigned) * 2 * np) != hipSuccess || hipMalloc(&L.perm, sizeof(int) * 2 * np) != hipSuccess)
        return fail(c, GKM_ENOMEM, "hipMalloc particles");
    std::vector<int> tags(np);
    for (long p = 0; p < np; p++) tags[p] = (int)p;
    std::vector<double> kept;
    if (L.drop) {  // locateParticles keeps a tracer only in the block that owns it (update.cpp:1304-1316)
        kept.assign(pos, pos + 3 * np);
        for (long p = 0; p < np; p++) {
            bool out = false;
            for (int a = 0; a < 3; a++) out |= (int)(kept[3 * p + a] / twopi) != 0;
            if (out)
                for (int a = 0; a < 3; a++) kept[3 * p + a] = std::nan("");
        }
        pos = kept.data();
    }
    HIPCHK(c, hipMemcpyAsync(L.pos, pos, sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(L.tag, tags.data(), sizeof(int) * np, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, lpt::sort_by_cell(L, geom_of(c), c->dx, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GKM_OK;
}

// copy a per-particle array (slot order, `width` doubles per particle) to the host in tag order
static int download_tag_order(gkm_ctx* c, const double* src, int width, double* host)
{
    lpt::State& L = c->lpt;
    HIPCHK(c, lpt::scatter_to_tags(L.np, L.tag, src, width, L.pos2, c->stream));
    HIPCHK(c, hipMemcpyAsync(host, L.pos2, sizeof(double) * width * L.np, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GKM_OK;
}

int lpt_set_reference_drop(gkm_ctx* c, int on)
{
    if (!c) return GKM_EINVAL;
    if (c->lpt.np > 0) return fail(c, GKM_ENOTREADY, "set the tracer ownership mode before lpt_upload");
    c->lpt.drop = on != 0;
    return GKM_OK;
}

int lpt_count(gkm_ctx* c, long* alive)
{
    if (!c || !alive) return fail(c, GKM_EINVAL, "null argument");
    const lpt::State& L = c->lpt;
    *alive = L.np;
    if (L.np == 0 || !L.drop) return GKM_OK;
    std::vector<double> pos((size_t)3 * L.np);
    HIPCHK(c, hipMemcpyAsync(pos.data(), L.pos, sizeof(double) * 3 * L.np, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    long n = 0;
    for (long p = 0; p < L.np; p++) n += !std::isnan(pos[(size_t)3 * p]);
    *alive = n;
    return GKM_OK;
}

int lpt_download(gkm_ctx* c, double* pos, double* vel)
{
    if (!c) return GKM_EINVAL;
    const lpt::State& L = c->lpt;
    if (L.np == 0) return GKM_OK;
    int rc = GKM_OK;
    if (pos) rc = download_tag_order(c, L.pos, 3, pos);
    if (!rc && vel) rc = download_tag_order(c, L.U, 3, vel);
    return rc;
}

static int sample_device(gkm_ctx* c)
{
    int rc = ensure_spline(c);
    if (rc) return rc;
    lpt::State& L = c->lpt;
    const lpt::Geom g = geom_of(c);
    const double* cu[3] = {L.coef[L.cur][0], L.coef[L.cur][1], L.coef[L.cur][2]};
    rc = timed(c, 5, [&] { return lpt::launch_sample(L.np, L.pos, cu, g, L.U, L.G, c->stream); });
    if (rc) return rc;
    L.sampled = true;
    return GKM_OK;
}

int lpt_sample(gkm_ctx* c, double* U, double* gU, double* gV, double* gW)
{
    if (!c) return GKM_EINVAL;
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    lpt::State& L = c->lpt;
    if (L.np == 0) return GKM_OK;
    int rc = sample_device(c);
    if (rc) return rc;
    if (U) {
        rc = download_tag_order(c, L.U, 3, U);
        if (rc) return rc;
    }
    if (gU || gV || gW) {
        std::vector<double> G((size_t)9 * L.np);
        rc = download_tag_order(c, L.G, 9, G.data());
        if (rc) return rc;
        double* outs[3] = {gU, gV, gW};
        for (int f = 0; f < 3; f++)
            if (outs[f])
                for (long p = 0; p < L.np; p++)
                    for (int a = 0; a < 3; a++) outs[f][3 * p + a] = G[(size_t)9 * p + 3 * f + a];
    }
    return GKM_OK;
}

// splines of rho, p and the velocity Laplacians of the resident state, and the
// device table of the 8 splines getParticleProp reads (update.cpp:1200-1233)
static int build_prop_splines(gkm_ctx* c)
{
    int rc = ensure_spline(c);
    if (rc) return rc;
    lpt::State& L = c->lpt;
    const long N = c->nc + 8;
    for (int f = 0; f < 5; f++)
        if (!L.pcoef[f] && hipMalloc(&L.pcoef[f], sizeof(double) * N * N * N) != hipSuccess)
            return fail(c, GKM_ENOMEM, "hipMalloc property splines");
    if (!L.pcs && hipMalloc(&L.pcs, sizeof(double*) * 8) != hipSuccess)
        return fail(c, GKM_ENOMEM, "hipMalloc spline table");
    const lpt::Geom g = geom_of(c);
    HIPCHK(c, lpt::launch_prefilter(c->W[c->cur], c->plane, 0, c->gas.gam, g, L.pivots, L.pcoef[0], c->stream));
    HIPCHK(c, lpt::launch_prefilter(c->W[c->cur], c->plane, 4, c->gas.gam, g, L.pivots, L.pcoef[1], c->stream));
    for (int f = 0; f < 3; f++)
        HIPCHK(c, lpt::launch_lap_spline(L.coef[L.cur][f], g, L.pivots, L.pcoef[2 + f], c->stream));
    const double* table[8] = {L.pcoef[0], L.coef[L.cur][0], L.coef[L.cur][1], L.coef[L.cur][2],
                              L.pcoef[1], L.pcoef[2], L.pcoef[3], L.pcoef[4]};
    HIPCHK(c, hipMemcpyAsync((void*)L.pcs, table, sizeof(table), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // `table` is a host stack array
    return GKM_OK;
}

int lpt_properties(gkm_ctx* c, double* out)
{
    if (!c || !out) return fail(c, GKM_EINVAL, "null argument");
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    lpt::State& L = c->lpt;
    if (L.np == 0) return GKM_OK;
    int rc = build_prop_splines(c);
    if (rc) return rc;
    if (!L.props && hipMalloc(&L.props, sizeof(double) * lpt::kPropWidth * L.np) != hipSuccess)
        return fail(c, GKM_ENOMEM, "hipMalloc particle properties");
    const lpt::Geom g = geom_of(c);
    HIPCHK(c, lpt::launch_props(L.np, L.pos, L.tag, L.pcs, g, L.props, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, L.props, sizeof(double) * lpt::kPropWidth * L.np, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GKM_OK;
}

int lpt_stats(gkm_ctx* c, gkm_pstats* out)
{
    if (!c || !out) return fail(c, GKM_EINVAL, "null argument");
    lpt::State& L = c->lpt;
    if (!L.sampled) return fail(c, GKM_ENOTREADY, "lpt_stats needs a sample");
    int rc = ensure_spline_mem(c);
    if (rc) return rc;
    hipError_t e = lpt::launch_pstats(L.np, L.U, L.G, L.red, c->stream);
    if (e != hipSuccess) return fail(c, GKM_EHIP, std::string("pstats: ") + hipGetErrorString(e));
    double s[6];
    rc = sum_partials(c, 8, 0, 6, s);
    if (rc) return rc;
    const double n = (double)L.np;  // update.cpp:1147-1152 divide by numparticles
    out->tke = s[0] / n;
    out->sijsij = s[1] / n;
    out->skew_u = (s[3] / n) / std::pow(s[2] / n, 1.5);
    out->skew_w = (s[5] / n) / std::pow(s[4] / n, 1.5);
    return GKM_OK;
}

int lpt_heun(gkm_ctx* c, double dt)
{
    if (!c) return GKM_EINVAL;
    lpt::State& L = c->lpt;
    if (!L.sampled) return fail(c, GKM_ENOTREADY, "lpt_heun needs the start-of-step sample");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, GKM_EINVAL, "dt must be positive and finite");
    int rc = ensure_spline(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->dscal + 2, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const lpt::Geom g = geom_of(c);
    const double* cn[3] = {L.coef[L.cur][0], L.coef[L.cur][1], L.coef[L.cur][2]};
    rc = timed(c, 5, [&] { return lpt::launch_heun(L.np, L.pos, L.U, cn, g, c->dscal + 2, c->stream); });
    if (rc) return rc;
    if (++L.since_sort >= L.sort_every) {
        rc = timed(c, 6, [&] { return lpt::sort_by_cell(L, g, c->dx, c->stream); });
        if (rc) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GKM_OK;
}

int lpt_coupled_step(gkm_ctx* c, double dt)
{
    if (!c) return GKM_EINVAL;
    if (!c->have_state) return fail(c, GKM_ENOTREADY, "no state uploaded");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, GKM_EINVAL, "dt must be positive and finite");
    lpt::State& L = c->lpt;
    int rc = sample_device(c);  // U^n at x^n from the spline of W^n (update.cpp:136-137)
    if (rc) return rc;
    rc = gkm_step(c, dt);       // W^{n+1} (update.cpp:145-156)
    if (rc) return rc;
    const int nb = 1 - L.cur;   // spline of W^{n+1} (update.cpp:157)
    rc = build_spline(c, nb);
    if (rc) return rc;
    L.cur = nb;
    L.valid = true;
    const lpt::Geom g = geom_of(c);
    const double* cn[3] = {L.coef[nb][0], L.coef[nb][1], L.coef[nb][2]};
    rc = timed(c, 5, [&] { return lpt::launch_heun(L.np, L.pos, L.U, cn, g, c->dscal, c->stream); });
    if (rc) return rc;
    if (++L.since_sort >= L.sort_every) {
        rc = timed(c, 6, [&] { return lpt::sort_by_cell(L, g, c->dx, c->stream); });
        if (rc) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GKM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Ensemble: independent members (own grid size, gas, IC, dt) stepped together.
// Per step one launch computes every member's CFL dt (two kernels), and one
// launch per direction sweeps every member (ens_sweep_kernel), so a step is
// five launches whatever the number of members; a step from each buffer
// parity is captured once into a HIP graph and replayed.  Each member is a
// gkm_ctx, so its state goes in and out through gkm_upload_state /
// gkm_download_state / gkm_stats_compute.  No reference counterpart.
// ---------------------------------------------------------------------------
struct gkm_ensemble {
    int n = 0;
    bool mixed = false;
    std::vector<gkm_ctx*> mem;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int tile[3] = {0, 0, 0};
    gkm::SweepArgs* dargs = nullptr;  // [parity][direction][member]
    int* dfirst = nullptr;            // [direction][member + 1] first block of each member
    int nblk[3] = {0, 0, 0};
    bool exact[3] = {false, false, false};  // per direction: tile and segment divide every member's grid
    bool compensated = false;               // every member's z sweep takes the compensated update
    gkm::CflArgs* dcfl = nullptr;     // [parity][member]
    double* partial = nullptr;        // n * kEnsCflBlocks
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // tracers of the members that hold particles ("entries"); the spline of
    // W[q] lives in coef[q], so both captured steps are fixed
    int npart = 0;
    lpt::EnsPart* dpart = nullptr;  // [parity][sample, Heun][entry]
    int* dpfirst = nullptr;         // [entry + 1] first block of each entry
    int pblk = 0;
    lpt::ZArgs* dz = nullptr;       // [parity][entry]: z pass of u, v, w of the next state
    int* dzfirst = nullptr;         // [entry + 1]
    int zblk = 0, zchunks = 0;
    lpt::EnsLine* dline = nullptr;  // [parity][3 entry]: x and y passes of u, v, w
    int* dlfirst = nullptr;         // [axis - 1][3 entry + 1]
    int lblk[2] = {0, 0};
    size_t pcap = 0;                // entries the device tables hold
    std::vector<char> psig;         // bytes of the uploaded tables
    // persistent mode: the whole step is one launch of ens_step_kernel
    bool persistent = false;
    int pgrid = 0;                     // resident workgroups
    unsigned long long* dbar = nullptr;  // barrier counter, base, timeout flag, work counters
    unsigned long long timeout_ticks = 0;
    std::string err;
};

namespace {

int efail(gkm_ensemble* e, int code, const std::string& msg)
{
    if (e) e->err = msg;
    return code;
}

#define EHIPCHK(ens, call)                                                                           \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return efail((ens), GKM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));         \
    } while (0)

void drop_graphs(gkm_ensemble* e)
{
    for (auto& g : e->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// sweep arguments of member m for parity p (state in W[p], next state W[1-p]), direction d, tile v
gkm::SweepArgs ens_sweep_args(const gkm_ctx* c, int p, int d, int v)
{
    gkm::SweepArgs a{};
    a.W = c->W[p];
    a.dW = c->dW;
    a.Wout = c->W[1 - p];
    a.cmp_in = c->compensate ? c->cmp[p] : nullptr;
    a.cmp_out = c->compensate ? c->cmp[1 - p] : nullptr;
    a.dt = c->dscal;
    a.plane = c->plane;
    a.nc = c->nc;
    a.dx = c->dx;
    a.inv_dx = 1.0 / c->dx;
    a.g = c->gas;
    a.tile = v;
    a.seg = gkm::sweep_seg_len(d, c->nc, v, 0);
    a.nseg = gkm::sweep_nseg(c->nc, a.seg);
    a.mode = d;
    return a;
}

// upload the argument tables for tiles t[3]
int ens_build(gkm_ensemble* e, const int t[3])
{
    const int n = e->n;
    std::vector<gkm::SweepArgs> args((size_t)6 * n);
    std::vector<int> first((size_t)3 * (n + 1));
    // the sweep kernels take the update's kind as a compile-time choice: the
    // members must agree (mixed members compensate unless switched off alone,
    // which members cannot be)
    int ncomp = 0;
    for (gkm_ctx* c : e->mem) ncomp += c->compensate ? 1 : 0;
    if (ncomp != 0 && ncomp != n) return efail(e, GKM_EINVAL, "ensemble members differ in compensation");
    for (int d = 0; d < 3; d++) {
        int acc = 0;
        bool exact = true;
        // table entries longest workgroups first (faces per workgroup along the
        // sweep, then member order), so that the short ones fill the last round
        // instead of the long ones trailing it: the persistent step hands the
        // pencils out in this order (64-member workload: 6.62 -> 6.42 ms per
        // step); the per-phase launches measured the same either way
        std::vector<int> order(n);
        std::vector<int> seg(n);
        for (int m = 0; m < n; m++) {
            order[m] = m;
            seg[m] = ens_sweep_args(e->mem[m], 0, d, t[d]).seg;
        }
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return seg[x] > seg[y]; });
        for (int k = 0; k < n; k++) {
            const int m = order[k];
            first[(size_t)d * (n + 1) + k] = acc;
            for (int p = 0; p < 2; p++) args[((size_t)p * 3 + d) * n + k] = ens_sweep_args(e->mem[m], p, d, t[d]);
            const gkm::SweepArgs& a = args[(size_t)d * n + k];
            acc += gkm::sweep_blocks(d, a.nc, t[d], a.seg);
            exact = exact && gkm::sweep_exact(d, t[d], a.nc, a.seg);
        }
        first[(size_t)d * (n + 1) + n] = acc;
        e->nblk[d] = acc;
        e->exact[d] = exact;
    }
    e->compensated = ncomp == n;
    std::vector<gkm::CflArgs> cfl((size_t)2 * n);
    for (int p = 0; p < 2; p++)
        for (int m = 0; m < n; m++) {
            const gkm_ctx* c = e->mem[m];
            cfl[(size_t)p * n + m] = gkm::CflArgs{c->W[p], c->plane, c->gas.gam, c->cfg.cfl, c->dx, c->dscal,
                                                  c->dscal + 1, c->dflag};
        }
    EHIPCHK(e, hipMemcpyAsync(e->dargs, args.data(), sizeof(gkm::SweepArgs) * args.size(), hipMemcpyHostToDevice,
                              e->stream));
    EHIPCHK(e, hipMemcpyAsync(e->dfirst, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice, e->stream));
    EHIPCHK(e, hipMemcpyAsync(e->dcfl, cfl.data(), sizeof(gkm::CflArgs) * cfl.size(), hipMemcpyHostToDevice,
                              e->stream));
    EHIPCHK(e, hipStreamSynchronize(e->stream));  // the host vectors go out of scope
    for (int d = 0; d < 3; d++) e->tile[d] = t[d];
    drop_graphs(e);
    return GKM_OK;
}

// one step of every member from parity p: tracer velocities at x^n from the
// spline of W^n (update.cpp:136-137), CFL dt, the x, y, z sweeps, the velocity
// splines of W^{n+1} (update.cpp:157) and the Heun update (ParticleTrack.cpp:48-73)
gkm::PersistStep persist_args(const gkm_ensemble* e, int p)
{
    const int n = e->n, np = e->npart;
    gkm::PersistStep a{};
    for (int d = 0; d < 3; d++) {
        a.sweep[d] = e->dargs + ((size_t)p * 3 + d) * n;
        a.sfirst[d] = e->dfirst + (size_t)d * (n + 1);
        a.nsweep[d] = e->nblk[d];
        a.tile[d] = e->tile[d];
    }
    a.cfl = e->dcfl + (size_t)p * n;
    a.partial = e->partial;
    a.nmem = n;
    if (np) {
        a.sample = e->dpart + (size_t)(2 * p) * np;
        a.heun = e->dpart + (size_t)(2 * p + 1) * np;
        a.pfirst = e->dpfirst;
        a.npart = np;
        a.pblk = e->pblk;
        a.z = e->dz + (size_t)p * np;
        a.zfirst = e->dzfirst;
        a.zblk = e->zblk;
        a.zchunks = e->zchunks;
        a.lines = e->dline + (size_t)p * 3 * np;
        a.lfirst = e->dlfirst;
        a.lblk[0] = e->lblk[0];
        a.lblk[1] = e->lblk[1];
    }
    a.count = e->dbar;
    a.base = e->dbar + 1;
    a.timeout = reinterpret_cast<int*>(e->dbar + 2);
    a.timeout_ticks = e->timeout_ticks;
    a.work = reinterpret_cast<unsigned int*>(e->dbar + 3);
    return a;
}

// capturing: the launches go into a HIP graph (the persistent step is then an
// ordinary launch; eager persistent steps are cooperative launches)
hipError_t ens_step(gkm_ensemble* e, int p, bool capturing)
{
    const int n = e->n, np = e->npart;
    if (e->persistent)
        return gkm::launch_persist_step(persist_args(e, p), e->mixed, e->pgrid, !capturing, e->stream);
    hipError_t r = hipSuccess;
    if (np) r = lpt::launch_ens_sample(e->dpart + (size_t)(2 * p) * np, e->dpfirst, np, e->pblk, e->stream);
    if (r == hipSuccess) r = gkm::launch_ens_cfl(e->dcfl + (size_t)p * n, n, e->partial, e->stream);
    for (int d = 0; d < 3 && r == hipSuccess; d++)
        r = gkm::launch_ens_sweep(d, e->tile[d], e->dargs + ((size_t)p * 3 + d) * n, e->dfirst + (size_t)d * (n + 1),
                                  n, e->nblk[d], e->mixed, e->exact[d], e->compensated, e->stream);
    if (np && r == hipSuccess)
        r = lpt::launch_ens_prefilter(e->dz + (size_t)p * np, e->dzfirst, np, e->zblk, e->zchunks,
                                      e->dline + (size_t)p * 3 * np, e->dlfirst, 3 * np, e->lblk, e->stream);
    if (np && r == hipSuccess)
        r = lpt::launch_ens_heun(e->dpart + (size_t)(2 * p + 1) * np, e->dpfirst, np, e->pblk, e->stream);
    return r;
}

// Tracer tables for a run starting at parity p.  Every member with particles
// gets the spline of its resident state W[p] in coef[p] (built here, outside
// any capture), then the tables are uploaded if they changed (new particle
// buffers, members gaining or losing particles), which drops the graphs.
int ens_particles_prepare(gkm_ensemble* e, int p)
{
    std::vector<gkm_ctx*> ent;
    for (gkm_ctx* c : e->mem)
        if (c->lpt.np > 0) ent.push_back(c);
    for (gkm_ctx* c : ent) {
        lpt::State& L = c->lpt;
        if (L.valid && L.cur == p) continue;
        const int rc = build_spline(c, p);
        if (rc) return efail(e, rc, std::string("member spline: ") + c->err);
        L.cur = p;
        L.valid = true;
    }
    const int np = (int)ent.size();
    std::vector<lpt::EnsPart> parts((size_t)4 * np);
    std::vector<lpt::ZArgs> zs((size_t)2 * np);
    std::vector<lpt::EnsLine> lines((size_t)6 * np);
    std::vector<int> pfirst(np + 1), zfirst(np + 1), lfirst((size_t)2 * (3 * np + 1));
    int acc = 0, zacc = 0, lacc[2] = {0, 0}, ntmax = 0;
    for (int k = 0; k < np; k++) {
        gkm_ctx* c = ent[k];
        lpt::State& L = c->lpt;
        const lpt::Geom g = geom_of(c);
        ntmax = std::max(ntmax, g.nt);
        for (int q = 0; q < 2; q++)
            for (int h = 0; h < 2; h++) {
                const int b = h ? 1 - q : q;  // sample on the spline of W[q], Heun on that of W[1-q]
                parts[(size_t)(2 * q + h) * np + k] =
                    lpt::EnsPart{L.np, L.pos, L.U, L.G, {L.coef[b][0], L.coef[b][1], L.coef[b][2]}, g, c->dscal};
            }
        for (int q = 0; q < 2; q++) {
            zs[(size_t)q * np + k] = lpt::ZArgs{c->W[1 - q], c->plane, 1, 3, c->gas.gam, g, L.pivots,
                                                {L.coef[1 - q][0], L.coef[1 - q][1], L.coef[1 - q][2]}};
            for (int f = 0; f < 3; f++)
                lines[(size_t)q * 3 * np + 3 * k + f] = lpt::EnsLine{g, L.pivots, L.coef[1 - q][f]};
        }
        pfirst[k] = acc;
        acc += lpt::ens_particle_blocks(L.np);
        zfirst[k] = zacc;
        zacc += lpt::ens_z_blocks(g);
        for (int f = 0; f < 3; f++)
            for (int ax = 0; ax < 2; ax++) {
                lfirst[(size_t)ax * (3 * np + 1) + 3 * k + f] = lacc[ax];
                lacc[ax] += lpt::ens_xy_blocks(g, ax + 1);
            }
    }
    pfirst[np] = acc;
    zfirst[np] = zacc;
    for (int ax = 0; ax < 2; ax++) lfirst[(size_t)ax * (3 * np + 1) + 3 * np] = lacc[ax];
    const int zchunks = np ? lpt::ens_z_chunks(ntmax) : 0;
    if (np && !zchunks) return efail(e, GKM_EINVAL, "splines need nc <= 1082");
    std::vector<char> sig;
    auto put = [&](const void* ptr, size_t bytes) {
        const char* b = (const char*)ptr;
        sig.insert(sig.end(), b, b + bytes);
    };
    put(&np, sizeof(np));
    put(parts.data(), sizeof(lpt::EnsPart) * parts.size());
    put(zs.data(), sizeof(lpt::ZArgs) * zs.size());
    put(lines.data(), sizeof(lpt::EnsLine) * lines.size());
    put(pfirst.data(), sizeof(int) * pfirst.size());
    put(zfirst.data(), sizeof(int) * zfirst.size());
    put(lfirst.data(), sizeof(int) * lfirst.size());
    if (sig == e->psig) return GKM_OK;
    drop_graphs(e);
    e->psig.clear();
    e->npart = 0;
    if ((size_t)np > e->pcap) {
        for (void* ptr : {(void*)e->dpart, (void*)e->dpfirst, (void*)e->dz, (void*)e->dzfirst, (void*)e->dline,
                          (void*)e->dlfirst})
            (void)hipFree(ptr);
        e->dpart = nullptr;
        e->dz = nullptr;
        e->dline = nullptr;
        e->dpfirst = e->dzfirst = e->dlfirst = nullptr;
        e->pcap = 0;
        if (hipMalloc(&e->dpart, sizeof(lpt::EnsPart) * 4 * (size_t)np) != hipSuccess ||
            hipMalloc(&e->dpfirst, sizeof(int) * (size_t)(np + 1)) != hipSuccess ||
            hipMalloc(&e->dz, sizeof(lpt::ZArgs) * 2 * (size_t)np) != hipSuccess ||
            hipMalloc(&e->dzfirst, sizeof(int) * (size_t)(np + 1)) != hipSuccess ||
            hipMalloc(&e->dline, sizeof(lpt::EnsLine) * 6 * (size_t)np) != hipSuccess ||
            hipMalloc(&e->dlfirst, sizeof(int) * 2 * (size_t)(3 * np + 1)) != hipSuccess)
            return efail(e, GKM_ENOMEM, "hipMalloc ensemble tracer tables");
        e->pcap = (size_t)np;
    }
    if (np) {
        auto up = [&](void* dst, const void* src, size_t bytes) {
            return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream);
        };
        EHIPCHK(e, up(e->dpart, parts.data(), sizeof(lpt::EnsPart) * parts.size()));
        EHIPCHK(e, up(e->dpfirst, pfirst.data(), sizeof(int) * pfirst.size()));
        EHIPCHK(e, up(e->dz, zs.data(), sizeof(lpt::ZArgs) * zs.size()));
        EHIPCHK(e, up(e->dzfirst, zfirst.data(), sizeof(int) * zfirst.size()));
        EHIPCHK(e, up(e->dline, lines.data(), sizeof(lpt::EnsLine) * lines.size()));
        EHIPCHK(e, up(e->dlfirst, lfirst.data(), sizeof(int) * lfirst.size()));
        EHIPCHK(e, hipStreamSynchronize(e->stream));  // the host vectors go out of scope
    }
    e->npart = np;
    e->pblk = acc;
    e->zblk = zacc;
    e->zchunks = zchunks;
    for (int ax = 0; ax < 2; ax++) e->lblk[ax] = lacc[ax];
    e->psig = std::move(sig);
    return GKM_OK;
}

int ens_parity(gkm_ensemble* e, int* p)
{
    for (int m = 0; m < e->n; m++) {
        if (!e->mem[m]->have_state) return efail(e, GKM_ENOTREADY, "member " + std::to_string(m) + " has no state");
        if (e->mem[m]->cur != e->mem[0]->cur)
            return efail(e, GKM_EINVAL, "members were stepped separately; their buffers are out of phase");
    }
    *p = e->mem[0]->cur;
    return GKM_OK;
}

}  // namespace

extern "C" {

int gkm_ensemble_create(int n, const gkm_config* cfgs, gkm_ensemble** out)
{
    if (!out) return GKM_EINVAL;
    *out = nullptr;
    if (n < 1 || n > 4096 || !cfgs) return GKM_EINVAL;
    auto* e = new (std::nothrow) gkm_ensemble;
    if (!e) return GKM_ENOMEM;
    *out = e;
    e->n = n;
    e->mixed = cfgs[0].precision == GKM_MIXED;
    for (int m = 0; m < n; m++) {
        if (cfgs[m].precision != cfgs[0].precision || cfgs[m].device != cfgs[0].device)
            return efail(e, GKM_EINVAL, "members must share precision and device");
        gkm_ctx* c = nullptr;
        const int rc = gkm_create(&cfgs[m], &c);
        if (rc) {
            const std::string msg = c ? c->err : "gkm_create failed";
            gkm_destroy(c);
            return efail(e, rc, "member " + std::to_string(m) + ": " + msg);
        }
        c->member = true;
        e->mem.push_back(c);
    }
    EHIPCHK(e, hipSetDevice(cfgs[0].device));
    EHIPCHK(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    EHIPCHK(e, hipEventCreate(&e->ev[0]));
    EHIPCHK(e, hipEventCreate(&e->ev[1]));
    if (hipMalloc(&e->dargs, sizeof(gkm::SweepArgs) * 6 * (size_t)n) != hipSuccess ||
        hipMalloc(&e->dfirst, sizeof(int) * 3 * (size_t)(n + 1)) != hipSuccess ||
        hipMalloc(&e->dcfl, sizeof(gkm::CflArgs) * 2 * (size_t)n) != hipSuccess ||
        hipMalloc(&e->partial, sizeof(double) * gkm::kEnsCflBlocks * (size_t)n) != hipSuccess)
        return efail(e, GKM_ENOMEM, "hipMalloc ensemble tables");
    int t0[3];
    for (int d = 0; d < 3; d++) {
        t0[d] = -1;
        for (int v = 0; v < gkm::kTileVariants && t0[d] < 0; v++) {
            bool ok = true;
            for (gkm_ctx* c : e->mem) ok = ok && gkm::sweep_tile_ok(d, v, c->nc);
            if (ok) t0[d] = v;
        }
        if (t0[d] < 0) return efail(e, GKM_EINVAL, "no sweep tile fits every member's grid");
    }
    return ens_build(e, t0);
}

void gkm_ensemble_destroy(gkm_ensemble* e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    for (gkm_ctx* c : e->mem) gkm_destroy(c);
    (void)hipFree(e->dargs);
    (void)hipFree(e->dfirst);
    (void)hipFree(e->dcfl);
    (void)hipFree(e->partial);
    for (void* ptr : {(void*)e->dpart, (void*)e->dpfirst, (void*)e->dz, (void*)e->dzfirst, (void*)e->dline,
                      (void*)e->dlfirst, (void*)e->dbar})
        (void)hipFree(ptr);
    if (e->ev[0]) (void)hipEventDestroy(e->ev[0]);
    if (e->ev[1]) (void)hipEventDestroy(e->ev[1]);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* gkm_ensemble_last_error(const gkm_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

gkm_ctx* gkm_ensemble_member(gkm_ensemble* e, int m)
{
    if (!e || m < 0 || m >= e->n) return nullptr;
    return e->mem[m];
}

int gkm_ensemble_set_tiles(gkm_ensemble* e, const int tile[3])
{
    if (!e || !tile) return efail(e, GKM_EINVAL, "null argument");
    for (int d = 0; d < 3; d++)
        for (gkm_ctx* c : e->mem)
            if (!gkm::sweep_tile_ok(d, tile[d], c->nc))
                return efail(e, GKM_EINVAL, "tile variant out of range or larger than a member's grid");
    return ens_build(e, tile);
}

int gkm_ensemble_autotune(gkm_ensemble* e, int reps, int tile[3], double ms[9])
{
    if (!e || reps < 1) return efail(e, GKM_EINVAL, "bad argument");
    int p = 0;
    int rc = ens_parity(e, &p);
    if (rc) return rc;
    const int n = e->n;
    EHIPCHK(e, gkm::launch_ens_cfl(e->dcfl + (size_t)p * n, n, e->partial, e->stream));  // every member's dt
    int best[3];
    for (int d = 0; d < 3; d++) {
        double best_ms = 0.0;
        best[d] = e->tile[d];
        bool first = true;
        for (int v = 0; v < gkm::kTileVariants; v++) {
            bool ok = true;
            for (gkm_ctx* c : e->mem) ok = ok && gkm::sweep_tile_ok(d, v, c->nc);
            if (!ok) {
                if (ms) ms[3 * d + v] = -1.0;
                continue;
            }
            int t[3] = {e->tile[0], e->tile[1], e->tile[2]};
            t[d] = v;
            rc = ens_build(e, t);
            if (rc) return rc;
            // dW is scratch and Wout the next state's buffer: the members' states are not changed
            auto launch = [&] {
                return gkm::launch_ens_sweep(d, v, e->dargs + ((size_t)p * 3 + d) * n, e->dfirst + (size_t)d * (n + 1),
                                             n, e->nblk[d], e->mixed, e->exact[d], e->compensated, e->stream);
            };
            EHIPCHK(e, launch());
            EHIPCHK(e, hipEventRecord(e->ev[0], e->stream));
            for (int r = 0; r < reps; r++) EHIPCHK(e, launch());
            EHIPCHK(e, hipEventRecord(e->ev[1], e->stream));
            EHIPCHK(e, hipEventSynchronize(e->ev[1]));
            float el = 0.f;
            EHIPCHK(e, hipEventElapsedTime(&el, e->ev[0], e->ev[1]));
            const double avg = el / reps;
            if (ms) ms[3 * d + v] = avg;
            if (first || avg < best_ms) {
                best_ms = avg;
                best[d] = v;
                first = false;
            }
        }
        int t[3] = {e->tile[0], e->tile[1], e->tile[2]};
        t[d] = best[d];
        rc = ens_build(e, t);
        if (rc) return rc;
    }
    if (tile)
        for (int d = 0; d < 3; d++) tile[d] = e->tile[d];
    return GKM_OK;
}

int gkm_ensemble_run(gkm_ensemble* e, int nsteps, int use_graph)
{
    if (!e || nsteps < 0) return efail(e, GKM_EINVAL, "bad argument");
    int p = 0;
    int rc = ens_parity(e, &p);
    if (rc) return rc;
    rc = ens_particles_prepare(e, p);
    if (rc) return rc;
    EHIPCHK(e, hipDeviceSynchronize());  // member streams (splines, uploads) are done before the ensemble stream runs
    if (use_graph) {
        for (int q = 0; q < 2; q++) {
            if (e->graph[q]) continue;
            hipGraph_t g = nullptr;
            EHIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t r = ens_step(e, q, true);
            const hipError_t r2 = hipStreamEndCapture(e->stream, &g);
            if (r != hipSuccess || r2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return efail(e, GKM_EHIP, std::string("graph capture: ") + hipGetErrorString(r != hipSuccess ? r : r2));
            }
            const hipError_t r3 = hipGraphInstantiate(&e->graph[q], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (r3 != hipSuccess) {
                e->graph[q] = nullptr;
                return efail(e, GKM_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(r3));
            }
        }
    }
    for (int s = 0; s < nsteps; s++) {
        if (use_graph)
            EHIPCHK(e, hipGraphLaunch(e->graph[p], e->stream));
        else
            EHIPCHK(e, ens_step(e, p, false));
        p = 1 - p;
    }
    EHIPCHK(e, hipStreamSynchronize(e->stream));
    if (e->persistent) {
        int timed_out = 0;
        EHIPCHK(e, hipMemcpy(&timed_out, e->dbar + 2, sizeof(int), hipMemcpyDeviceToHost));
        if (timed_out) {
            // the counter, base and work counters of the failed launches are
            // stale: start them afresh (graph replays read the same device words)
            EHIPCHK(e, hipMemset(e->dbar, 0, gkm::kPersistSyncWords * sizeof(unsigned long long)));
            return efail(e, GKM_EHIP, "persistent step: a grid barrier timed out (workgroups not all resident)");
        }
    }
    std::vector<int> flags(e->n, 0);
    for (int m = 0; m < e->n; m++) {
        gkm_ctx* c = e->mem[m];
        c->cur = p;
        lpt::State& L = c->lpt;
        if (L.np > 0 && nsteps > 0) {  // the last step left the spline of W[p] in coef[p] and U, G at its start
            L.cur = p;
            L.valid = true;
            L.sampled = true;
        } else if (nsteps > 0) {
            L.valid = false;
        }
        EHIPCHK(e, hipMemcpy(&flags[m], c->dflag, sizeof(int), hipMemcpyDeviceToHost));
    }
    for (int m = 0; m < e->n; m++)
        if (flags[m] && e->mem[m]->cfg.check_state)
            return efail(e, GKM_ESTATE, "member " + std::to_string(m) + ": non-physical state (rho <= 0, p <= 0 or NaN)");
    return GKM_OK;
}

int gkm_ensemble_set_persistent(gkm_ensemble* e, int on)
{
    if (!e) return GKM_EINVAL;
    if (on && !e->dbar) {
        const int dev = e->mem[0]->cfg.device;
        e->pgrid = gkm::persist_grid(e->mixed, dev);
        if (e->pgrid < 1) return efail(e, GKM_EHIP, "persistent step: occupancy query failed");
        int khz = 0;
        EHIPCHK(e, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) return efail(e, GKM_EHIP, "persistent step: no wall-clock rate");
        e->timeout_ticks = 2ull * 1000ull * (unsigned long long)khz;  // 2 s per barrier wait
        if (hipMalloc(&e->dbar, gkm::kPersistSyncWords * sizeof(unsigned long long)) != hipSuccess)
            return efail(e, GKM_ENOMEM, "hipMalloc barrier");
        EHIPCHK(e, hipMemset(e->dbar, 0, gkm::kPersistSyncWords * sizeof(unsigned long long)));
    }
    if (e->persistent != (on != 0)) drop_graphs(e);
    e->persistent = on != 0;
    return GKM_OK;
}

int gkm_ensemble_persistent_profile(gkm_ensemble* e, double ms[9], int* nphases)
{
    if (!e || !ms || !nphases) return efail(e, GKM_EINVAL, "null argument");
    if (!e->persistent) return efail(e, GKM_EINVAL, "the ensemble is not in persistent mode");
    int p = 0;
    int rc = ens_parity(e, &p);
    if (rc) return rc;
    rc = ens_particles_prepare(e, p);
    if (rc) return rc;
    EHIPCHK(e, hipDeviceSynchronize());
    unsigned long long* dtr = nullptr;
    if (hipMalloc(&dtr, 10 * sizeof(unsigned long long)) != hipSuccess) return efail(e, GKM_ENOMEM, "hipMalloc trace");
    gkm::PersistStep a = persist_args(e, p);
    a.trace = dtr;
    hipError_t r = gkm::launch_persist_step(a, e->mixed, e->pgrid, true, e->stream);
    unsigned long long t[10] = {0};
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    if (r == hipSuccess) r = hipMemcpy(t, dtr, sizeof(t), hipMemcpyDeviceToHost);
    (void)hipFree(dtr);
    if (r != hipSuccess) return efail(e, GKM_EHIP, std::string("persistent profile: ") + hipGetErrorString(r));
    // the traced step advanced every member by one step
    for (gkm_ctx* c : e->mem) {
        c->cur = 1 - p;
        if (c->lpt.np > 0) {
            c->lpt.cur = 1 - p;
            c->lpt.valid = true;
            c->lpt.sampled = true;
        } else {
            c->lpt.valid = false;
        }
    }
    int khz = 0;
    EHIPCHK(e, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->mem[0]->cfg.device));
    const int nb = e->npart ? 9 : 5;
    for (int k = 0; k < nb; k++) ms[k] = (double)(t[k + 1] - t[k]) / (double)khz;
    *nphases = nb;
    return GKM_OK;
}

int gkm_ensemble_dt(gkm_ensemble* e, double* dt)
{
    if (!e || !dt) return efail(e, GKM_EINVAL, "null argument");
    for (int m = 0; m < e->n; m++) EHIPCHK(e, hipMemcpy(dt + m, e->mem[m]->dscal, sizeof(double), hipMemcpyDeviceToHost));
    return GKM_OK;
}

}  // extern "C"
