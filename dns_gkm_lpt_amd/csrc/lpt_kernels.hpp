// lpt_kernels.hpp -- natural tricubic B-spline prefilter, particle sampling /
// Heun advection and flow statistics for gfx950.
#pragma once
#include <hip/hip_runtime.h>

namespace lpt {

// Device particle and spline state owned by a gkm context.
struct State {
    long np = 0;
    double* pos = nullptr;     // np x 3 (y, x, z), tag order
    double* U = nullptr;       // np x 3 velocity sampled at the start of the step
    double* G = nullptr;       // np x 9 velocity gradients gU, gV, gW over (y, x, z)
    double* coef[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};  // (nc+8)^3 each
    // full-property splines of the resident state {rho, p, lap u, lap v, lap w}, built on demand
    double* pcoef[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    const double** pcs = nullptr;  // device array of the 8 spline pointers for props_kernel
    double* props = nullptr;       // np x kPropWidth, tag order
    int cur = 0;               // coef[cur] = spline of the resident state when valid
    bool valid = false;
    double* pivots = nullptr;  // Thomas pivots 1/(4 - p_{r-1}) of the [1 4 1] system
    double* red = nullptr;     // reduction scratch
    bool sampled = false;
    // cell sort: particles are kept ordered by the cell holding them, tag[i]
    // is the upload index of the particle stored in slot i
    int* tag = nullptr;
    int* tag2 = nullptr;
    double* pos2 = nullptr;    // np x 9 scratch (gathers, tag-order outputs)
    unsigned* keys = nullptr;  // 2 x np
    int* perm = nullptr;       // 2 x np
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int sort_every = 10;       // re-sort after this many Heun steps
    int since_sort = 0;
    int drop = 0;              // reference tracer ownership (Geom::drop)
};

void release(State& s);

// Spline geometry on the padded cell-centre grid (Interpolation.cpp:3-31):
// nt = nc+6 knots per axis from start = -3dx + dx/2, spacing delta.
struct Geom {
    int nc, nt, N;      // N = nt + 2 coefficients per axis
    double start, dinv; // 1/delta
    // 1: the reference's single-rank tracer ownership -- a tracer whose block
    // id int(pos / 2 pi) is not 0 on some axis (after the wrap: a coordinate
    // of exactly 2 pi) is dropped (its position set to NaN) as evolveParticles
    // pops it (update.cpp:1056-1109, ParticleTrack.cpp:21-40); dropped tracers
    // sample zero velocity and gradients and are not advanced
    int drop;
};

// field q of the unpadded state W (0 = rho, 1..3 = velocity component,
// 4 = pressure) -> coefficients of its natural tricubic spline on the padded
// grid (periodic ghosts read by wrap).
hipError_t launch_prefilter(const double* W, long plane, int q, double gam, const Geom& g, const double* pivots,
                            double* coef, hipStream_t s);
// the three velocity components at once (rho read once per knot)
hipError_t launch_prefilter_velocity(const double* W, long plane, double gam, const Geom& g, const double* pivots,
                                     double* const coef[3], hipStream_t s);
// spline of the Laplacian of spline c (knot values, then prefilter in place)
hipError_t launch_lap_spline(const double* c, const Geom& g, const double* pivots, double* lap, hipStream_t s);
// full particle properties, kPropWidth doubles per particle written to row tag[p];
// cs_dev is a device array of the 8 spline pointers {rho, u, v, w, p, lap u, lap v, lap w}
constexpr int kPropWidth = 41;
hipError_t launch_props(long np, const double* pos, const int* tag, const double* const* cs_dev, const Geom& g,
                        double* out, hipStream_t s);
hipError_t launch_sample(long np, const double* pos, const double* const cu[3], const Geom& g, double* U, double* G,
                         hipStream_t s);
hipError_t launch_heun(long np, double* pos, const double* U, const double* const cn[3], const Geom& g,
                       const double* dt, hipStream_t s);

// Reorder the particles by cell index (y, x, z cell of dx) with a radix sort;
// pos and tag follow, U and G are resampled after a sort.
hipError_t sort_by_cell(State& s, const Geom& g, double dx, hipStream_t st);
// dst[tag[i]] = src[i] for rows of `width` doubles (slot order -> tag order)
hipError_t scatter_to_tags(long np, const int* tag, const double* src, int width, double* dst, hipStream_t st);

// ---- batched over the members of an ensemble (no reference counterpart) ----
// One launch covers every entry; entry e owns blocks [first[e], first[e+1]),
// and its block runs the same body as the single-context kernel.
// z pass of the state's splines: nf = 1 (field q0 into coef[0]) or nf = 3
// (q0 = 1: u, v, w into coef[0..2])
struct ZArgs {
    const double* W;
    long plane;
    int q0, nf;
    double gam;
    Geom g;
    const double* piv;
    double* coef[3];
};
struct EnsLine {  // x and y passes of one spline, in place
    Geom g;
    const double* piv;
    double* coef;
};
struct EnsPart {  // one member's tracers
    long np;
    double* pos;
    double* U;
    double* G;
    const double* c[3];  // velocity spline to sample (sample) or of the new field (Heun)
    Geom g;
    const double* dt;
};
// blocks of one entry: z pass (256-thread blocks, a z line per wave), x and y
// passes (64, axis 1 or 2), particles (256 threads, four per tracer: 64
// tracers); ens_z_chunks(largest nt) is the z kernel's chunk count (0: nt too
// large)
int ens_z_blocks(const Geom& g);
int ens_xy_blocks(const Geom& g, int axis);
int ens_z_chunks(int nt_max);
int ens_particle_blocks(long np);
// z pass of every ZArgs entry, then the x and y passes of every EnsLine entry;
// lfirst[(axis - 1) * (nl + 1) + e]
hipError_t launch_ens_prefilter(const ZArgs* zs, const int* zfirst, int nz, int zblk, int zchunks,
                                const EnsLine* lines, const int* lfirst, int nl, const int lblk[2], hipStream_t s);
hipError_t launch_ens_sample(const EnsPart* parts, const int* first, int n, int nblk, hipStream_t s);
hipError_t launch_ens_heun(const EnsPart* parts, const int* first, int n, int nblk, hipStream_t s);

// flow statistics (PrintStats, update.cpp:688-842); red holds >= kStatBlocks*32 doubles
constexpr int kStatBlocks = 512;
struct StatParams {
    double gam, R, mu0, T0;
};
hipError_t launch_stats(const double* W, long plane, const double* const cu[3], const Geom& g, const StatParams& sp,
                        double* red, hipStream_t s);
// second pass: squared deviations from the means (P, T, 1/rho) into red[32*b + 16..18]
hipError_t launch_stats2(const double* W, long plane, const StatParams& sp, double pm, double tm, double vm,
                         double* red, hipStream_t s);
// particle statistics (GetParticleStats, update.cpp:1111-1162); red >= kStatBlocks*8
hipError_t launch_pstats(long np, const double* U, const double* G, double* red, hipStream_t s);

}  // namespace lpt
