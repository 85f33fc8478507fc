// gkm_kernels.hip -- gfx950 kernels of the GKM-WENO Eulerian step.
//
// One fused kernel per sweep direction replaces the reference's derivs*_Kernel,
// flip_Kernel, flux, flipBack_Kernel and update_Kernel chain
// (GPUupdate.cu:119-202): WENO states of every face a workgroup needs are
// built once into LDS, each thread then evaluates one GKS face flux, and the
// flux differences are formed in LDS and accumulated per cell.  Interface
// states and fluxes never round-trip HBM.
#include "gkm_kernels.hpp"
#include "gkm_sweep.hpp"

namespace gkm {

template <int DIR, bool MIX, int V, bool GEN, int CMP>
__global__ __launch_bounds__(256, 2) void sweep_kernel(SweepArgs p)
{
    sweep_body<DIR, MIX, V, false, GEN, CMP>(p, blockIdx.x, gridDim.x, true);
}

// One launch for the sweeps of every member of an ensemble: member m owns the
// blocks [first[m], first[m+1]) of the grid, in launch order.  Unlike the
// single-grid sweep there is no XCD-grouped reorder: members' workgroups
// differ in length (a 64^3 pencil has twice the chunks of a 32^3 one), and
// contiguous ranges per XCD would hand one XCD the long ones; dealt round-robin,
// every XCD gets the same mix (64 members at 32^3-64^3: 5.34 against 5.90 ms
// per step).
template <int DIR, bool MIX, int V, bool GEN, int CMP>
__global__ __launch_bounds__(256, 2) void ens_sweep_kernel(const SweepArgs* __restrict__ members,
                                                          const int* __restrict__ first, int nmem)
{
    const int g = blockIdx.x;
    int lo = 0, hi = nmem;  // first[lo] <= g < first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= g) lo = mid;
        else hi = mid;
    }
    const SweepArgs p = members[lo];
    sweep_body<DIR, MIX, V, false, GEN, CMP>(p, g - first[lo], first[lo + 1] - first[lo], false);
}

static int cdiv(int a, int b) { return (a + b - 1) / b; }

bool sweep_tile_ok(int dir, int tile, int nc)
{
    // Any grid size: the pencils and segments of the last tiles reach past nc
    // (those lines and faces are computed with wrapped cells and not written).
    // Every wrapped index must stay in [-nc, 2nc), the range wrapi handles:
    // stencils reach CN + 2 cells before a segment and past the grid's end,
    // pencils one line past a tile.
    if (dir < 0 || dir > 2 || tile < 0 || tile >= kTileVariants) return false;
    const int CN = kTiles[dir][tile][0], CA = kTiles[dir][tile][1], CB = kTiles[dir][tile][2];
    return CN + 3 <= nc && CA <= nc && CB <= nc;
}

int sweep_first_tile(int dir, int nc)
{
    for (int v = 0; v < kTileVariants; v++)
        if (sweep_tile_ok(dir, v, nc)) return v;
    return -1;
}

bool sweep_grid_ok(int nc)
{
    return nc >= 16 && sweep_first_tile(0, nc) >= 0 && sweep_first_tile(1, nc) >= 0 && sweep_first_tile(2, nc) >= 0;
}

int sweep_seg_len(int dir, int nc, int tile, int slots)
{
    // faces per workgroup along the sweep axis (a multiple of CN); a segment
    // costs seg/CN chunks plus one lead-in chunk (the face before the segment).
    if (!sweep_tile_ok(dir, tile, nc)) return 0;
    const int CN = kTiles[dir][tile][0], CA = kTiles[dir][tile][1], CB = kTiles[dir][tile][2];
    const long pencils = (long)cdiv(nc, CA) * cdiv(nc, CB);
    const int lchunks = cdiv(nc, CN);  // chunks of a whole line
    if (slots > 0) {
        // one grid: the length with the fewest chunk-times, counting rounds of
        // `slots` concurrent workgroups; the longer segment on ties
        int best = lchunks * CN;
        long best_cost = -1;
        for (int parts = 1;; parts *= 2) {
            const int seg = cdiv(lchunks, parts) * CN;
            const long blocks = pencils * cdiv(nc, seg);
            const long cost = (blocks + slots - 1) / slots * (seg / CN + 1);
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                best = seg;
            }
            if (seg == CN) break;
        }
        return best;
    }
    // members of an ensemble (their workgroups add up): the whole line when
    // that already gives >= 1024 workgroups (4 per CU), otherwise halve while
    // segments stay >= 8 chunks and >= 64 faces long.
    const int min_seg = 8 * CN > 64 ? 8 * CN : 64;
    int seg = lchunks * CN;
    while (pencils * cdiv(nc, seg) < 1024) {
        const int half = cdiv(seg / CN, 2) * CN;
        if (half >= seg || half < min_seg) break;
        seg = half;
    }
    return seg;
}

template <int DIR, int V>
static hipError_t launch_tile(const SweepArgs& a, bool mixed, hipStream_t s)
{
    constexpr int CN = Tile<DIR, V>::CN, CA = Tile<DIR, V>::CA, CB = Tile<DIR, V>::CB;
    if (!sweep_tile_ok(DIR, V, a.nc) || a.seg <= 0 || a.seg % CN || a.nseg != cdiv(a.nc, a.seg))
        return hipErrorInvalidValue;
    const long nblk = (long)cdiv(a.nc, CA) * cdiv(a.nc, CB) * a.nseg;
    // the compensated update is the mixed z sweep's (an FP64 context has no carry)
    if ((a.cmp_out != nullptr) != (a.cmp_in != nullptr) || (a.cmp_out && !mixed)) return hipErrorInvalidValue;
    const bool gen = !sweep_exact(DIR, V, a.nc, a.seg);  // some workgroup reaches past the grid
    const bool cmp = a.cmp_out && a.mode == 2;
    const dim3 grid((unsigned)nblk), block(CN * CA * CB);
    if (mixed) {
        if constexpr (DIR == 2) {
            if (cmp) {
                if (gen) hipLaunchKernelGGL((sweep_kernel<DIR, true, V, true, 1>), grid, block, 0, s, a);
                else hipLaunchKernelGGL((sweep_kernel<DIR, true, V, false, 1>), grid, block, 0, s, a);
                return hipGetLastError();
            }
        }
        if (gen) hipLaunchKernelGGL((sweep_kernel<DIR, true, V, true, 0>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((sweep_kernel<DIR, true, V, false, 0>), grid, block, 0, s, a);
    } else {
        if (gen) hipLaunchKernelGGL((sweep_kernel<DIR, false, V, true, 0>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((sweep_kernel<DIR, false, V, false, 0>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

template <int DIR>
static hipError_t launch_dir(const SweepArgs& a, bool mixed, hipStream_t s)
{
    static_assert(kTileVariants == 3, "one case per tile variant");
    switch (a.tile) {
    case 0: return launch_tile<DIR, 0>(a, mixed, s);
    case 1: return launch_tile<DIR, 1>(a, mixed, s);
    case 2: return launch_tile<DIR, 2>(a, mixed, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sweep(int dir, const SweepArgs& a, bool mixed, hipStream_t s)
{
    if (dir == 0) return launch_dir<0>(a, mixed, s);
    if (dir == 1) return launch_dir<1>(a, mixed, s);
    if (dir == 2) return launch_dir<2>(a, mixed, s);
    return hipErrorInvalidValue;
}

int sweep_blocks(int dir, int nc, int tile, int seg)
{
    if (!sweep_tile_ok(dir, tile, nc) || seg <= 0) return 0;
    const int CA = kTiles[dir][tile][1], CB = kTiles[dir][tile][2];
    return cdiv(nc, CA) * cdiv(nc, CB) * cdiv(nc, seg);
}

bool sweep_exact(int dir, int tile, int nc, int seg)
{
    if (!sweep_tile_ok(dir, tile, nc) || seg <= 0) return false;
    return nc % kTiles[dir][tile][1] == 0 && nc % kTiles[dir][tile][2] == 0 && nc % seg == 0;
}

template <int DIR, int V>
static hipError_t launch_ens_tile(const SweepArgs* members, const int* first, int nmem, int nblk, bool mixed,
                                  bool exact, bool cmp, hipStream_t s)
{
    const dim3 grid((unsigned)nblk), block(256);
    if (mixed) {
        if constexpr (DIR == 2) {
            if (cmp) {
                if (exact) hipLaunchKernelGGL((ens_sweep_kernel<DIR, true, V, false, 1>), grid, block, 0, s, members, first, nmem);
                else hipLaunchKernelGGL((ens_sweep_kernel<DIR, true, V, true, 1>), grid, block, 0, s, members, first, nmem);
                return hipGetLastError();
            }
        }
        if (exact) hipLaunchKernelGGL((ens_sweep_kernel<DIR, true, V, false, 0>), grid, block, 0, s, members, first, nmem);
        else hipLaunchKernelGGL((ens_sweep_kernel<DIR, true, V, true, 0>), grid, block, 0, s, members, first, nmem);
    } else {
        if (exact) hipLaunchKernelGGL((ens_sweep_kernel<DIR, false, V, false, 0>), grid, block, 0, s, members, first, nmem);
        else hipLaunchKernelGGL((ens_sweep_kernel<DIR, false, V, true, 0>), grid, block, 0, s, members, first, nmem);
    }
    return hipGetLastError();
}

template <int DIR>
static hipError_t launch_ens_dir(int tile, const SweepArgs* members, const int* first, int nmem, int nblk, bool mixed,
                                 bool exact, bool cmp, hipStream_t s)
{
    switch (tile) {
    case 0: return launch_ens_tile<DIR, 0>(members, first, nmem, nblk, mixed, exact, cmp, s);
    case 1: return launch_ens_tile<DIR, 1>(members, first, nmem, nblk, mixed, exact, cmp, s);
    case 2: return launch_ens_tile<DIR, 2>(members, first, nmem, nblk, mixed, exact, cmp, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ens_sweep(int dir, int tile, const SweepArgs* members, const int* first, int nmem, int nblk,
                            bool mixed, bool exact, bool compensated, hipStream_t s)
{
    if (nmem <= 0 || nblk <= 0 || (compensated && !mixed)) return hipErrorInvalidValue;
    if (dir == 0) return launch_ens_dir<0>(tile, members, first, nmem, nblk, mixed, exact, compensated, s);
    if (dir == 1) return launch_ens_dir<1>(tile, members, first, nmem, nblk, mixed, exact, compensated, s);
    if (dir == 2) return launch_ens_dir<2>(tile, members, first, nmem, nblk, mixed, exact, compensated, s);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// CFL: dt = cfl dx / max(|U| + sqrt(gam p / rho)) over all cells
// (timestep, update.cpp:1031-1054).  A block maximum goes to partial[]; a
// non-physical cell (rho <= 0, p <= 0 or not finite) records -1 instead.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void cfl_partial_kernel(const double* __restrict__ W, long plane, double gam,
                                                          double* partial)
{
    cfl_cells(W, plane, gam, (long)blockIdx.x * blockDim.x, (long)gridDim.x * blockDim.x, partial + blockIdx.x);
}

__global__ __launch_bounds__(256) void cfl_final_kernel(const double* partial, int n, double cfl, double dx,
                                                        double* dt, double* vmax, int* bad)
{
    cfl_reduce(partial, n, cfl, dx, dt, vmax, bad);
}

hipError_t launch_cfl(const double* W, long plane, double gam, double cfl, double dx, double* partial, double* dt,
                      double* vmax, int* bad, hipStream_t s)
{
    hipLaunchKernelGGL(cfl_partial_kernel, dim3(kCflBlocks), dim3(256), 0, s, W, plane, gam, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cfl_final_kernel, dim3(1), dim3(256), 0, s, partial, kCflBlocks, cfl, dx, dt, vmax, bad);
    return hipGetLastError();
}

// Ensemble CFL: kEnsCflBlocks blocks per member reduce max(|U|+c) of that
// member's cells into partial[m * kEnsCflBlocks + b]; one block per member then
// writes dt = cfl dx / max (0 and the flag when a state is non-physical).
__global__ __launch_bounds__(256) void ens_cfl_partial_kernel(const CflArgs* __restrict__ members, double* partial)
{
    const int m = blockIdx.x / kEnsCflBlocks, b = blockIdx.x % kEnsCflBlocks;
    const CflArgs a = members[m];
    cfl_cells(a.W, a.plane, a.gam, (long)b * blockDim.x, (long)kEnsCflBlocks * blockDim.x, partial + blockIdx.x);
}

__global__ __launch_bounds__(256) void ens_cfl_final_kernel(const CflArgs* __restrict__ members, const double* partial)
{
    const CflArgs a = members[blockIdx.x];
    cfl_reduce(partial + (long)blockIdx.x * kEnsCflBlocks, kEnsCflBlocks, a.cfl, a.dx, a.dt, a.vmax, a.bad);
}

hipError_t launch_ens_cfl(const CflArgs* members, int nmem, double* partial, hipStream_t s)
{
    if (nmem <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ens_cfl_partial_kernel, dim3((unsigned)(nmem * kEnsCflBlocks)), dim3(256), 0, s, members,
                       partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ens_cfl_final_kernel, dim3((unsigned)nmem), dim3(256), 0, s, members, partial);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Layout conversion.  pack fills the host layout's 3 ghost layers per face
// periodically: ghost(i) = cell(((i-3) mod nc) + 3) on every axis (the
// single-domain result of HaloTransfer, update.cpp:855-1029).
// ---------------------------------------------------------------------------
__global__ void unpack_kernel(const double* __restrict__ padded, double* __restrict__ W, int nc)
{
    const long plane = (long)nc * nc * nc;
    const int nt = nc + 6;
    const long ntp = (long)nt * nt * nt;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < 5 * plane; c += (long)gridDim.x * blockDim.x) {
        const int q = (int)(c / plane);
        const long r = c - q * plane;
        const int k = (int)(r % nc), i = (int)((r / nc) % nc), j = (int)(r / ((long)nc * nc));
        W[c] = padded[q * ntp + ((long)(j + 3) * nt + (i + 3)) * nt + (k + 3)];
    }
}

__global__ void pack_kernel(const double* __restrict__ W, double* __restrict__ padded, int nc)
{
    const long plane = (long)nc * nc * nc;
    const int nt = nc + 6;
    const long ntp = (long)nt * nt * nt;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < 5 * ntp; c += (long)gridDim.x * blockDim.x) {
        const int q = (int)(c / ntp);
        const long r = c - q * ntp;
        const int k = (int)(r % nt), i = (int)((r / nt) % nt), j = (int)(r / ((long)nt * nt));
        const int jj = (j - 3 + nc) % nc, ii = (i - 3 + nc) % nc, kk = (k - 3 + nc) % nc;
        padded[c] = W[q * plane + ((long)jj * nc + ii) * nc + kk];
    }
}

static unsigned grid_for(long n, int block)
{
    long g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

hipError_t launch_unpack(const double* padded, double* W, int nc, hipStream_t s)
{
    const long n = 5L * nc * nc * nc;
    hipLaunchKernelGGL(unpack_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, padded, W, nc);
    return hipGetLastError();
}

hipError_t launch_pack(const double* W, double* padded, int nc, hipStream_t s)
{
    const long nt = nc + 6;
    const long n = 5L * nt * nt * nt;
    hipLaunchKernelGGL(pack_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, W, padded, nc);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pointwise kernels for parity checks of the device math
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void flux_batch_kernel(long n, const double* __restrict__ in, const double* __restrict__ tau, double dt,
                                  double dx, Gas<T> g, double* __restrict__ F)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* v = in + 50 * i;
    T WL[5], WR[5], Wl[5], Wr[5], D[6][5], Fo[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        WL[q] = (T)v[q];
        WR[q] = (T)v[5 + q];
        Wl[q] = (T)v[10 + q];
        Wr[q] = (T)v[15 + q];
#pragma unroll
        for (int d = 0; d < 6; d++) D[d][q] = (T)v[20 + 5 * d + q];
    }
    gks_flux_tf<T, T>(WL, WR, Wl, Wr, D, (T)tau[i], (T)dt, (T)dx, g, Fo);
#pragma unroll
    for (int q = 0; q < 5; q++) F[5 * i + q] = (double)Fo[q];
}

__global__ __launch_bounds__(64) void flux_batch_mixed_kernel(long n, const double* __restrict__ in,
                                                              const double* __restrict__ tau, double dt, double dx,
                                                              Gas<double> g, double* __restrict__ F)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* v = in + 50 * i;
    double WL[5], WR[5], Wl[5], Wr[5], D[6][5], Fo[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        WL[q] = v[q];
        WR[q] = v[5 + q];
        Wl[q] = v[10 + q];
        Wr[q] = v[15 + q];
#pragma unroll
        for (int d = 0; d < 6; d++) D[d][q] = v[20 + 5 * d + q];
    }
    gks_flux_tf<double, float>(WL, WR, Wl, Wr, D, tau[i], dt, dx, g, Fo);
#pragma unroll
    for (int q = 0; q < 5; q++) F[5 * i + q] = Fo[q];
}

template <typename T>
__global__ __launch_bounds__(64) void weno_batch_kernel(long n, const double* __restrict__ in, Gas<T> g, double* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T S[6][5], Wl[5], Wr[5];
#pragma unroll
    for (int s = 0; s < 6; s++)
#pragma unroll
        for (int q = 0; q < 5; q++) S[s][q] = (T)in[30 * i + 5 * s + q];
    weno5<T>(S, g, Wl, Wr);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        out[10 * i + q] = (double)Wl[q];
        out[10 * i + 5 + q] = (double)Wr[q];
    }
}

__global__ __launch_bounds__(64) void weno_batch_mixed_kernel(long n, const double* __restrict__ in, Gas<double> g,
                                                              double* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double S[6][5], Wl[5], Wr[5];
#pragma unroll
    for (int s = 0; s < 6; s++)
#pragma unroll
        for (int q = 0; q < 5; q++) S[s][q] = in[30 * i + 5 * s + q];
    weno5_mixed(S, g, Wl, Wr);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        out[10 * i + q] = Wl[q];
        out[10 * i + 5 + q] = Wr[q];
    }
}


hipError_t launch_flux_batch(long n, const double* in, const double* tau, double dt, double dx, const Gas<double>& g,
                             int precision, double* F, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 63) / 64);
    if (precision == 0)
        hipLaunchKernelGGL(flux_batch_kernel<double>, dim3(blocks), dim3(64), 0, s, n, in, tau, dt, dx, g, F);
    else if (precision == 1)
        hipLaunchKernelGGL(flux_batch_mixed_kernel, dim3(blocks), dim3(64), 0, s, n, in, tau, dt, dx, g, F);
    else
        hipLaunchKernelGGL(flux_batch_kernel<float>, dim3(blocks), dim3(64), 0, s, n, in, tau, dt, dx, to_f32(g), F);
    return hipGetLastError();
}

hipError_t launch_weno_batch(long n, const double* in, const Gas<double>& g, int precision, double* out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 63) / 64);
    if (precision == 0)
        hipLaunchKernelGGL(weno_batch_kernel<double>, dim3(blocks), dim3(64), 0, s, n, in, g, out);
    else if (precision == 1)
        hipLaunchKernelGGL(weno_batch_mixed_kernel, dim3(blocks), dim3(64), 0, s, n, in, g, out);
    else
        hipLaunchKernelGGL(weno_batch_kernel<float>, dim3(blocks), dim3(64), 0, s, n, in, to_f32(g), out);
    return hipGetLastError();
}

}  // namespace gkm
