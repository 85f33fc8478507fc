// gkm_kernels.hpp -- launch interface of the gfx950 kernels of the Eulerian step.
#pragma once
#include <hip/hip_runtime.h>
#include "gkm_flux.hpp"

namespace gkm {

// Device state: 5 planes (rho, rho u, rho v, rho w, rho E) of nc^3 doubles,
// unpadded and periodic, cell (y,x,z) at (y*nc + x)*nc + z.  Ghost layers exist
// only in the host layout; the kernels wrap indices instead.
struct SweepArgs {
    const double* W;     // state at t^n
    double* dW;          // sum of flux differences of the sweeps done so far
    double* Wout;        // state at t^{n+1} (written by the last sweep)
    // Compensated accumulation of the update (mode 2), or nullptr for the plain
    // FP64 update: cmp_in holds, per cell and variable, the rounding error the
    // previous step's update left out of W (exact sum = W + cmp_in); the update
    // adds it to the increment and stores its own error in cmp_out (TwoSum).
    // Layout: carry_plane(nc) floats per variable, cell at carry_index.
    const float* cmp_in;
    float* cmp_out;
    const double* dt;    // device scalar (graph-capturable)
    long plane;          // nc^3
    int nc;
    int seg;             // faces per workgroup along the sweep axis
    int nseg;            // nc / seg
    int mode;            // 0: dW = dF   1: dW += dF   2: Wout = W - (dW + dF)/dx
    int tile;            // tile shape variant, 0 .. kTileVariants-1
    double dx, inv_dx;
    Gas<double> g;
};

// Layout of the compensated update's FP32 carry, which only the z sweep's
// update reads and writes: blocks of 16 cells along z, within a block the
// x-y plane x-major, cell (y, x, z) at ((z / 16 nc + x) nc + y) 16 + z mod 16.
// A wave of the 16 x 4 x 4 z tile updates 16 z cells of 4 lines of
// consecutive y, so each of its carry loads and stores is one 256 B run (the
// state's layout gives four 64 B runs, halves of lines that the chunk along z
// after it fills).  Blocks past nc (nc not a multiple of 16) leave gaps.
__host__ __device__ __forceinline__ long carry_plane(int nc) { return (long)nc * nc * ((nc + 15) & ~15); }
__host__ __device__ __forceinline__ long carry_index(int y, int x, int z, int nc)
{
    return (((long)(z >> 4) * nc + x) * nc + y) * 16 + (z & 15);
}

// Sweep tile shapes: CN faces along the sweep axis x CA x CB tangential lines
// (256 threads, one face each).  A chunk reconstructs CN x (CA CB + 2 CA + 2 CB)
// WENO states for CN x CA x CB fluxes; the tangential edge lines are the price
// of the y/z slopes.  x sweep: a = y, b = z; y sweep: a = x, b = z; z sweep:
// a = x, b = y.
constexpr int kTileVariants = 3;
constexpr int kTiles[3][kTileVariants][3] = {
    {{4, 4, 16}, {1, 16, 16}, {2, 8, 16}},
    {{4, 4, 16}, {1, 16, 16}, {2, 8, 16}},
    {{16, 4, 4}, {4, 8, 8}, {8, 8, 4}},
};
template <int DIR, int V>
struct Tile {
    static constexpr int CN = kTiles[DIR][V][0], CA = kTiles[DIR][V][1], CB = kTiles[DIR][V][2];
};

hipError_t launch_sweep(int dir, const SweepArgs& a, bool mixed, hipStream_t s);
// blocks of one grid's sweep launch for a tile variant and segment length
int sweep_blocks(int dir, int nc, int tile, int seg);
// one launch sweeping every member of an ensemble: members[m] (device array) with
// blocks [first[m], first[m+1]) of nblk = first[nmem]; one tile variant for all.
// exact: the tile and segment divide every member's grid (sweep_exact);
// compensated: every member's z sweep takes the compensated update (mixed
// members), else none does.
hipError_t launch_ens_sweep(int dir, int tile, const SweepArgs* members, const int* first, int nmem, int nblk,
                            bool mixed, bool exact, bool compensated, hipStream_t s);
// the tile variant and the segment length divide nc: no workgroup reaches past the grid
bool sweep_exact(int dir, int tile, int nc, int seg);
// faces per workgroup chosen for a grid: slots = concurrent workgroups of the
// device for a single grid, 0 for the members of an ensemble
int sweep_seg_len(int dir, int nc, int tile, int slots);
// any nc in [16, 1024] has a tile variant per direction; a variant fits nc when
// CN + 3 <= nc and CA, CB <= nc
bool sweep_grid_ok(int nc);
bool sweep_tile_ok(int dir, int tile, int nc);
int sweep_first_tile(int dir, int nc);           // the first variant that fits, or -1
// segments along the sweep axis: nseg = ceil(nc / seg); the last may reach past nc
inline int sweep_nseg(int nc, int seg) { return seg > 0 ? (nc + seg - 1) / seg : 0; }

// CFL: partial maxima (one per block) then a one-block final pass writing
// dt = cfl*dx/max and a non-physical-state flag.
constexpr int kCflBlocks = 1024;
hipError_t launch_cfl(const double* W, long plane, double gam, double cfl, double dx, double* partial, double* dt,
                      double* vmax, int* bad, hipStream_t s);
// ensemble CFL: one dt per member (partial holds nmem * kEnsCflBlocks doubles)
struct CflArgs {
    const double* W;
    long plane;
    double gam, cfl, dx;
    double *dt, *vmax;
    int* bad;
};
constexpr int kEnsCflBlocks = 64;
hipError_t launch_ens_cfl(const CflArgs* members, int nmem, double* partial, hipStream_t s);

// padded (nt^3 per plane, ghosts) <-> unpadded (nc^3 per plane)
hipError_t launch_unpack(const double* padded, double* W, int nc, hipStream_t s);
hipError_t launch_pack(const double* W, double* padded, int nc, hipStream_t s);

hipError_t launch_flux_batch(long n, const double* in, const double* tau, double dt, double dx, const Gas<double>& g,
                             int precision, double* F, hipStream_t s);
hipError_t launch_weno_batch(long n, const double* in, const Gas<double>& g, int precision, double* out,
                             hipStream_t s);

}  // namespace gkm
