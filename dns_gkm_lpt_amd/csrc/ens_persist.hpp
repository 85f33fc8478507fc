// ens_persist.hpp -- the ensemble step as one persistent kernel (no reference
// counterpart).
#pragma once
#include <hip/hip_runtime.h>

#include "gkm_kernels.hpp"
#include "lpt_kernels.hpp"

namespace gkm {

// phases of the persistent step that hand out their items dynamically
constexpr int kPersistPhases = 6;
// device words of the persistent step's synchronisation: barrier counter,
// base, timeout flag (64-bit words), then the 2 kPersistPhases 32-bit work
// counters and tallies
constexpr int kPersistSyncWords = 3 + kPersistPhases;

// Everything one step of every member reads, for one buffer parity.  The
// tables are the ones the per-phase launches use (gkm_api.hip builds them).
struct PersistStep {
    const SweepArgs* sweep[3];  // [dir]: member arguments
    const int* sfirst[3];       // [dir]: first workgroup of each member, nmem + 1 entries
    int nsweep[3];              // workgroups (items) per direction
    int tile[3];                // tile variant per direction
    const CflArgs* cfl;         // [member]
    double* partial;            // nmem * kEnsCflBlocks block maxima
    int nmem;
    // tracers of the members that hold particles (npart entries; 0: none)
    const lpt::EnsPart* sample;
    const lpt::EnsPart* heun;
    const int* pfirst;
    int npart, pblk;
    const lpt::ZArgs* z;
    const int* zfirst;
    int zblk, zchunks;
    const lpt::EnsLine* lines;  // 3 npart entries (u, v, w of each)
    const int* lfirst;          // [axis - 1][3 npart + 1]
    int lblk[2];
    // grid barrier: arrivals so far (monotonic), its value when the launch
    // started, the timeout flag, and the wait limit in wall-clock ticks
    unsigned long long* count;
    unsigned long long* base;
    int* timeout;
    unsigned long long timeout_ticks;
    // work counters of the phases whose items differ in cost (x, y, z sweeps;
    // z, x, y passes of the splines) and, from work + kPersistPhases, the
    // number of workgroups that have run out of each phase's items; both zero
    // between launches
    unsigned int* work;
    // optional: workgroup 0 writes the wall clock at the start and after every
    // barrier (10 entries), for per-phase timings
    unsigned long long* trace;
};

// workgroups of the persistent kernel that are resident at once (0: unknown)
int persist_grid(bool mixed, int device);
// cooperative: launch with hipLaunchCooperativeKernel (not inside a graph capture)
hipError_t launch_persist_step(const PersistStep& a, bool mixed, int grid, bool cooperative, hipStream_t s);

}  // namespace gkm
