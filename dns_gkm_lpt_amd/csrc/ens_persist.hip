// ens_persist.hip -- one step of every ensemble member in one persistent
// kernel launch (no reference counterpart).
//
// The grid holds as many workgroups as the GPU keeps resident at once.  Each
// phase of the step is a list of work items -- the workgroups the per-phase
// launches would run.  Items whose cost differs (sweep pencils of members of
// different grid sizes; spline line groups) are handed out one at a time from
// a counter per phase, the next to whichever workgroup is free, as the
// hardware dispatcher does for the per-phase launches; a static split into
// contiguous ranges gave the workgroups that drew the 64^3 members' pencils
// twice the work of the others (sweep phases 1.9-2.1 ms against 1.1-1.3 ms
// per launch).  Uniform items (CFL blocks, members' dt, tracer blocks) keep
// the contiguous range [n b / G, n (b + 1) / G) of workgroup b.  Phases are
// separated by a grid barrier:
//   1  CFL block maxima of every member; tracer velocities at x^n
//   2  dt of every member
//   3  x sweeps   4  y sweeps   5  z sweeps (with the update)
//   6  z pass of the velocity splines of W^{n+1}   7  x pass   8  y pass
//   9  Heun update of the tracers
// (6-9 only when members hold tracers).  An item runs exactly the code of the
// per-phase kernels, so results equal theirs bit for bit.
//
// Grid barrier: every workgroup adds one to a monotonic counter per barrier
// and waits until it reaches base + k G (base: the counter when the launch
// started, republished by workgroup 0 after the last barrier).  Eager steps
// are cooperative launches, which fail up front when the grid cannot be
// resident at once; graph-captured steps are ordinary launches of the same
// occupancy-sized grid.  Each wait is bounded by a wall-clock limit: on expiry
// the workgroup sets the timeout flag and carries on, every other workgroup
// stops waiting as soon as it sees the flag, later launches of the run return
// at entry, and the host reports the error and resets the barrier.
#include "ens_persist.hpp"
#include "gkm_sweep.hpp"
#include "lpt_device.hpp"

namespace gkm {

namespace {

__device__ __forceinline__ int entry_of(const int* __restrict__ first, int n, int i)
{
    int lo = 0, hi = n - 1;  // largest e with first[e] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Each wave first waits until its own stores have reached L2; one wave per
// workgroup then arrives with an agent-scope release (the XCD's L2 is written
// back before the counter moves), polls the counter with relaxed loads, and
// after the wait invalidates this CU's L1 and the L2 (agent-scope acquire)
// before the workgroup continues.  Once any workgroup has timed out (the flag
// is set) the others stop waiting: they neither arrive nor wait at this or any
// later barrier, so a failed launch ends after one wait limit.
// Returns true (in every thread of the workgroup) when the launch has failed.
__device__ __forceinline__ bool grid_barrier(const PersistStep& a, unsigned long long target, int* failed)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0 && __hip_atomic_load(a.timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // L2 write-back
        __builtin_amdgcn_s_waitcnt(0);                      // ... completed before the arrival is seen
        __hip_atomic_fetch_add(a.count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long t0 = wall_clock64();
        while (__hip_atomic_load(a.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (__hip_atomic_load(a.timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (wall_clock64() - t0 > a.timeout_ticks) {
                __hip_atomic_store(a.timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (threadIdx.x == 0) *failed = __hip_atomic_load(a.timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return *failed != 0;
}

// The next item of phase `ph` for the whole workgroup, or -1 once its n items
// are handed out.  Thread 0 takes it from the phase's counter; the workgroup
// that is the last of the G to run out resets the counter and the tally for
// the next launch (every other workgroup has by then made its last take).
__device__ __forceinline__ int next_item(const PersistStep& a, int ph, int n, int* slot)
{
    if (threadIdx.x == 0) {
        unsigned int* ctr = a.work + ph;
        unsigned int* tally = a.work + kPersistPhases + ph;
        const int i = (int)__hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (i >= n &&
            __hip_atomic_fetch_add(tally, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
            __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(tally, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        *slot = i;
    }
    __syncthreads();
    const int i = *slot;
    __syncthreads();  // every thread has read the slot before thread 0 rewrites it
    return i < n ? i : -1;
}

template <int DIR, bool MIX, int V>
__device__ __forceinline__ void sweep_items_v(const PersistStep& a, int* slot, double* lds)
{
    for (int i; (i = next_item(a, DIR, a.nsweep[DIR], slot)) >= 0;) {
        const int m = entry_of(a.sfirst[DIR], a.nmem, i);
        const SweepArgs p = a.sweep[DIR][m];
        const int f0 = a.sfirst[DIR][m], nb = a.sfirst[DIR][m + 1] - f0;
        sweep_body<DIR, MIX, V, true>(p, i - f0, nb, false, lds);
        __syncthreads();  // the next item rewrites the LDS buffer
    }
}

template <int DIR, bool MIX>
__device__ __forceinline__ void sweep_items(const PersistStep& a, int* slot, double* lds)
{
    switch (a.tile[DIR]) {
    case 0: sweep_items_v<DIR, MIX, 0>(a, slot, lds); break;
    case 1: sweep_items_v<DIR, MIX, 1>(a, slot, lds); break;
    default: sweep_items_v<DIR, MIX, 2>(a, slot, lds); break;
    }
}

template <int K>
__device__ __forceinline__ void zpass_item(const PersistStep& a, int i)
{
    const int e = lpt::ens_entry(a.zfirst, a.npart, i);
    const lpt::ZArgs& z = a.z[e];
    const long line = (long)(i - a.zfirst[e]) * 4 + (threadIdx.x >> 6);
    if (line < (long)z.g.nt * z.g.nt) lpt::prefilter_z_line<K>(line, z);  // uniform over the wave
}

template <bool MIX>
__global__ __launch_bounds__(256, 2) void ens_step_kernel(PersistStep a)
{
    __shared__ double lds[kSweepLdsMax];
    __shared__ int failed, slot;
    // a launch after a timed-out one (same run, flag not yet cleared by the
    // host) does nothing; thread 0 reads the flag for the whole workgroup
    if (threadIdx.x == 0) failed = __hip_atomic_load(a.timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (failed) return;
    const int G = gridDim.x;
    // uniform phases: consecutive workgroups take consecutive slices
    const int lb = blockIdx.x;
    auto range = [&](int n, int& lo, int& hi) {
        lo = (int)((long)n * lb / G);
        hi = (int)((long)n * (lb + 1) / G);
    };
    const unsigned long long base = __hip_atomic_load(a.base, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    int nbar = 0;
    if (a.trace && blockIdx.x == 0 && threadIdx.x == 0) a.trace[0] = wall_clock64();
    // true: a barrier of this launch timed out, and the workgroup ends here
    auto sync = [&] {
        if (grid_barrier(a, base + (unsigned long long)(++nbar) * G, &failed)) return true;
        if (a.trace && blockIdx.x == 0 && threadIdx.x == 0) a.trace[nbar] = wall_clock64();
        return false;
    };
    const int tid = threadIdx.x;
    int lo, hi;

    // 1: CFL block maxima; tracer velocities and gradients at x^n
    range(a.nmem * kEnsCflBlocks, lo, hi);
    for (int i = lo; i < hi; i++) {
        const int m = i / kEnsCflBlocks, b = i % kEnsCflBlocks;
        const CflArgs c = a.cfl[m];
        cfl_cells(c.W, c.plane, c.gam, (long)b * 256, (long)kEnsCflBlocks * 256, a.partial + i);
    }
    range(a.npart ? a.pblk : 0, lo, hi);
    for (int i = lo; i < hi; i++) {
        const int e = lpt::ens_entry(a.pfirst, a.npart, i);
        const lpt::EnsPart& P = a.sample[e];
        const long t = (long)(i - a.pfirst[e]) * 256 + tid;  // four threads per tracer
        if ((t >> 2) < P.np) lpt::sample_quad(t >> 2, tid & 3, P.pos, P.c[0], P.c[1], P.c[2], P.g, P.U, P.G);
    }
    if (sync()) return;
    // 2: dt of every member
    range(a.nmem, lo, hi);
    for (int m = lo; m < hi; m++) {
        const CflArgs c = a.cfl[m];
        cfl_reduce(a.partial + (long)m * kEnsCflBlocks, kEnsCflBlocks, c.cfl, c.dx, c.dt, c.vmax, c.bad);
    }
    if (sync()) return;
    // 3-5: sweeps
    sweep_items<0, MIX>(a, &slot, lds);
    if (sync()) return;
    sweep_items<1, MIX>(a, &slot, lds);
    if (sync()) return;
    sweep_items<2, MIX>(a, &slot, lds);
    if (a.npart) {
        // 6: z pass of the splines of u, v, w of W^{n+1}
        if (sync()) return;
        for (int i; (i = next_item(a, 3, a.zblk, &slot)) >= 0;) {
            switch (a.zchunks) {
            case 1: zpass_item<1>(a, i); break;
            case 2: zpass_item<2>(a, i); break;
            case 3: zpass_item<3>(a, i); break;
            case 5: zpass_item<5>(a, i); break;
            case 9: zpass_item<9>(a, i); break;
            default: zpass_item<lpt::kMaxZChunks>(a, i); break;
            }
        }
        // 7, 8: x and y passes; an item is four 64-line groups, one per wave
        for (int axis = 1; axis <= 2; axis++) {
            if (sync()) return;
            const int nsub = a.lblk[axis - 1];
            const int* first = a.lfirst + (axis - 1) * (3 * a.npart + 1);
            for (int i; (i = next_item(a, 3 + axis, (nsub + 3) / 4, &slot)) >= 0;) {
                const int sub = 4 * i + (tid >> 6);
                if (sub >= nsub) continue;
                const int e = lpt::ens_entry(first, 3 * a.npart, sub);
                const lpt::EnsLine& L = a.lines[e];
                lpt::prefilter_inplace_one((long)(sub - first[e]) * 64 + (tid & 63), L.coef, L.g, axis, L.piv);
            }
        }
        // 9: Heun
        if (sync()) return;
        range(a.pblk, lo, hi);
        for (int i = lo; i < hi; i++) {
            const int e = lpt::ens_entry(a.pfirst, a.npart, i);
            const lpt::EnsPart& P = a.heun[e];
            const long t = (long)(i - a.pfirst[e]) * 256 + tid;  // four threads per tracer
            if ((t >> 2) < P.np) lpt::heun_quad(t >> 2, tid & 3, P.pos, P.U, P.c[0], P.c[1], P.c[2], P.g, *P.dt);
        }
    }
    if (a.trace) {  // the end of the last phase
        if (sync()) return;
    }
    // every workgroup reached the last barrier, so every one has read base:
    // publish the base of the next launch (the counter ends at base + nbar G)
    if (blockIdx.x == 0 && tid == 0)
        __hip_atomic_store(a.base, base + (unsigned long long)nbar * G, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

int persist_grid(bool mixed, int device)
{
    int per_cu = 0, cus = 0;
    const void* k = mixed ? reinterpret_cast<const void*>(&ens_step_kernel<true>)
                          : reinterpret_cast<const void*>(&ens_step_kernel<false>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    int g = per_cu * cus;
    if (g >= 8) g -= g % 8;
    return g;
}

hipError_t launch_persist_step(const PersistStep& a, bool mixed, int grid, bool cooperative, hipStream_t s)
{
    if (grid < 1 || a.nmem < 1) return hipErrorInvalidValue;
    const void* k = mixed ? reinterpret_cast<const void*>(&ens_step_kernel<true>)
                          : reinterpret_cast<const void*>(&ens_step_kernel<false>);
    if (cooperative) {
        // fails up front when the grid cannot be resident at once
        PersistStep arg = a;
        void* params[] = {&arg};
        return hipLaunchCooperativeKernel(k, dim3(grid), dim3(256), params, 0, s);
    }
    if (mixed)
        hipLaunchKernelGGL(ens_step_kernel<true>, dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(ens_step_kernel<false>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gkm
