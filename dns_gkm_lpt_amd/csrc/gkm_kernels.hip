// gkm_kernels.hip -- gfx950 kernels of the GKM-WENO Eulerian step.
//
// One fused kernel per sweep direction replaces the reference's derivs*_Kernel,
// flip_Kernel, flux, flipBack_Kernel and update_Kernel chain
// (GPUupdate.cu:119-202): WENO states of every face a workgroup needs are
// built once into LDS, each thread then evaluates one GKS face flux, and the
// flux differences are formed in LDS and accumulated per cell.  Interface
// states and fluxes never round-trip HBM.
#include "gkm_kernels.hpp"

namespace gkm {

__device__ __forceinline__ int wrapi(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

// physical strides: y -> nc*nc, x -> nc, z -> 1.
// DIR 0 (x sweep): n = x, a = y, b = z.   DWy pairs with a, DWz with b.
// DIR 1 (y sweep): n = y, a = x, b = z.
// DIR 2 (z sweep): n = z, a = x, b = y (the reference's tangential pairing
//                  of derivsZ, update.cpp:402-455, kept as is).
template <int DIR>
__device__ __forceinline__ long cell_index(int cn, int ca, int cb, int nc)
{
    if (DIR == 0) return ((long)ca * nc + cn) * nc + cb;
    if (DIR == 1) return ((long)cn * nc + ca) * nc + cb;
    return ((long)cb * nc + ca) * nc + cn;
}

// Lines of the plus-shaped set a workgroup reconstructs: the CA x CB interior
// lines plus one line beyond each tangential edge (no corners).
template <int CA, int CB>
struct Lines {
    static constexpr int L = CA * CB + 2 * CA + 2 * CB;
    __device__ static __forceinline__ int id(int ia, int ib)
    {
        if (ia < 0) return CA * CB + ib;
        if (ia >= CA) return CA * CB + CB + ib;
        if (ib < 0) return CA * CB + 2 * CB + ia;
        if (ib >= CB) return CA * CB + 2 * CB + CA + ia;
        return ia * CB + ib;
    }
    __device__ static __forceinline__ void coords(int l, int& ia, int& ib)
    {
        if (l < CA * CB) { ia = l / CB; ib = l % CB; return; }
        l -= CA * CB;
        if (l < CB) { ia = -1; ib = l; return; }
        l -= CB;
        if (l < CB) { ia = CA; ib = l; return; }
        l -= CB;
        if (l < CA) { ia = l; ib = -1; return; }
        l -= CA;
        ia = l;
        ib = CB;
    }
};

// LDS layout of the WENO states (10 values per face and line) and of the face
// fluxes (5 values per face slot and interior line).  The fastest thread index
// is contiguous: the tangential line for the x and y sweeps, the face for the z
// sweep (whose lines are short and whose faces are consecutive in memory).
template <int DIR, int CN, int L>
__device__ __forceinline__ int sw_at(int v, int tn, int l)
{
    return DIR == 2 ? (v * L + l) * CN + tn : (v * CN + tn) * L + l;
}
template <int DIR, int CN, int NL>
__device__ __forceinline__ int sf_at(int q, int slot, int lc)
{
    return DIR == 2 ? (q * NL + lc) * (CN + 1) + slot : (q * (CN + 1) + slot) * NL + lc;
}

// Flux inputs of one face, read where the flux uses them: cell states from the
// (cached) state array, WENO states and their tangential differences from LDS.
// q is the rotated-frame slot (normal momentum in slot 1; update.cpp:361-398, 461-497).
template <int DIR, int CN, int L>
struct SweepIn {
    const double* __restrict__ W;
    long plane, iL, iR;
    const double* s;
    int tn, lc, l_am, l_ap, l_bm, l_bp;
    double two_idx, half_idx;
    __device__ static int phys(int q)
    {
        constexpr int sw = DIR == 1 ? 2 : 3;
        return DIR == 0 ? q : (q == 1 ? sw : (q == sw ? 1 : q));
    }
    __device__ double at(int v, int l) const { return s[sw_at<DIR, CN, L>(v, tn, l)]; }
    __device__ double WL(int q) const { return W[phys(q) * plane + iL]; }
    __device__ double WR(int q) const { return W[phys(q) * plane + iR]; }
    __device__ double Wl(int q) const { return at(phys(q), lc); }
    __device__ double Wr(int q) const { return at(5 + phys(q), lc); }
    __device__ double D(int d, int q) const
    {
        const int p = phys(q);
        switch (d) {
        case 0: return (at(p, lc) - WL(q)) * two_idx;                 // DWxl
        case 1: return (WR(q) - at(5 + p, lc)) * two_idx;             // DWxr
        case 2: return (at(p, l_ap) - at(p, l_am)) * half_idx;        // DWyl
        case 3: return (at(5 + p, l_ap) - at(5 + p, l_am)) * half_idx;  // DWyr
        case 4: return (at(p, l_bp) - at(p, l_bm)) * half_idx;        // DWzl
        default: return (at(5 + p, l_bp) - at(5 + p, l_bm)) * half_idx;  // DWzr
        }
    }
};

// MIX = false: the FP64 path.  MIX = true: WENO in FP32 on FP64 stencil
// differences and the flux with FP32 slope terms (gks_flux_tf<double, float>).
// Two waves per SIMD are requested: the register budget is then 256 per lane.
// With one wave per SIMD the compiler (ROCm 7.2, gfx950) spilled into AGPRs and
// the x and y instances produced wrong flux differences on the GPU (the z
// instance and every instance built this way match the FP64 oracle to 1e-14).
//
// sweep_body is the work of one workgroup: block `bid` of the `nblk` blocks of
// one grid (member).  remap = true applies the XCD-aware order to bid.
template <int DIR, bool MIX, int V>
__device__ __forceinline__ void sweep_body(const SweepArgs& p, int bid, int nblk, bool remap)
{
    constexpr int CN = Tile<DIR, V>::CN, CA = Tile<DIR, V>::CA, CB = Tile<DIR, V>::CB;
    constexpr int NT = CN * CA * CB;
    static_assert(NT == 256, "every tile variant runs 256 threads, one face each");
    using LN = Lines<CA, CB>;
    constexpr int L = LN::L;
    __shared__ double sW[10 * CN * L];          // WENO states: Wl (0..4), Wr (5..9)
    __shared__ double sF[5 * (CN + 1) * CA * CB]; // face fluxes; slot 0 = face before the chunk

    const int nc = p.nc;
    const long plane = p.plane;
    const double* __restrict__ W = p.W;

    // workgroup -> (pencil, segment); XCD-aware: the hardware deals workgroups
    // round-robin over the 8 XCDs, so give each XCD a contiguous range of pencils.
    if (remap && (nblk & 7) == 0) bid = (bid & 7) * (nblk >> 3) + (bid >> 3);
    const int s = bid % p.nseg;
    const int pencil = bid / p.nseg;
    const int npb = nc / CB;
    const int a0 = (pencil / npb) * CA;
    const int b0 = (pencil % npb) * CB;

    const int tid = threadIdx.x;
    int tn, ta, tb;
    if (DIR == 2) {
        tn = tid % CN;
        tb = (tid / CN) % CB;
        ta = tid / (CN * CB);
    } else {
        tb = tid % CB;
        ta = (tid / CB) % CA;
        tn = tid / (CA * CB);
    }
    const int lc = ta * CB + tb;

    const double dt = *p.dt;
    const double dx = p.dx;
    const int nchunk = p.seg / CN;
    const int seg0 = s * p.seg;

    SweepIn<DIR, CN, L> in;
    in.W = W;
    in.plane = plane;
    in.s = sW;
    in.tn = tn;
    in.lc = lc;
    in.l_am = LN::id(ta - 1, tb);
    in.l_ap = LN::id(ta + 1, tb);
    in.l_bm = LN::id(ta, tb - 1);
    in.l_bp = LN::id(ta, tb + 1);
    in.two_idx = 2.0 * p.inv_dx;
    in.half_idx = 0.5 * p.inv_dx;

    for (int c = -1; c < nchunk; ++c) {
        const int n0 = seg0 + c * CN;  // first face of the chunk (may be < 0 for the lead-in)

        // ---- WENO states of every (face, line) of the chunk into LDS
        for (int i = tid; i < CN * L; i += NT) {
            int itn, il;
            if (DIR == 2) { itn = i % CN; il = i / CN; }
            else { il = i % L; itn = i / L; }
            int ia, ib;
            LN::coords(il, ia, ib);
            const int ca = wrapi(a0 + ia, nc), cb = wrapi(b0 + ib, nc);
            const int fn = n0 + itn;
            double S[6][5];
#pragma unroll
            for (int t = 0; t < 6; t++) {
                const long idx = cell_index<DIR>(wrapi(fn - 2 + t, nc), ca, cb, nc);
#pragma unroll
                for (int q = 0; q < 5; q++) S[t][q] = W[q * plane + idx];
            }
            double Wl[5], Wr[5];
            if (MIX)
                weno5_mixed(S, p.g, Wl, Wr);
            else
                weno5<double>(S, p.g, Wl, Wr);
#pragma unroll
            for (int q = 0; q < 5; q++) {
                sW[sw_at<DIR, CN, L>(q, itn, il)] = Wl[q];
                sW[sw_at<DIR, CN, L>(5 + q, itn, il)] = Wr[q];
            }
        }
        __syncthreads();

        // ---- one GKS flux per thread
        {
            const int fn = n0 + tn;
            const int ca = a0 + ta, cb = b0 + tb;  // interior lines: in range
            in.iL = cell_index<DIR>(wrapi(fn, nc), ca, cb, nc);
            in.iR = cell_index<DIR>(wrapi(fn + 1, nc), ca, cb, nc);
            double tau;
            {
                double WL[5];
#pragma unroll
                for (int q = 0; q < 5; q++) WL[q] = W[q * plane + in.iL];
                tau = collision_time<double>(WL, p.g);  // symmetric in the momentum slots
            }
            double F[5];
            if (MIX)
                gks_flux_tf<double, float>(in, tau, dt, dx, p.g, F);
            else
                gks_flux_tf<double, double>(in, tau, dt, dx, p.g, F);
#pragma unroll
            for (int q = 0; q < 5; q++) sF[sf_at<DIR, CN, CA * CB>(in.phys(q), tn + 1, lc)] = F[q];
        }
        __syncthreads();

        // ---- flux differences of the chunk's cells
        if (c >= 0) {
            const long idx = cell_index<DIR>(n0 + tn, a0 + ta, b0 + tb, nc);
#pragma unroll
            for (int q = 0; q < 5; q++) {
                const double dF = sF[sf_at<DIR, CN, CA * CB>(q, tn + 1, lc)] - sF[sf_at<DIR, CN, CA * CB>(q, tn, lc)];
                const long o = q * plane + idx;
                if (p.mode == 0) p.dW[o] = dF;
                else if (p.mode == 1) p.dW[o] += dF;
                else p.Wout[o] = W[o] - p.inv_dx * (p.dW[o] + dF);
            }
        }
        __syncthreads();
        if (tn == CN - 1) {
#pragma unroll
            for (int q = 0; q < 5; q++) sF[sf_at<DIR, CN, CA * CB>(q, 0, lc)] = sF[sf_at<DIR, CN, CA * CB>(q, CN, lc)];
        }
        // the next write of sF slots 1..CN happens after the next chunk's first barrier
    }
}

template <int DIR, bool MIX, int V>
__global__ __launch_bounds__(256, 2) void sweep_kernel(SweepArgs p)
{
    sweep_body<DIR, MIX, V>(p, blockIdx.x, gridDim.x, true);
}

// One launch for the sweeps of every member of an ensemble: member m owns the
// blocks [first[m], first[m+1]) of the grid (after the XCD-aware reorder of
// the whole grid, so each XCD gets a contiguous range of members' pencils).
template <int DIR, bool MIX, int V>
__global__ __launch_bounds__(256, 2) void ens_sweep_kernel(const SweepArgs* __restrict__ members,
                                                          const int* __restrict__ first, int nmem)
{
    const int nblk = gridDim.x;
    int g = blockIdx.x;
    if ((nblk & 7) == 0) g = (g & 7) * (nblk >> 3) + (g >> 3);
    int lo = 0, hi = nmem;  // first[lo] <= g < first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= g) lo = mid;
        else hi = mid;
    }
    const SweepArgs p = members[lo];
    sweep_body<DIR, MIX, V>(p, g - first[lo], first[lo + 1] - first[lo], false);
}

bool sweep_grid_ok(int nc)
{
    // multiples of 16 fit every tile shape; nc >= 32 keeps every wrapped index
    // (faces of the lead-in chunk reach 16 + 2 cells before the segment) in
    // [-nc, 2nc), the range wrapi handles.
    return nc >= 32 && nc % 16 == 0;
}

int sweep_seg_len(int dir, int nc, int tile)
{
    // faces per workgroup: the whole line when that already gives >= 1024
    // workgroups (4 per CU), otherwise split; each split costs one lead-in
    // chunk, so segments stay >= 8 chunks and >= 64 faces long.
    if (dir < 0 || dir > 2 || tile < 0 || tile >= kTileVariants) return 0;
    const int CN = kTiles[dir][tile][0], CA = kTiles[dir][tile][1], CB = kTiles[dir][tile][2];
    const long pencils = (long)(nc / CA) * (nc / CB);
    const int min_seg = 8 * CN > 64 ? 8 * CN : 64;
    int seg = nc;
    while (pencils * (nc / seg) < 1024 && seg % 2 == 0 && (seg / 2) % CN == 0 && seg / 2 >= min_seg) seg /= 2;
    return seg;
}

template <int DIR, int V>
static hipError_t launch_tile(const SweepArgs& a, bool mixed, hipStream_t s)
{
    constexpr int CN = Tile<DIR, V>::CN, CA = Tile<DIR, V>::CA, CB = Tile<DIR, V>::CB;
    if (a.nc % CA || a.nc % CB || a.seg % CN || a.nc % a.seg || a.nseg != a.nc / a.seg) return hipErrorInvalidValue;
    const long nblk = (long)(a.nc / CA) * (a.nc / CB) * a.nseg;
    if (mixed)
        hipLaunchKernelGGL((sweep_kernel<DIR, true, V>), dim3((unsigned)nblk), dim3(CN * CA * CB), 0, s, a);
    else
        hipLaunchKernelGGL((sweep_kernel<DIR, false, V>), dim3((unsigned)nblk), dim3(CN * CA * CB), 0, s, a);
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
    if (dir < 0 || dir > 2 || tile < 0 || tile >= kTileVariants || seg <= 0) return 0;
    const int CA = kTiles[dir][tile][1], CB = kTiles[dir][tile][2];
    return (nc / CA) * (nc / CB) * (nc / seg);
}

template <int DIR, int V>
static hipError_t launch_ens_tile(const SweepArgs* members, const int* first, int nmem, int nblk, bool mixed,
                                  hipStream_t s)
{
    if (mixed)
        hipLaunchKernelGGL((ens_sweep_kernel<DIR, true, V>), dim3((unsigned)nblk), dim3(256), 0, s, members, first, nmem);
    else
        hipLaunchKernelGGL((ens_sweep_kernel<DIR, false, V>), dim3((unsigned)nblk), dim3(256), 0, s, members, first, nmem);
    return hipGetLastError();
}

template <int DIR>
static hipError_t launch_ens_dir(int tile, const SweepArgs* members, const int* first, int nmem, int nblk, bool mixed,
                                 hipStream_t s)
{
    switch (tile) {
    case 0: return launch_ens_tile<DIR, 0>(members, first, nmem, nblk, mixed, s);
    case 1: return launch_ens_tile<DIR, 1>(members, first, nmem, nblk, mixed, s);
    case 2: return launch_ens_tile<DIR, 2>(members, first, nmem, nblk, mixed, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ens_sweep(int dir, int tile, const SweepArgs* members, const int* first, int nmem, int nblk,
                            bool mixed, hipStream_t s)
{
    if (nmem <= 0 || nblk <= 0) return hipErrorInvalidValue;
    if (dir == 0) return launch_ens_dir<0>(tile, members, first, nmem, nblk, mixed, s);
    if (dir == 1) return launch_ens_dir<1>(tile, members, first, nmem, nblk, mixed, s);
    if (dir == 2) return launch_ens_dir<2>(tile, members, first, nmem, nblk, mixed, s);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// CFL: dt = cfl dx / max(|U| + sqrt(gam p / rho)) over all cells
// (timestep, update.cpp:1031-1054).  A block maximum goes to partial[]; a
// non-physical cell (rho <= 0, p <= 0 or not finite) records -1 instead.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ __launch_bounds__(256) void cfl_partial_kernel(const double* __restrict__ W, long plane, double gam,
                                                          double* partial)
{
    double m = 0.0;
    bool bad = false;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < plane; c += (long)gridDim.x * blockDim.x) {
        const double den = W[c];
        const double ir = 1.0 / den;
        const double u = W[plane + c] * ir, v = W[2 * plane + c] * ir, w = W[3 * plane + c] * ir;
        const double Vsqr = u * u + v * v + w * w;
        const double P = (den * (gam - 1)) * (W[4 * plane + c] * ir - 0.5 * Vsqr);
        const double val = sqrt(Vsqr) + sqrt(gam * P * ir);
        if (!(den > 0.0) || !(P > 0.0) || !isfinite(val)) bad = true;
        else m = fmax(m, val);
    }
    __shared__ double sm[4];
    __shared__ int sbad;
    if (threadIdx.x == 0) sbad = 0;
    __syncthreads();
    if (bad) sbad = 1;
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
        partial[blockIdx.x] = sbad ? -1.0 : r;
    }
}

// block reduction of n partial maxima (-1 = non-physical state seen) -> dt
__device__ __forceinline__ void cfl_reduce(const double* partial, int n, double cfl, double dx, double* dt, double* vmax,
                                           int* bad)
{
    double m = 0.0;
    int b = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = partial[i];
        if (v < 0.0) b = 1;
        else m = fmax(m, v);
    }
    __shared__ double sm[4];
    __shared__ int sb[4];
    m = wave_max(m);
    // any-reduction of the flag
    unsigned long long ballot = __ballot(b);
    if ((threadIdx.x & 63) == 0) {
        sm[threadIdx.x >> 6] = m;
        sb[threadIdx.x >> 6] = ballot != 0ull;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double r = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
        const int anyb = sb[0] | sb[1] | sb[2] | sb[3];
        *vmax = r;
        *bad = anyb;
        *dt = (anyb || !(r > 0.0)) ? 0.0 : cfl * dx / r;
    }
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
    double mx = 0.0;
    bool bad = false;
    for (long c = (long)b * blockDim.x + threadIdx.x; c < a.plane; c += (long)kEnsCflBlocks * blockDim.x) {
        const double* W = a.W;
        const long plane = a.plane;
        const double den = W[c];
        const double ir = 1.0 / den;
        const double u = W[plane + c] * ir, v = W[2 * plane + c] * ir, w = W[3 * plane + c] * ir;
        const double Vsqr = u * u + v * v + w * w;
        const double P = (den * (a.gam - 1)) * (W[4 * plane + c] * ir - 0.5 * Vsqr);
        const double val = sqrt(Vsqr) + sqrt(a.gam * P * ir);
        if (!(den > 0.0) || !(P > 0.0) || !isfinite(val)) bad = true;
        else mx = fmax(mx, val);
    }
    __shared__ double sm[4];
    __shared__ int sbad;
    if (threadIdx.x == 0) sbad = 0;
    __syncthreads();
    if (bad) sbad = 1;
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double r = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
        partial[blockIdx.x] = sbad ? -1.0 : r;
    }
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
