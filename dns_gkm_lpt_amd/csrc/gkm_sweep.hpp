// gkm_sweep.hpp -- device code of the fused sweep (one workgroup's work) and of
// the CFL reductions, shared by the per-launch kernels (gkm_kernels.hip) and
// the persistent ensemble step (ens_persist.hip).
#pragma once
#include <hip/hip_runtime.h>

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

// LDS doubles of one tile variant: WENO states (10 per face and line) and
// face fluxes (5 per face slot and interior line)
template <int DIR, int V>
constexpr int sweep_lds_doubles()
{
    constexpr int CN = Tile<DIR, V>::CN, CA = Tile<DIR, V>::CA, CB = Tile<DIR, V>::CB;
    return 10 * CN * Lines<CA, CB>::L + 5 * (CN + 1) * CA * CB;
}
template <int DIR>
constexpr int sweep_lds_max()
{
    constexpr int a = sweep_lds_doubles<DIR, 0>(), b = sweep_lds_doubles<DIR, 1>(), c = sweep_lds_doubles<DIR, 2>();
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
}
constexpr int kSweepLdsMax = sweep_lds_max<0>() > sweep_lds_max<1>()
                                 ? (sweep_lds_max<0>() > sweep_lds_max<2>() ? sweep_lds_max<0>() : sweep_lds_max<2>())
                                 : (sweep_lds_max<1>() > sweep_lds_max<2>() ? sweep_lds_max<1>() : sweep_lds_max<2>());

// MIX = false: the FP64 path.  MIX = true: WENO in FP32 on FP64 stencil
// differences and the flux with FP32 slope terms (gks_flux_tf<double, float>).
// Two waves per SIMD are requested: the register budget is then 256 per lane.
// With one wave per SIMD the compiler (ROCm 7.2, gfx950) spilled into AGPRs and
// the x and y instances produced wrong flux differences on the GPU (the z
// instance and every instance built this way match the FP64 oracle to 1e-14).
//
// sweep_body is the work of one workgroup: block `bid` of the `nblk` blocks of
// one grid (member).  remap = true applies the XCD-aware order to bid.
// EXT = false: the body's own LDS arrays; EXT = true: the caller's buffer lds
// of at least sweep_lds_doubles<DIR, V>() doubles (a kernel running several
// bodies shares one buffer between them).
// GEN = true: any grid (tiles and segments may reach past nc); GEN = false:
// only for grids the tile and the segment divide (no line or face of the
// workgroup lies past the grid, so none is masked or wrapped into range: the
// general form's extra live values made the mixed x and y sweeps spill).
// CMP (the z sweep's update): -1 compensated when p.cmp_out is set, 0 plain,
// 1 compensated (the fixed forms leave the other update out of the kernel).
template <int DIR, bool MIX, int V, bool EXT = false, bool GEN = true, int CMP = -1>
__device__ __forceinline__ void sweep_body(const SweepArgs& p, int bid, int nblk, bool remap, double* lds = nullptr)
{
    constexpr int CN = Tile<DIR, V>::CN, CA = Tile<DIR, V>::CA, CB = Tile<DIR, V>::CB;
    constexpr int NT = CN * CA * CB;
    static_assert(NT == 256, "every tile variant runs 256 threads, one face each");
    using LN = Lines<CA, CB>;
    constexpr int L = LN::L;
    double* sW;  // WENO states: Wl (0..4), Wr (5..9)
    double* sF;  // face fluxes; slot 0 = face before the chunk
    if constexpr (EXT) {
        sW = lds;
    } else {
        __shared__ double own[sweep_lds_doubles<DIR, V>()];
        sW = own;
    }
    sF = sW + 10 * CN * L;

    const int nc = p.nc;
    const long plane = p.plane;
    const double* __restrict__ W = p.W;

    // workgroup -> (pencil, segment); XCD-aware: the hardware deals workgroups
    // round-robin over the 8 XCDs, so give each XCD a contiguous range of pencils.
    if (remap && (nblk & 7) == 0) bid = (bid & 7) * (nblk >> 3) + (bid >> 3);
    const int s = bid % p.nseg;
    const int pencil = bid / p.nseg;
    const int npb = GEN ? (nc + CB - 1) / CB : nc / CB;  // pencils across b; with GEN the last may reach past nc
    // z sweep on an exact grid: pencils in 8 x 8 patches, so that the ~64
    // workgroups an XCD runs at once cover a square of lines whose tangential
    // edge lines are mostly interior lines of the same patch, read once into
    // that XCD's L2 (the row order re-read the x edge lines of every strip;
    // 512^3: z sweep 19.55 -> 19.22 ms mixed, 25.44 -> 24.79 ms FP64).  The
    // x and y sweeps keep the row order, which measured faster for them.
    int pa, pb;
    if (DIR == 2 && !GEN && (npb & 7) == 0 && ((nc / CA) & 7) == 0) {
        const int t = pencil >> 6, w = pencil & 63, tpb = npb >> 3;
        pa = (t / tpb) * 8 + (w >> 3);
        pb = (t % tpb) * 8 + (w & 7);
    } else {
        pa = pencil / npb;
        pb = pencil % npb;
    }
    const int a0 = pa * CA;
    const int b0 = pb * CB;

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
    // a grid that is not a multiple of the tile: lines past nc (their cells
    // wrap) and faces past the segment's end are computed, not written
    const bool live = !GEN || (a0 + ta < nc && b0 + tb < nc);

    const double dt = *p.dt;
    const double dx = p.dx;
    const int nchunk = p.seg / CN;
    const int seg0 = s * p.seg;
    const int seg_end = seg0 + p.seg < nc ? seg0 + p.seg : nc;

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

        // ---- WENO states of every (face, line) of the chunk into LDS.  In the x
        // and y sweeps with an even face count a thread takes two neighbouring
        // faces of a line: their stencils share five of seven cells, loaded and
        // turned into temperatures once; each face's reconstruction is the same
        // arithmetic as a face alone (512^3, 4 x 4 x 16 and 2 x 8 x 16 tiles:
        // x 15.66 -> 14.45 ms, y 17.17 -> 16.42 ms mixed; 22.6 -> 21.2 and
        // 24.3 -> 23.2 ms FP64).  The z sweep keeps one face per thread: with
        // pairs, neighbouring lanes read cells two apart along z and it slowed
        // (19.25 -> 20.16 ms mixed).
        if constexpr (DIR != 2 && CN % 2 == 0) {
            constexpr int CH = CN / 2;
            const double gR = (p.g.gam - 1.0) / p.g.R;
            for (int i = tid; i < CH * L; i += NT) {
                const int il = i % L, itn = 2 * (i / L);
                int ia, ib;
                LN::coords(il, ia, ib);
                const int ca = wrapi(a0 + ia, nc), cb = wrapi(b0 + ib, nc);
                const int fn = n0 + itn;
                double S[7][5], Tc[7];
#pragma unroll
                for (int t = 0; t < 7; t++) {
                    const long idx = cell_index<DIR>(wrapi(fn - 2 + t, nc), ca, cb, nc);
#pragma unroll
                    for (int q = 0; q < 5; q++) S[t][q] = W[q * plane + idx];
                    Tc[t] = temp_gr<double>(S[t], gR);
                }
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (!MIX && h == 1) {
                        // FP64: the second face starts from fresh copies of the
                        // shared cells and temperatures, so the compiler cannot
                        // merge its arithmetic with the first face's (merged, the
                        // two faces' second differences and candidates were shared
                        // and the fused multiply-adds fell differently: a face's
                        // rounding depended on its place in the pair, and a
                        // periodic shift of the state no longer commuted with the
                        // step bit for bit).  The mixed reconstruction works on
                        // differences from each face's own base cell, so the two
                        // faces share only the loaded cells and temperatures
                        // (whose copies made the mixed sweeps spill).
#pragma unroll
                        for (int t = 1; t < 7; t++) {
#pragma unroll
                            for (int q = 0; q < 5; q++) asm volatile("" : "+v"(S[t][q])::"memory");
                            asm volatile("" : "+v"(Tc[t])::"memory");
                        }
                    }
                    double Wl[5], Wr[5];
                    if (MIX)
                        weno5_mixed_t(S + h, Tc + h, p.g, Wl, Wr);
                    else
                        weno5_t<double>(S + h, Tc + h, p.g, Wl, Wr);
#pragma unroll
                    for (int q = 0; q < 5; q++) {
                        sW[sw_at<DIR, CN, L>(q, itn + h, il)] = Wl[q];
                        sW[sw_at<DIR, CN, L>(5 + q, itn + h, il)] = Wr[q];
                    }
                }
            }
        } else
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
            const int ca = GEN ? wrapi(a0 + ta, nc) : a0 + ta, cb = GEN ? wrapi(b0 + tb, nc) : b0 + tb;
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
        if (c >= 0 && live && (!GEN || n0 + tn < seg_end)) {
            const long idx = cell_index<DIR>(n0 + tn, a0 + ta, b0 + tb, nc);
#pragma unroll
            for (int q = 0; q < 5; q++) {
                const double dF = sF[sf_at<DIR, CN, CA * CB>(q, tn + 1, lc)] - sF[sf_at<DIR, CN, CA * CB>(q, tn, lc)];
                const long o = q * plane + idx;
                // the x sweep only ever starts the sum; the y sweep adds to it
                // unless asked for its flux differences alone (mode 0), the z
                // sweep updates unless asked for its flux differences alone
                if (DIR == 0 || p.mode == 0) p.dW[o] = dF;
                else if (DIR == 1) p.dW[o] += dF;
                else if (CMP == 1 || (CMP < 0 && p.cmp_out)) {
                    // W^{n+1} + e = W^n + (increment + carried error), exactly (TwoSum);
                    // z sweep: n = z, a = x, b = y
                    const long co = q * carry_plane(nc) + carry_index(b0 + tb, a0 + ta, n0 + tn, nc);
                    const double w = W[o];
                    const double y = (double)p.cmp_in[co] - p.inv_dx * (p.dW[o] + dF);
                    const double t = w + y;
                    const double yv = t - w;
                    const double e = (w - (t - yv)) + (y - yv);
                    p.Wout[o] = t;
                    p.cmp_out[co] = (float)e;
                } else {
                    p.Wout[o] = W[o] - p.inv_dx * (p.dW[o] + dF);
                }
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

// ---------------------------------------------------------------------------
// CFL reductions: dt = cfl dx / max(|U| + sqrt(gam p / rho)) over all cells
// (timestep, update.cpp:1031-1054).  A non-physical cell (rho <= 0, p <= 0 or
// not finite) records -1 instead of its block maximum.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// maximum of |U| + c over cells c0, c0 + stride, ... of one state; thread 0
// writes it (or -1) to *out
__device__ __forceinline__ void cfl_cells(const double* __restrict__ W, long plane, double gam, long c0, long stride,
                                          double* out)
{
    double m = 0.0;
    bool bad = false;
    for (long c = c0 + threadIdx.x; c < plane; c += stride) {
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
        const double r = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
        *out = sbad ? -1.0 : r;
    }
    __syncthreads();  // sm, sbad are reused by the block's next call
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
    __syncthreads();  // sm, sb are reused by the block's next call
}

}  // namespace gkm
