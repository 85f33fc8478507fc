// adj_kernels.hip -- tangent-linear and discrete adjoint of the fixed-dt FP64
// GKM-WENO step (no reference counterpart).
//
// The step being differentiated is the one the sweeps evaluate
// (update.cpp:214-516 restated): for every direction, WENO states of every
// face and line, the face flux of (WL, WR, Wl, Wr, six slopes) with the
// collision time of the left cell, and W' = W - (1/dx) sum_d (F_d(i+1/2) -
// F_d(i-1/2)).  Derivatives come from running the same templates on dual
// numbers (gkm_dual.hpp):
//   tangent-linear: one dual pass per face with the perturbation as tangent;
//   adjoint: per face, the flux's 50 partials contracted with the face's
//   adjoint flux (face_flux_vjp, adj_flux.hpp: one forward and one reverse pass
//   through the whole flux); the slopes are linear in the WENO states, so
//   their adjoint is an explicit gather; per face and line, WENO in reverse
//   mode (one pass per one-dimensional reconstruction, the energy slot by the
//   chain rule); then a gather per cell.
// One thread per face or cell (they serve the 64^3 adjoint configuration).
// The flux adjoint runs one workgroup per CU (launch bounds 256, 1): with 512
// registers per lane and its side blocks in a function of their own it spills
// least, and at 64^3 that measured fastest (0.60 s per 200-step gradient,
// against 0.78 s inlined and 0.89 s at two workgroups per CU, with the WENO
// adjoint at two workgroups per CU in all three).
//
// Face arrays hold one value per face and direction at the index of the face's
// left cell.  Direction d: sweep axis n, tangential axes a, b as in the sweep
// kernels (x: n = x, a = y, b = z; y: n = y, a = x, b = z; z: n = z, a = x, b = y).
#include "adj_kernels.hpp"

#include "adj_flux.hpp"
#include "adj_math.hpp"
#include "gkm_dual.hpp"
#include "gkm_flux.hpp"

namespace gkm {

namespace {

__device__ __forceinline__ int wrap(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

template <int DIR>
__device__ __forceinline__ long cell_at(int n, int a, int b, int nc)
{
    if (DIR == 0) return ((long)a * nc + n) * nc + b;
    if (DIR == 1) return ((long)n * nc + a) * nc + b;
    return ((long)b * nc + a) * nc + n;
}

// (n, a, b) of linear index c (inverse of cell_at)
template <int DIR>
__device__ __forceinline__ void coords_of(long c, int nc, int& n, int& a, int& b)
{
    const int z = (int)(c % nc), x = (int)((c / nc) % nc), y = (int)(c / ((long)nc * nc));
    if (DIR == 0) { n = x; a = y; b = z; }
    else if (DIR == 1) { n = y; a = x; b = z; }
    else { n = z; a = x; b = y; }
}

// rotated-frame slot -> lab slot (normal momentum in slot 1; update.cpp:361-398, 461-497)
template <int DIR>
__device__ __forceinline__ int phys(int q)
{
    constexpr int sw = DIR == 1 ? 2 : 3;
    return DIR == 0 ? q : (q == 1 ? sw : (q == sw ? 1 : q));
}

// the 50 flux inputs of face (n, a, b) in the rotated frame: values (and tangents
// when Wt / WSt are given)
template <int DIR, typename T, typename Get>
__device__ __forceinline__ void gather_inputs(int n, int a, int b, int nc, long plane, double dx, Get get, T x[50])
{
    const long f = cell_at<DIR>(n, a, b, nc);
    const long fr = cell_at<DIR>(wrap(n + 1, nc), a, b, nc);
    const long fam = cell_at<DIR>(n, wrap(a - 1, nc), b, nc), fap = cell_at<DIR>(n, wrap(a + 1, nc), b, nc);
    const long fbm = cell_at<DIR>(n, a, wrap(b - 1, nc), nc), fbp = cell_at<DIR>(n, a, wrap(b + 1, nc), nc);
    const double two_idx = 2.0 / dx, half_idx = 0.5 / dx;
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const int p = phys<DIR>(q);
        const T WL = get(0, p, f), WR = get(0, p, fr);   // cells
        const T Wl = get(1, p, f), Wr = get(1, 5 + p, f); // WENO states
        x[q] = WL;
        x[5 + q] = WR;
        x[10 + q] = Wl;
        x[15 + q] = Wr;
        x[20 + q] = (Wl - WL) * T(two_idx);
        x[25 + q] = (WR - Wr) * T(two_idx);
        x[30 + q] = (get(1, p, fap) - get(1, p, fam)) * T(half_idx);
        x[35 + q] = (get(1, 5 + p, fap) - get(1, 5 + p, fam)) * T(half_idx);
        x[40 + q] = (get(1, p, fbp) - get(1, p, fbm)) * T(half_idx);
        x[45 + q] = (get(1, 5 + p, fbp) - get(1, 5 + p, fbm)) * T(half_idx);
    }
}

// ---------------------------------------------------------------------------
// forward intermediates and tangent-linear
// ---------------------------------------------------------------------------
template <int DIR>
__global__ __launch_bounds__(256, 2) void weno_faces_kernel(const double* __restrict__ W, const double* __restrict__ Wt,
                                                         long plane, int nc, Gas<double> g, double* __restrict__ WS,
                                                         double* __restrict__ WSt)
{
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= plane) return;
    int n, a, b;
    coords_of<DIR>(f, nc, n, a, b);
    if (!Wt) {
        double S[6][5], Wl[5], Wr[5];
#pragma unroll
        for (int t = 0; t < 6; t++) {
            const long c = cell_at<DIR>(wrap(n - 2 + t, nc), a, b, nc);
#pragma unroll
            for (int q = 0; q < 5; q++) S[t][q] = W[q * plane + c];
        }
        weno5<double>(S, g, Wl, Wr);
#pragma unroll
        for (int q = 0; q < 5; q++) {
            WS[q * plane + f] = Wl[q];
            WS[(5 + q) * plane + f] = Wr[q];
        }
        return;
    }
    const Gas<Dual> gd = dual_gas(g);
    Dual S[6][5], Wl[5], Wr[5];
    for (int t = 0; t < 6; t++) {
        const long c = cell_at<DIR>(wrap(n - 2 + t, nc), a, b, nc);
        for (int q = 0; q < 5; q++) S[t][q] = Dual(W[q * plane + c], Wt[q * plane + c]);
    }
    weno5<Dual>(S, gd, Wl, Wr);
    for (int q = 0; q < 5; q++) {
        WS[q * plane + f] = Wl[q].v;
        WS[(5 + q) * plane + f] = Wr[q].v;
        WSt[q * plane + f] = Wl[q].d;
        WSt[(5 + q) * plane + f] = Wr[q].d;
    }
}

template <int DIR>
__global__ __launch_bounds__(256, 2) void flux_tl_kernel(const double* __restrict__ W, const double* __restrict__ Wt,
                                                      const double* __restrict__ WS, const double* __restrict__ WSt,
                                                      long plane, int nc, Gas<double> g, double dt, double dx,
                                                      double* __restrict__ Ft)
{
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= plane) return;
    int n, a, b;
    coords_of<DIR>(f, nc, n, a, b);
    auto get = [&](int arr, int v, long idx) {
        return arr == 0 ? Dual(W[v * plane + idx], Wt[v * plane + idx]) : Dual(WS[v * plane + idx], WSt[v * plane + idx]);
    };
    Dual x[50], F[5];
    gather_inputs<DIR, Dual>(n, a, b, nc, plane, dx, get, x);
    face_flux<Dual>(x, Dual(dt), Dual(dx), dual_gas(g), F);
    for (int q = 0; q < 5; q++) Ft[phys<DIR>(q) * plane + f] = F[q].d;
}

// Wout = W - (1/dx) sum_d (F_d(face of the cell) - F_d(face before it))
__global__ __launch_bounds__(256, 2) void tl_update_kernel(const double* __restrict__ Wt, const double* __restrict__ F0,
                                                        const double* __restrict__ F1, const double* __restrict__ F2,
                                                        long plane, int nc, double inv_dx, double* __restrict__ out)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= plane) return;
    const int z = (int)(c % nc), x = (int)((c / nc) % nc), y = (int)(c / ((long)nc * nc));
    const long cx = ((long)y * nc + wrap(x - 1, nc)) * nc + z;  // previous face along x
    const long cy = ((long)wrap(y - 1, nc) * nc + x) * nc + z;
    const long cz = ((long)y * nc + x) * nc + wrap(z - 1, nc);
    for (int q = 0; q < 5; q++) {
        const long o = q * plane;
        const double s = (F0[o + c] - F0[o + cx]) + (F1[o + c] - F1[o + cy]) + (F2[o + c] - F2[o + cz]);
        out[o + c] = Wt[o + c] - inv_dx * s;
    }
}

// ---------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------
// Xbar[k][f] = sum_q Fbar_q dF_q / dx_k for the 50 flux inputs of face f, with
// Fbar = -(1/dx) (lam(cell n) - lam(cell n+1)); stored by input group and lab slot.
template <int DIR>
__global__ __launch_bounds__(256, 1) void flux_adj_kernel(const double* __restrict__ W, const double* __restrict__ WS,
                                                       const double* __restrict__ lam, long plane, int nc,
                                                       Gas<double> g, double dt, double dx,
                                                       double* __restrict__ Xbar)
{
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= plane) return;
    int n, a, b;
    coords_of<DIR>(f, nc, n, a, b);
    const long fr = cell_at<DIR>(wrap(n + 1, nc), a, b, nc);
    const double inv_dx = 1.0 / dx;
    double Fbar[5];
    for (int q = 0; q < 5; q++) {
        const int p = phys<DIR>(q);
        Fbar[q] = -inv_dx * (lam[p * plane + f] - lam[p * plane + fr]);
    }
    auto get = [&](int arr, int v, long idx) { return arr == 0 ? W[v * plane + idx] : WS[v * plane + idx]; };
    double x[50];
    gather_inputs<DIR, double>(n, a, b, nc, plane, dx, get, x);
    double xbar[50];
    face_flux_vjp(x, dt, dx, g, Fbar, xbar);
    for (int k = 0; k < 50; k++) Xbar[(long)(5 * (k / 5) + phys<DIR>(k % 5)) * plane + f] = xbar[k];
}

// adjoints of the WENO states (WSbar, 10 per face) and of the face's two cells
// (CL for cell n, CR for cell n+1), from Xbar through the slope definitions
// D0 = 2 (Wl - WL)/dx, D1 = 2 (WR - Wr)/dx, D2/D3 = (W(a+1) - W(a-1))/(2 dx) of
// Wl/Wr, D4/D5 the same along b.
template <int DIR>
__global__ __launch_bounds__(256, 2) void adj_assemble_kernel(const double* __restrict__ Xbar, long plane, int nc,
                                                           double dx, double* __restrict__ WSbar,
                                                           double* __restrict__ CL, double* __restrict__ CR)
{
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= plane) return;
    int n, a, b;
    coords_of<DIR>(f, nc, n, a, b);
    const long fam = cell_at<DIR>(n, wrap(a - 1, nc), b, nc), fap = cell_at<DIR>(n, wrap(a + 1, nc), b, nc);
    const long fbm = cell_at<DIR>(n, a, wrap(b - 1, nc), nc), fbp = cell_at<DIR>(n, a, wrap(b + 1, nc), nc);
    const double two_idx = 2.0 / dx, half_idx = 0.5 / dx;
    auto X = [&](int grp, int p, long idx) { return Xbar[(long)(5 * grp + p) * plane + idx]; };
    for (int p = 0; p < 5; p++) {
        WSbar[p * plane + f] = X(2, p, f) + two_idx * X(4, p, f) + half_idx * (X(6, p, fam) - X(6, p, fap)) +
                               half_idx * (X(8, p, fbm) - X(8, p, fbp));
        WSbar[(5 + p) * plane + f] = X(3, p, f) - two_idx * X(5, p, f) + half_idx * (X(7, p, fam) - X(7, p, fap)) +
                                     half_idx * (X(9, p, fbm) - X(9, p, fbp));
        CL[p * plane + f] = X(0, p, f) - two_idx * X(4, p, f);
        CR[p * plane + f] = X(1, p, f) + two_idx * X(5, p, f);
    }
}

// Sbar[5 t + q][f] = sum_o WSbar_o dWS_o / dS_{t,q}.  WENO-SYMOO reconstructs
// rho, rho u, rho v, rho w one component at a time; only the energy slot
// couples them, E = W0 R T igm + |W1..3|^2 / (2 W0) with T the reconstruction of
// the cells' temperatures.  The adjoints of the energy slots are first moved
// onto the other four slots and the reconstructed temperatures (closed-form
// partial derivatives), then each one-dimensional reconstruction is
// differentiated in reverse mode (weno_side_grad: all six cells in one pass),
// and the temperature gradient reaches the cells through T(W).  One workgroup
// per CU: with 512 registers per lane it spills 60 B per lane instead of
// 1060 B (0.53 s per 64^3 200-step gradient against 0.60 s).
template <int DIR>
__global__ __launch_bounds__(256, 1) void weno_adj_kernel(const double* __restrict__ W, const double* __restrict__ WSbar,
                                                       long plane, int nc, Gas<double> g, double* __restrict__ Sbar)
{
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= plane) return;
    int n, a, b;
    coords_of<DIR>(f, nc, n, a, b);
    double S[6][5], Tc[6];
#pragma unroll
    for (int t = 0; t < 6; t++) {
        const long c = cell_at<DIR>(wrap(n - 2 + t, nc), a, b, nc);
#pragma unroll
        for (int q = 0; q < 5; q++) S[t][q] = W[q * plane + c];
        Tc[t] = temp_of(S[t], g.gam, g.R);
    }
    const double eps = g.eps, igm = 1.0 / (g.gam - 1.0), gR = (g.gam - 1.0) / g.R;
    double gS[6][5], gT[6];
#pragma unroll
    for (int t = 0; t < 6; t++) {
        gT[t] = 0.0;
#pragma unroll
        for (int q = 0; q < 5; q++) gS[t][q] = 0.0;
    }
    // side 0: the left state Wl (stencil i-2..i+3); side 1: the right state Wr
    // (the same formula on the mirrored stencil)
#pragma unroll
    for (int side = 0; side < 2; side++) {
        auto cell = [&](int k) { return side == 0 ? k : 5 - k; };
        double w[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
            w[q] = weno_side(S[cell(0)][q], S[cell(1)][q], S[cell(2)][q], S[cell(3)][q], S[cell(4)][q], S[cell(5)][q],
                             eps);
        const double T = weno_side(Tc[cell(0)], Tc[cell(1)], Tc[cell(2)], Tc[cell(3)], Tc[cell(4)], Tc[cell(5)], eps);
        const double eb = WSbar[(5 * side + 4) * plane + f];
        const double m2 = w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double wbar = WSbar[(5 * side + q) * plane + f] +
                                eb * (q == 0 ? g.R * T * igm - 0.5 * m2 / (w[0] * w[0]) : w[q] / w[0]);
            double gq[6] = {0, 0, 0, 0, 0, 0};
            weno_side_grad(S[cell(0)][q], S[cell(1)][q], S[cell(2)][q], S[cell(3)][q], S[cell(4)][q], S[cell(5)][q],
                           eps, wbar, gq);
#pragma unroll
            for (int k = 0; k < 6; k++) gS[cell(k)][q] += gq[k];
        }
        double gt[6] = {0, 0, 0, 0, 0, 0};
        weno_side_grad(Tc[cell(0)], Tc[cell(1)], Tc[cell(2)], Tc[cell(3)], Tc[cell(4)], Tc[cell(5)], eps,
                       eb * (w[0] * g.R * igm), gt);
#pragma unroll
        for (int k = 0; k < 6; k++) gT[cell(k)] += gt[k];
    }
    // T(W) = (gam-1)/R (W4/W0 - |W1..3|^2/(2 W0^2)) per cell
#pragma unroll
    for (int t = 0; t < 6; t++) {
        const double ir = 1.0 / S[t][0], k2 = S[t][1] * S[t][1] + S[t][2] * S[t][2] + S[t][3] * S[t][3];
        const double dT = gT[t] * gR;
        gS[t][0] += dT * (-S[t][4] * ir * ir + k2 * ir * ir * ir);
#pragma unroll
        for (int q = 1; q < 4; q++) gS[t][q] += dT * (-S[t][q] * ir * ir);
        gS[t][4] += dT * ir;
#pragma unroll
        for (int q = 0; q < 5; q++) Sbar[(long)(5 * t + q) * plane + f] = gS[t][q];
    }
}

// lam_acc(cell) += CL(face of the cell) + CR(face before it) + sum_t Sbar_t(face cell + 2 - t)
template <int DIR>
__global__ __launch_bounds__(256, 2) void adj_gather_kernel(const double* __restrict__ CL, const double* __restrict__ CR,
                                                         const double* __restrict__ Sbar, long plane, int nc,
                                                         double* __restrict__ lam_acc)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= plane) return;
    int n, a, b;
    coords_of<DIR>(c, nc, n, a, b);
    const long fprev = cell_at<DIR>(wrap(n - 1, nc), a, b, nc);
    long fs[6];
    for (int t = 0; t < 6; t++) fs[t] = cell_at<DIR>(wrap(n + 2 - t, nc), a, b, nc);
    for (int q = 0; q < 5; q++) {
        double s = CL[q * plane + c] + CR[q * plane + fprev];
        for (int t = 0; t < 6; t++) s += Sbar[(long)(5 * t + q) * plane + fs[t]];
        lam_acc[q * plane + c] += s;
    }
}

unsigned blocks_of(long n) { return (unsigned)((n + 255) / 256); }

template <int DIR>
hipError_t tl_dir(const double* W, const double* Wt, long plane, int nc, const Gas<double>& g, double dt, double dx,
                  const AdjScratch& s, double* Ft, hipStream_t st)
{
    hipLaunchKernelGGL(weno_faces_kernel<DIR>, dim3(blocks_of(plane)), dim3(256), 0, st, W, Wt, plane, nc, g, s.WS,
                       s.WSt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(flux_tl_kernel<DIR>, dim3(blocks_of(plane)), dim3(256), 0, st, W, Wt, s.WS, s.WSt, plane, nc, g,
                       dt, dx, Ft);
    return hipGetLastError();
}

template <int DIR>
hipError_t adj_dir(const double* W, const double* lam, long plane, int nc, const Gas<double>& g, double dt, double dx,
                   const AdjScratch& s, double* lam_acc, hipStream_t st)
{
    const unsigned nb = blocks_of(plane);
    hipLaunchKernelGGL(weno_faces_kernel<DIR>, dim3(nb), dim3(256), 0, st, W, nullptr, plane, nc, g, s.WS, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(flux_adj_kernel<DIR>, dim3(nb), dim3(256), 0, st, W, s.WS, lam, plane, nc, g, dt, dx, s.Xbar);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(adj_assemble_kernel<DIR>, dim3(nb), dim3(256), 0, st, s.Xbar, plane, nc, dx, s.WSbar, s.CL, s.CR);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(weno_adj_kernel<DIR>, dim3(nb), dim3(256), 0, st, W, s.WSbar, plane, nc, g, s.Sbar);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(adj_gather_kernel<DIR>, dim3(nb), dim3(256), 0, st, s.CL, s.CR, s.Sbar, plane, nc, lam_acc);
    return hipGetLastError();
}

}  // namespace

size_t adj_scratch_doubles(long plane) { return (size_t)plane * (10 + 10 + 3 * 5 + 50 + 10 + 5 + 5 + 30); }

void adj_scratch_carve(double* base, long plane, AdjScratch& s)
{
    double* p = base;
    auto take = [&](int k) {
        double* r = p;
        p += (size_t)k * plane;
        return r;
    };
    s.WS = take(10);
    s.WSt = take(10);
    for (int d = 0; d < 3; d++) s.Ft[d] = take(5);
    s.Xbar = take(50);
    s.WSbar = take(10);
    s.CL = take(5);
    s.CR = take(5);
    s.Sbar = take(30);
}

hipError_t launch_tl_step(const double* W, const double* Wt, long plane, int nc, const Gas<double>& g, double dt,
                          double dx, const AdjScratch& s, double* Wt_out, hipStream_t st)
{
    if (Wt == Wt_out) return hipErrorInvalidValue;
    hipError_t e = tl_dir<0>(W, Wt, plane, nc, g, dt, dx, s, s.Ft[0], st);
    if (e == hipSuccess) e = tl_dir<1>(W, Wt, plane, nc, g, dt, dx, s, s.Ft[1], st);
    if (e == hipSuccess) e = tl_dir<2>(W, Wt, plane, nc, g, dt, dx, s, s.Ft[2], st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(tl_update_kernel, dim3(blocks_of(plane)), dim3(256), 0, st, Wt, s.Ft[0], s.Ft[1], s.Ft[2], plane,
                       nc, 1.0 / dx, Wt_out);
    return hipGetLastError();
}

hipError_t launch_adj_step(const double* W, const double* lam, long plane, int nc, const Gas<double>& g, double dt,
                           double dx, const AdjScratch& s, double* lam_out, hipStream_t st)
{
    if (lam == lam_out) return hipErrorInvalidValue;
    hipError_t e = hipMemcpyAsync(lam_out, lam, sizeof(double) * 5 * plane, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = adj_dir<0>(W, lam, plane, nc, g, dt, dx, s, lam_out, st);
    if (e == hipSuccess) e = adj_dir<1>(W, lam, plane, nc, g, dt, dx, s, lam_out, st);
    if (e == hipSuccess) e = adj_dir<2>(W, lam, plane, nc, g, dt, dx, s, lam_out, st);
    return e;
}

}  // namespace gkm
