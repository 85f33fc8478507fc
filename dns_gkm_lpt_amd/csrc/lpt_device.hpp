// lpt_device.hpp -- device code of the spline prefilter and of the tracer
// sample and Heun steps, shared by the per-launch kernels (lpt_kernels.hip)
// and the persistent ensemble step (ens_persist.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "gkm_math.hpp"
#include "lpt_kernels.hpp"

namespace lpt {

__device__ __forceinline__ int wrapi(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

// Natural-spline line solve in place.  The line holds f_0..f_{M-1} at
// positions 0..M-1 (stride st) and receives c_0..c_{M+1}.  With n = M - 2
// unknowns (row r reads f_{r+1} and writes position r + 2) the Thomas sweeps
// are
//   forward   dp_r = (6 f_{r+1} - [r = 0] c_1 - [r = n-1] c_M - dp_{r-1}) piv_r
//   backward  x_{n-1} = dp_{n-1},  x_r = dp_r - piv_r x_{r+1},
// then c_1 = f_0, c_M = f_{M-1}, c_0 = 2 c_1 - x_0, c_{M+1} = 2 c_M - x_{n-1}
// (einspline natural boundary; piv_r = 1/(4 - piv_{r-1}), piv_0 = 1/4, formed on
// the host).
//
// The rows go in blocks of B, three kept in registers, so every position is
// read once and written once (the two-sweep form writes dp and reads it back).
// Block j's backward sweep enters with x_{r0+B}, obtained by running the
// backward recurrence over blocks j+2 and j+1 from x = dp at the last row of
// block j+2: exact when that row is n - 1, otherwise off by piv x_{r+1}, an
// error each later step multiplies by piv_r <= 2 - sqrt(3) = 0.268, so by
// 0.268^32 = 5e-19 of the line's values after the 2B = 32 rows of the two
// blocks (below rounding: results equal the sequential sweeps' to a few ulp).
// The forward sweep is the sequential one.
template <int B, bool FULL>
__host__ __device__ __forceinline__ void fwd_block(double (&D)[B], int r0, int n, const double* line, long st,
                                          const double* __restrict__ piv, double c1, double cM, double& dp)
{
#pragma unroll
    for (int k = 0; k < B; k++) D[k] = (FULL || r0 + k < n) ? line[(long)(r0 + k + 1) * st] : 0.0;
#pragma unroll
    for (int k = 0; k < B; k++) {
        const int r = r0 + k;
        if (FULL || r < n) {
            double rhs = 6.0 * D[k];
            if (r == 0) rhs -= c1;
            if (r == n - 1) rhs -= cM;
            dp = (rhs - dp) * piv[r];
            D[k] = dp;
        }
    }
}

// backward sweep over the rows of D (rows r0.. below n) entering with
// x = x_{r0+B} (0 for the line's last block); WRITE stores x_r at position r + 2
template <int B, bool FULL, bool WRITE>
__host__ __device__ __forceinline__ double bwd_block(const double (&D)[B], int r0, int n, double x, double* line, long st,
                                            const double* __restrict__ piv)
{
#pragma unroll
    for (int k = B - 1; k >= 0; k--) {
        const int r = r0 + k;
        if (FULL || r < n) {
            x = D[k] - piv[r] * x;
            if (WRITE) line[(long)(r + 2) * st] = x;
        }
    }
    return x;
}

// One step of the ring: D holds rows r0.., E rows r0 + B.. (when r0 + B < n);
// F receives rows r0 + 2B.. (read forward here), then F and E are swept
// backward as the warm-up of D's backward sweep, which writes D's rows.  Returns
// true when D was the line's last block.
template <int B>
__host__ __device__ __forceinline__ bool ring_step(double (&D)[B], double (&E)[B], double (&F)[B], int r0, int n,
                                                   double* line, long st, const double* __restrict__ piv, double c1,
                                                   double cM, double& dp, double& x0)
{
    double x;
    if (r0 + B >= n) {  // D holds the end of the line
        x = bwd_block<B, false, true>(D, r0, n, 0.0, line, st, piv);
        if (r0 == 0) x0 = x;
        return true;
    }
    if (r0 + 2 * B >= n) {  // E holds the end of the line: its backward sweep starts exactly at n - 1
        x = bwd_block<B, false, false>(E, r0 + B, n, 0.0, line, st, piv);
    } else {
        if (r0 + 3 * B <= n) {
            fwd_block<B, true>(F, r0 + 2 * B, n, line, st, piv, c1, cM, dp);
            x = bwd_block<B, true, false>(F, r0 + 2 * B, n, 0.0, line, st, piv);
        } else {
            fwd_block<B, false>(F, r0 + 2 * B, n, line, st, piv, c1, cM, dp);
            x = bwd_block<B, false, false>(F, r0 + 2 * B, n, 0.0, line, st, piv);
        }
        x = bwd_block<B, true, false>(E, r0 + B, n, x, line, st, piv);
    }
    x = bwd_block<B, true, true>(D, r0, n, x, line, st, piv);
    if (r0 == 0) x0 = x;
    return false;
}

template <int B>
__host__ __device__ __forceinline__ void line_solve_blocked(int M, double* line, long st, const double* __restrict__ piv)
{
    const int n = M - 2;
    const double c1 = line[0], cM = line[(long)(M - 1) * st];
    double P[B], Q[B], R[B];
    double dp = 0.0, x0 = 0.0;
    if (B <= n) fwd_block<B, true>(P, 0, n, line, st, piv, c1, cM, dp);
    else fwd_block<B, false>(P, 0, n, line, st, piv, c1, cM, dp);
    if (B < n) {
        if (2 * B <= n) fwd_block<B, true>(Q, B, n, line, st, piv, c1, cM, dp);
        else fwd_block<B, false>(Q, B, n, line, st, piv, c1, cM, dp);
    }
    for (int r0 = 0;; r0 += 3 * B) {
        if (ring_step<B>(P, Q, R, r0, n, line, st, piv, c1, cM, dp, x0)) break;
        if (ring_step<B>(Q, R, P, r0 + B, n, line, st, piv, c1, cM, dp, x0)) break;
        if (ring_step<B>(R, P, Q, r0 + 2 * B, n, line, st, piv, c1, cM, dp, x0)) break;
    }
    line[st] = c1;
    line[(long)M * st] = cM;
    line[0] = 2.0 * c1 - x0;                       // x0 = c_2
    line[(long)(M + 1) * st] = 2.0 * cM - dp;      // dp = dp_{n-1} = x_{n-1} = c_{M-1}
}

// rows per register block of line_solve_blocked (three blocks in registers:
// the warm-up spans two)
constexpr int kLineBlock = 16;

// Pass 1: z lines straight from the state (periodic wrap).  Source field q as
// generateSplines forms it (update.cpp:1202-1216): 0 = rho, 1..3 = velocity
// W_q / rho, 4 = pressure (rho (gam-1)) (E/rho - |U|^2/2) with U = W/rho.
// Field q of the state at padded position (a, b, k) of the knot grid; W_q / rho
// as W_q times the reciprocal of rho (gkm::g_rcp, within 1 ulp of the quotient).
__device__ __forceinline__ double field_at(const double* __restrict__ W, long plane, int q, double gam, int nc, int a,
                                          int b, int k)
{
    const long c = ((long)wrapi(a - 3, nc) * nc + wrapi(b - 3, nc)) * nc + wrapi(k - 3, nc);
    const double den = W[c];
    if (q == 0) return den;
    const double ir = gkm::g_rcp(den);
    if (q <= 3) return W[q * plane + c] * ir;
    const double u = W[plane + c] * ir, v = W[2 * plane + c] * ir, w = W[3 * plane + c] * ir;
    const double Vsqr = u * u + v * v + w * w;
    return (den * (gam - 1)) * (W[4 * plane + c] * ir - 0.5 * Vsqr);
}

// ---------------------------------------------------------------------------
// z pass: one wave per z line.  Lane l of chunk j holds knot position
// 64 j + l, so loads and stores are contiguous along z.  Both sweeps of
// the natural-spline solve (line_solve_blocked) are first-order linear recurrences,
//   forward   dp_r = -piv_r dp_{r-1} + piv_r rhs_r      (dp_{-1} = 0)
//   backward  x_r  = -piv_r x_{r+1}  + dp_r             (x_{n-1} = dp_{n-1}),
// so each 64-position chunk is an inclusive wave scan of the affine maps
// x -> A x + B (DPP row shifts and row broadcasts, no LDS) applied to the
// carry from the previous chunk.  The backward sweep scans lane-reversed
// chunks.  Results equal the sequential sweeps' to rounding (the scan associates the
// same products differently: measured <= 4 ulp of the line's largest value).
// ---------------------------------------------------------------------------
template <int CTRL, int ROW>
__device__ __forceinline__ double dpp64(double old, double v)
{
    const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, ROW, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, ROW, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// compose with the map of the lane CTRL selects (identity where there is none)
template <int CTRL, int ROW>
__device__ __forceinline__ void scan_step(double& A, double& B)
{
    const double Ap = dpp64<CTRL, ROW>(1.0, A), Bp = dpp64<CTRL, ROW>(0.0, B);
    B = A * Bp + B;
    A = A * Ap;
}

// inclusive scan over the 64 lanes (lane 0's map applied first); all lanes active
__device__ __forceinline__ void wave_scan(double& A, double& B)
{
    scan_step<0x111, 0xf>(A, B);  // row_shr:1
    scan_step<0x112, 0xf>(A, B);  // row_shr:2
    scan_step<0x114, 0xf>(A, B);  // row_shr:4
    scan_step<0x118, 0xf>(A, B);  // row_shr:8
    scan_step<0x142, 0xa>(A, B);  // row_bcast:15 into rows 1 and 3
    scan_step<0x143, 0xc>(A, B);  // row_bcast:31 into rows 2 and 3
}

// largest knot count per line a wave holds (17 chunks of 64: nc <= 1082)
constexpr int kMaxZChunks = 17;

// Chunks away from both ends of a line see one pivot: piv_r = 1/(4 - piv_{r-1})
// reaches its FP64 fixed point pc (2 - sqrt(3) rounded) at r = 13 and stays
// there, so the maps of such a chunk are x -> a x + B_l with one a = -pc.  The
// scan then carries B alone, B_l += a^k B_{l-k}, with a^k a constant per step
// (a^1, a^2, a^4, a^8 for the row shifts; a^(l mod 16 + 1) and a^(l - 31) per
// lane for the row broadcasts), and the carry enters with a^(l + 1).
struct ScanW {
    double a1, a2, a4, a8, w15, w31, wc;
};

__device__ __forceinline__ double pow_small(double a, int e)  // a^e, 0 <= e <= 64, by squaring
{
    double r = 1.0;
    for (; e; e >>= 1, a *= a)
        if (e & 1) r *= a;
    return r;
}

__device__ __forceinline__ ScanW scan_weights(double a, int lane)
{
    ScanW w;
    w.a1 = a;
    w.a2 = a * a;
    w.a4 = w.a2 * w.a2;
    w.a8 = w.a4 * w.a4;
    const int row = lane >> 4;
    w.w15 = (row & 1) ? pow_small(a, (lane & 15) + 1) : 0.0;
    w.w31 = row >= 2 ? pow_small(a, lane - 31) : 0.0;
    w.wc = pow_small(a, lane + 1);
    return w;
}

// inclusive scan over the 64 lanes of maps x -> a x + B with a common a (see
// above): returns sum_{m <= l} a^(l-m) B_m at lane l; all lanes active
__device__ __forceinline__ double wave_scan_const(double B, const ScanW& w)
{
    B += w.a1 * dpp64<0x111, 0xf>(0.0, B);   // row_shr:1
    B += w.a2 * dpp64<0x112, 0xf>(0.0, B);   // row_shr:2
    B += w.a4 * dpp64<0x114, 0xf>(0.0, B);   // row_shr:4
    B += w.a8 * dpp64<0x118, 0xf>(0.0, B);   // row_shr:8
    B += w.w15 * dpp64<0x142, 0xa>(0.0, B);  // row_bcast:15 into rows 1 and 3
    B += w.w31 * dpp64<0x143, 0xc>(0.0, B);  // row_bcast:31 into rows 2 and 3
    return B;
}

__device__ __forceinline__ double lane_of(double v, int l)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, l), hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// natural-spline solve of NF z lines of one knot grid: f[q] holds the M knot
// values of line q (zero beyond M), out[q] receives its M + 2 coefficients.
// The NF lines go through each step together: every line's operations are
// those of a solve on its own, and the lines' independent scan chains give
// the VALU instructions to issue while one chain waits.  Called by every lane
// of a full wave.
template <int K, int NF>
__device__ __forceinline__ void zline_solve_n(double (&f)[NF][K], int M, const double* __restrict__ piv,
                                              double* const (&out)[NF])
{
    const int lane = threadIdx.x & 63, n = M - 2;
    // chunk j (rows 64 j - 1 .. 64 j + 62) is "inner" when j >= 1 (rows >= 63,
    // past the pivots' convergence) and its last row is below n - 1 (no end
    // condition, no row past the line); lines of at most two chunks have none,
    // and lines of 17 chunks (over 576 knots) keep the general scan, with which
    // their kernel needs no scratch
    constexpr bool kInner = K >= 3 && K < kMaxZChunks;
    const double pc = kInner ? piv[n - 1 < 63 ? n - 1 : 63] : 0.0;
    const ScanW sw = scan_weights(-pc, lane);
    double c1[NF], cM[NF], carry[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) {
        c1[q] = lane_of(f[q][0], 0);
        cM[q] = 0.0;
        carry[q] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < K; j++)
        if (j == (M - 1) / 64)
#pragma unroll
            for (int q = 0; q < NF; q++) cM[q] = lane_of(f[q][j], (M - 1) % 64);
#pragma unroll
    for (int j = 0; j < K; j++) {  // forward: r = p - 1 for position p = 64 j + lane
        if (kInner && j >= 1 && 64 * j + 62 <= n - 2) {
#pragma unroll
            for (int q = 0; q < NF; q++) f[q][j] = sw.wc * carry[q] + wave_scan_const(pc * (6.0 * f[q][j]), sw);
#pragma unroll
            for (int q = 0; q < NF; q++) carry[q] = lane_of(f[q][j], 63);
            continue;
        }
        const int r = 64 * j + lane - 1;
        double A[NF], B[NF];
#pragma unroll
        for (int q = 0; q < NF; q++) {
            A[q] = 1.0;
            B[q] = 0.0;
        }
        if (r >= 0 && r < n) {
            const double pv = piv[r];
#pragma unroll
            for (int q = 0; q < NF; q++) {
                double rhs = 6.0 * f[q][j];
                if (r == 0) rhs -= c1[q];
                if (r == n - 1) rhs -= cM[q];
                A[q] = -pv;
                B[q] = pv * rhs;
            }
        }
#pragma unroll
        for (int q = 0; q < NF; q++) wave_scan(A[q], B[q]);
#pragma unroll
        for (int q = 0; q < NF; q++) {
            f[q][j] = A[q] * carry[q] + B[q];
            carry[q] = lane_of(f[q][j], 63);
        }
    }
    double xl[NF], x0[NF];  // x_{n-1} = dp_{n-1}; x_0
#pragma unroll
    for (int q = 0; q < NF; q++) {
        xl[q] = 0.0;
        x0[q] = 0.0;
        carry[q] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < K; j++)
        if (j == n / 64)
#pragma unroll
            for (int q = 0; q < NF; q++) xl[q] = lane_of(f[q][j], n % 64);
#pragma unroll
    for (int j = K - 1; j >= 0; j--) {  // backward on reversed chunks: lane l holds r = 64 j + 62 - l
        double d[NF];
#pragma unroll
        for (int q = 0; q < NF; q++) d[q] = __shfl(f[q][j], 63 - lane);
        const int r = 64 * j + 62 - lane;
        if (kInner && j >= 1 && 64 * j + 62 <= n - 2) {
#pragma unroll
            for (int q = 0; q < NF; q++) {
                const double x = sw.wc * carry[q] + wave_scan_const(d[q], sw);
                carry[q] = lane_of(x, 63);
                out[q][r + 2] = x;
            }
            continue;
        }
        double A[NF], B[NF];
#pragma unroll
        for (int q = 0; q < NF; q++) {
            A[q] = 1.0;
            B[q] = 0.0;
        }
        if (r >= 0 && r < n) {
            const double a = r == n - 1 ? 0.0 : -piv[r];
#pragma unroll
            for (int q = 0; q < NF; q++) {
                A[q] = a;
                B[q] = d[q];
            }
        }
#pragma unroll
        for (int q = 0; q < NF; q++) wave_scan(A[q], B[q]);
#pragma unroll
        for (int q = 0; q < NF; q++) {
            const double x = A[q] * carry[q] + B[q];
            carry[q] = lane_of(x, 63);
            if (r >= 0 && r < n) out[q][r + 2] = x;
            if (j == 0) x0[q] = lane_of(x, 62);  // x_0 = c_2
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NF; q++) {
            out[q][1] = c1[q];
            out[q][M] = cM[q];
            out[q][0] = 2.0 * c1[q] - x0[q];
            out[q][M + 1] = 2.0 * cM[q] - xl[q];
        }
    }
}

// one z line (see zline_solve_n)
template <int K>
__device__ __forceinline__ void zline_solve(double (&f)[K], int M, const double* __restrict__ piv,
                                            double* __restrict__ out)
{
    double g[1][K];
#pragma unroll
    for (int j = 0; j < K; j++) g[0][j] = f[j];
    double* const o[1] = {out};
    zline_solve_n<K, 1>(g, M, piv, o);
}

template <int K>
__device__ __forceinline__ void prefilter_z_line(long line, const ZArgs& z)
{
    const int lane = threadIdx.x & 63, nt = z.g.nt, nc = z.g.nc;
    const int a = (int)(line / nt), b = (int)(line % nt);
    const long N = z.g.N;
    const long rowoff = ((long)a * N + b) * N;
    if (z.nf == 1) {
        double f[K];
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int p = 64 * j + lane;
            f[j] = p < nt ? field_at(z.W, z.plane, z.q0, z.gam, nc, a, b, p) : 0.0;
        }
        zline_solve<K>(f, nt, z.piv, z.coef[0] + rowoff);
        return;
    }
    double f[3][K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int p = 64 * j + lane;
        if (p < nt) {
            const long c = ((long)wrapi(a - 3, nc) * nc + wrapi(b - 3, nc)) * nc + wrapi(p - 3, nc);
            const double ir = gkm::g_rcp(z.W[c]);
#pragma unroll
            for (int q = 0; q < 3; q++) f[q][j] = z.W[(q + 1) * z.plane + c] * ir;  // as field_at
        } else {
#pragma unroll
            for (int q = 0; q < 3; q++) f[q][j] = 0.0;
        }
    }
    double* const out[3] = {z.coef[0] + rowoff, z.coef[1] + rowoff, z.coef[2] + rowoff};
    zline_solve_n<K, 3>(f, nt, z.piv, out);
}

// in place along z (stride 1; values at the knots of the coefficient box), x
// (stride N) or y (stride N^2)
__device__ __forceinline__ void prefilter_inplace_one(long t, double* __restrict__ coef, const Geom& g, int axis,
                                                      const double* __restrict__ piv)
{
    long base, st;
    if (axis == 0) {  // z lines: a, b in [0, nt)
        if (t >= (long)g.nt * g.nt) return;
        const int a = (int)(t / g.nt), b = (int)(t % g.nt);
        base = ((long)a * g.N + b) * g.N;
        st = 1;
    } else if (axis == 1) {  // x lines: a in [0, nt), c in [0, N)
        if (t >= (long)g.nt * g.N) return;
        const int a = (int)(t / g.N), c = (int)(t % g.N);
        base = (long)a * g.N * g.N + c;
        st = g.N;
    } else {          // y lines: b, c in [0, N)
        if (t >= (long)g.N * g.N) return;
        base = t;
        st = (long)g.N * g.N;
    }
    line_solve_blocked<kLineBlock>(g.nt, coef + base, st, piv);
}

// ---------------------------------------------------------------------------
// Evaluation (einspline eval_UBspline_3d_d / _vg semantics): u = (p-start)/delta,
// i = floor(u), t = u - i, weights of the uniform cubic B-spline.  Indices are
// clamped to the coefficient box so no input can read outside it; every point
// of the padded grid range is inside the box unclamped.
// ---------------------------------------------------------------------------
struct Loc {
    int i;
    double t;
    double w[4], d[4];
};

__device__ __forceinline__ Loc locate(double p, const Geom& g, bool deriv)
{
    double u = (p - g.start) * g.dinv;
    double ip = floor(u);
    double t = u - ip;
    int i = (int)ip;
    if (i < 0) { i = 0; t = 0.0; }
    if (i > g.N - 4) { i = g.N - 4; t = 1.0; }
    Loc L;
    L.i = i;
    L.t = t;
    const double t2 = t * t, t3 = t2 * t;
    const double s6 = 1.0 / 6.0;
    L.w[0] = (-t3 + 3 * t2 - 3 * t + 1) * s6;
    L.w[1] = (3 * t3 - 6 * t2 + 4) * s6;
    L.w[2] = (-3 * t3 + 3 * t2 + 3 * t + 1) * s6;
    L.w[3] = t3 * s6;
    if (deriv) {
        L.d[0] = (-0.5 * t2 + t - 0.5) * g.dinv;
        L.d[1] = (1.5 * t2 - 2 * t) * g.dinv;
        L.d[2] = (-1.5 * t2 + t + 0.5) * g.dinv;
        L.d[3] = 0.5 * t2 * g.dinv;
    }
    return L;
}

// The tricubic gathers of a tracer are done by the four lanes of a quad
// (k = lane & 3; the whole quad calls together, with the same p): lane k reads
// element k of each of the stencil's 4 x 4 rows of 4 coefficients along z, so
// one load instruction of a wave touches the cache lines of 16 tracers rather
// than 64 (one line per lane and load before); each lane sums its column over
// the 16 rows and two quad exchanges add the four columns with their z
// weights.  The sums are the same terms as eval_UBspline_3d_d_vg's, grouped
// column first; every lane of the quad ends with the same values.
__device__ __forceinline__ double quad_sum(double v)
{
    v += dpp64<0xB1, 0xf>(0.0, v);         // quad_perm [1, 0, 3, 2]
    return v + dpp64<0x4E, 0xf>(0.0, v);  // quad_perm [2, 3, 0, 1]
}

__device__ __forceinline__ double pick4(const double w[4], int k)
{
    return k == 0 ? w[0] : (k == 1 ? w[1] : (k == 2 ? w[2] : w[3]));
}

__device__ __forceinline__ void sample_quad(long p, int k, const double* __restrict__ pos,
                                            const double* __restrict__ cu, const double* __restrict__ cv,
                                            const double* __restrict__ cw, const Geom& g, double* __restrict__ U,
                                            double* __restrict__ G)
{
    if (g.drop && isnan(pos[3 * p])) {  // a dropped tracer contributes nothing
        if (k == 0) {
#pragma unroll
            for (int f = 0; f < 3; f++) {
                if (U) U[3 * p + f] = 0.0;
                if (G) G[9 * p + 3 * f + 0] = G[9 * p + 3 * f + 1] = G[9 * p + 3 * f + 2] = 0.0;
            }
        }
        return;
    }
    const Loc Ly = locate(pos[3 * p + 0], g, true);
    const Loc Lx = locate(pos[3 * p + 1], g, true);
    const Loc Lz = locate(pos[3 * p + 2], g, true);
    const double wz = pick4(Lz.w, k), dz = pick4(Lz.d, k);
    const long N = g.N;
    const long c0 = ((long)Ly.i * N + Lx.i) * N + Lz.i + k;
    const double* cs[3] = {cu, cv, cw};
#pragma unroll
    for (int f = 0; f < 3; f++) {
        // column k of the value, d/dy and d/dx sums, row by row: the sums over
        // x of a row with the x weights and their derivatives, then the y
        // weights (16 live weights where the products wy_a wx_b take 48: 64
        // VGPRs against 136)
        double P = 0, Py = 0, Px = 0;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            double s = 0, sx = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const double c = cs[f][c0 + ((long)a * N + b) * N];
                s += Lx.w[b] * c;
                sx += Lx.d[b] * c;
            }
            P += Ly.w[a] * s;
            Py += Ly.d[a] * s;
            Px += Ly.w[a] * sx;
        }
        const double v = quad_sum(wz * P), gy = quad_sum(wz * Py), gx = quad_sum(wz * Px), gz = quad_sum(dz * P);
        if (k == 0) {
            if (U) U[3 * p + f] = v;
            if (G) {
                G[9 * p + 3 * f + 0] = gy;
                G[9 * p + 3 * f + 1] = gx;
                G[9 * p + 3 * f + 2] = gz;
            }
        }
    }
}

// Heun RK2 of Particle::updatePosition (ParticleTrack.cpp:48-73): predictor
// x + U dt (not wrapped) sampled on the new field, corrector, wrap to [0, 2pi].
// Called by the four lanes of a quad like sample_quad; lane 0 stores.
__device__ __forceinline__ void heun_quad(long p, int k, double* __restrict__ pos, const double* __restrict__ U,
                                          const double* __restrict__ cu, const double* __restrict__ cv,
                                          const double* __restrict__ cw, const Geom& g, double dt)
{
    const double u = U[3 * p + 0], v = U[3 * p + 1], w = U[3 * p + 2];
    double y = pos[3 * p + 0], x = pos[3 * p + 1], z = pos[3 * p + 2];
    if (g.drop && isnan(y)) return;  // dropped earlier
    const Loc Ly = locate(y + v * dt, g, false);
    const Loc Lx = locate(x + u * dt, g, false);
    const Loc Lz = locate(z + w * dt, g, false);
    const double wz = pick4(Lz.w, k);
    const long N = g.N;
    const long c0 = ((long)Ly.i * N + Lx.i) * N + Lz.i + k;
    const double* cs[3] = {cu, cv, cw};
    double U2[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        double P = 0;
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) P += (Ly.w[a] * Lx.w[b]) * cs[f][c0 + ((long)a * N + b) * N];
        U2[f] = quad_sum(wz * P);
    }
    if (k != 0) return;
    y += 0.5 * (v + U2[1]) * dt;
    x += 0.5 * (u + U2[0]) * dt;
    z += 0.5 * (w + U2[2]) * dt;
    const double twopi = 2.0 * 3.141592653589793238462643383279502884;
    double r[3] = {y, x, z};
    bool out = false;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (r[c] > twopi) r[c] = fmod(r[c], twopi);
        if (r[c] < 0) r[c] = twopi - fmod(-r[c], twopi);
        out |= (int)(r[c] / twopi) != 0;  // block id with procDim 1 (ParticleTrack.cpp:24-25)
    }
    const double nan = __builtin_nan("");
#pragma unroll
    for (int c = 0; c < 3; c++) pos[3 * p + c] = (g.drop && out) ? nan : r[c];
}

__device__ __forceinline__ int ens_entry(const int* __restrict__ first, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace lpt
