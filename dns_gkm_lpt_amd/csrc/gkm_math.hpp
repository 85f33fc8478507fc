// gkm_math.hpp -- pointwise math of the GKM-WENO step for gfx950 kernels.
//
// Everything here is __host__ __device__ and templated on the arithmetic type
// T (double for the FP64 path, float for the mixed-precision path), so the same
// source runs in the sweep kernels and in host-side unit checks.
//
// Reference semantics followed (nishantp9/DNS_GKM_LPT, src/):
//   c2p                GKMflux.cpp:302-309
//   WENO-SYMOO         update.cpp:556-680 (4 candidate stencils, linear weights
//                      1/20, 9/20, 9/20, 1/20, IS3 := max IS, alpha = d/(eps+IS),
//                      energy reconstructed through temperature)
//   collision time     update.cpp:276-278 (Sutherland, left cell of the face)
//   Maxwellian moment tables (GKMflux.cpp:40-68), used by the GKS flux in
//   gkm_flux.hpp.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace gkm {

template <typename T>
struct Gas {
    T gam;      // ratio of specific heats
    T R;        // gas constant (param.h:21)
    T inv_prn;  // 1/Pr (param.h:19)
    T mu0;      // Sutherland reference viscosity (FlowField.cpp:50)
    T T0;       // reference temperature p0/(R rho0) (FlowField.cpp:47)
    T eps;      // WENO epsilon (param.h:18)
    int K;      // internal degrees of freedom (FlowField.cpp:42)
};

template <typename T>
__host__ __device__ __forceinline__ T sq(T a) { return a * a; }

template <typename T> __host__ __device__ __forceinline__ T g_sqrt(T x) { return sqrt(x); }
template <typename T> __host__ __device__ __forceinline__ T g_exp(T x) { return exp(x); }
template <typename T> __host__ __device__ __forceinline__ T g_erfc(T x) { return erfc(x); }
template <> __host__ __device__ __forceinline__ float g_sqrt<float>(float x) { return sqrtf(x); }
template <> __host__ __device__ __forceinline__ float g_exp<float>(float x) { return expf(x); }
template <> __host__ __device__ __forceinline__ float g_erfc<float>(float x) { return erfcf(x); }

// 1/b and a/b.  FP64 device code skips the correctly rounded division's
// operand scaling and fix-up (v_div_scale, v_div_fmas, v_div_fixup; 11 VALU
// instructions on gfx950): the hardware reciprocal is refined by two Newton
// steps r <- r + r (1 - b r), each squaring the relative error, and a/b takes
// the residual correction q <- q + r (a - b q) of q = a r (Markstein): 5 and 8
// instructions.  On the MI355X both equal the correctly rounded results on
// 4.2e6 operand pairs with magnitudes log-uniform in [e^-40, e^40] (no
// difference at all); the error bound of the method is 1 ulp for normal
// operands with a normal quotient.  A zero, infinite or NaN divisor gives NaN
// (the flux's states are positive and finite, and a non-physical state is
// caught by the CFL reduction).  Host code and other types divide.
template <typename T> __host__ __device__ __forceinline__ T g_rcp(T b) { return T(1) / b; }
template <typename T> __host__ __device__ __forceinline__ T g_div(T a, T b) { return a / b; }
template <>
__host__ __device__ __forceinline__ double g_rcp<double>(double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    return fma(fma(-b, r, 1.0), r, r);
#else
    return 1.0 / b;
#endif
}
template <>
__host__ __device__ __forceinline__ double g_div<double>(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = g_rcp(b);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
#else
    return a / b;
#endif
}

// 0.5 erfc(x) given E = exp(-x^2), which the half-space moments of a
// Maxwellian form anyway (gkm_flux.hpp).  erfc(t) = E erfcx(t) for t = |x|,
// erfc(-t) = 2 - erfc(t), and erfcx(t) = 2/(1 + 2t) P((t - 3)/(t + 3)) with P
// of degree 24: a Chebyshev fit in (t - 3)/(t + 3) in [-1, 1) of 40-digit
// erfcx values (dns_gkm_lpt_amd/erfcx_fit.py prints these coefficients; the
// transform follows Weideman, SIAM J. Numer. Anal. 31 (1994) 1497-1518).
// Largest absolute difference from the C library's erfc 5.6e-16 (about 5 ulp
// of 0.5, near x = 0 where the polynomial's alternating terms cancel; FP64,
// this operation order; tests/host/erfc_check.cpp).  About 50 VALU instructions with the
// shared exponential, against 140 for erfc itself on gfx950.  Other types
// (FP32 diagnostic path, dual numbers) call erfc.
constexpr double kErfcxP[25] = {
    0.6265040291348647,
    -0.06781055306229068,
    -0.023781147176723055,
    0.06482257935138193,
    -0.05963683170780693,
    0.03415401367197528,
    -0.01188525745600576,
    0.001264695907930164,
    0.0009443451443442583,
    -0.00038991613413261634,
    -5.2579684014024e-05,
    6.254966469219208e-05,
    2.473913930914913e-06,
    -1.0050245745815767e-05,
    -3.6367751223556406e-07,
    1.7361279106766794e-06,
    1.7496135370769662e-07,
    -3.0821947269270045e-07,
    -6.701414299241337e-08,
    5.084828360580839e-08,
    1.8283091137946232e-08,
    -6.622869822921796e-09,
    -3.2107595865462734e-09,
    4.862158760421362e-10,
    2.6325770446883475e-10,
};

template <typename T>
__host__ __device__ __forceinline__ T half_erfc(T x, T E)
{
    (void)E;
    return T(0.5) * g_erfc(x);
}
template <>
__host__ __device__ __forceinline__ double half_erfc<double>(double x, double E)
{
    const double t = fabs(x);
    // one division for both y = (t - 3)/(t + 3) and 1/(1 + 2t)
    const double a = t + 3.0, b = 1.0 + 2.0 * t;
    const double r = g_rcp(a * b);
    const double y = (t - 3.0) * b * r;
    double acc = kErfcxP[24];
#pragma unroll
    for (int i = 23; i >= 0; i--) acc = acc * y + kErfcxP[i];
    const double v = E * acc * a * r;  // 0.5 erfc(t) = 0.5 E 2/(1 + 2t) P(y)
    return x >= 0.0 ? v : 1.0 - v;
}

// ---------------------------------------------------------------------------
// WENO-SYMOO: reconstruction at face i+1/2 seen from the left, from cells
// a..f = i-2..i+3 (update.cpp:561-621).  The right state is the same formula
// on the mirrored stencil (f..a).
// ---------------------------------------------------------------------------
template <typename T>
__host__ __device__ __forceinline__ T weno_side(T a, T b, T c, T d, T e, T f, T eps)
{
    const T r3 = T(1) / T(3), r6 = T(1) / T(6), r12 = T(1) / T(12);
    T q0 = a * r3 - T(7) * b * r6 + T(11) * c * r6;
    T q1 = -b * r6 + T(5) * c * r6 + d * r3;
    T q2 = c * r3 + T(5) * d * r6 - e * r6;
    T q3 = T(11) * d * r6 - T(7) * e * r6 + f * r3;
    T s0 = T(13) * sq(a - T(2) * b + c) * r12 + sq(a - T(4) * b + T(3) * c) * T(0.25);
    T s1 = T(13) * sq(b - T(2) * c + d) * r12 + sq(-b + d) * T(0.25);
    T s2 = T(13) * sq(c - T(2) * d + e) * r12 + sq(-T(3) * c + T(4) * d - e) * T(0.25);
    T s3 = T(13) * sq(d - T(2) * e + f) * r12 + sq(-T(5) * d + T(8) * e - T(3) * f) * T(0.25);
    T smax = s0 > s1 ? s0 : s1;
    smax = smax > s2 ? smax : s2;
    s3 = smax > s3 ? smax : s3;  // update.cpp:577-584
    // alpha_k = d_k/(eps+IS_k) normalised: with D_j = eps+IS_j the weights are
    // d_k prod_{j!=k} D_j / sum_m d_m prod_{j!=m} D_j -- one division.
    const T D0 = eps + s0, D1 = eps + s1, D2 = eps + s2, D3 = eps + s3;
    const T p01 = D0 * D1, p23 = D2 * D3;
    const T n0 = T(0.05) * D1 * p23, n1 = T(0.45) * D0 * p23;
    const T n2 = T(0.45) * p01 * D3, n3 = T(0.05) * p01 * D2;
    return (n0 * q0 + n1 * q1 + n2 * q2 + n3 * q3) / (n0 + n1 + n2 + n3);
}

// Pieces of weno_pair.  wmax: the larger of two indicators (fmax for float
// and double: one instruction, equal to a > b ? a : b for non-NaN values).
// wratio: the weights' normalising division (g_div); FP32 (mixed precision,
// device) multiplies by the hardware reciprocal (1 ulp) instead of dividing.
template <typename T>
__host__ __device__ __forceinline__ T wmax(T a, T b) { return a > b ? a : b; }
template <> __host__ __device__ __forceinline__ float wmax<float>(float a, float b) { return fmaxf(a, b); }
template <> __host__ __device__ __forceinline__ double wmax<double>(double a, double b) { return fmax(a, b); }
template <typename T>
__host__ __device__ __forceinline__ T wratio(T n, T d) { return g_div(n, d); }
__device__ __forceinline__ float wratio(float n, float d) { return n * __builtin_amdgcn_rcpf(d); }

// nonlinear combination of the four candidates (weno_side's second half)
template <typename T>
__host__ __device__ __forceinline__ T weno_combine(T q0, T q1, T q2, T q3, T s0, T s1, T s2, T s3, T eps)
{
    s3 = wmax(wmax(wmax(s0, s1), s2), s3);  // update.cpp:577-584
    const T D0 = eps + s0, D1 = eps + s1, D2 = eps + s2, D3 = eps + s3;
    const T p01 = D0 * D1, p23 = D2 * D3;
    const T n0 = T(0.05) * D1 * p23, n1 = T(0.45) * D0 * p23;
    const T n2 = T(0.45) * p01 * D3, n3 = T(0.05) * p01 * D2;
    return wratio(n0 * q0 + n1 * q1 + n2 * q2 + n3 * q3, n0 + n1 + n2 + n3);
}

// Both reconstructions of one face in one pass: wl = weno_side(a..f) and
// wr = weno_side(f..a), the same formulas with the rational coefficients
// folded into single constants (b * 7/6 rather than 7 b / 6, and half of each
// first-difference combination squared in place of a quarter of its square)
// and the four 13/12-weighted squared second differences, which the two sides
// share, formed once.  Equal to the two weno_side calls up to rounding.
template <typename T>
__host__ __device__ __forceinline__ void weno_pair(T a, T b, T c, T d, T e, T f, T eps, T& wl, T& wr)
{
    const T k3 = T(1.0 / 3.0), k6 = T(1.0 / 6.0), k56 = T(5.0 / 6.0), k76 = T(7.0 / 6.0), k116 = T(11.0 / 6.0);
    const T k13 = T(13.0 / 12.0);
    const T t0 = a - T(2) * b + c, t1 = b - T(2) * c + d, t2 = c - T(2) * d + e, t3 = d - T(2) * e + f;
    const T K0 = k13 * t0 * t0, K1 = k13 * t1 * t1, K2 = k13 * t2 * t2, K3 = k13 * t3 * t3;
    {
        const T q0 = a * k3 - b * k76 + c * k116;
        const T q1 = c * k56 - b * k6 + d * k3;
        const T q2 = c * k3 + d * k56 - e * k6;
        const T q3 = d * k116 - e * k76 + f * k3;
        const T h0 = T(0.5) * a - T(2) * b + T(1.5) * c;
        const T h1 = T(0.5) * (d - b);
        const T h2 = T(2) * d - T(1.5) * c - T(0.5) * e;
        const T h3 = T(4) * e - T(2.5) * d - T(1.5) * f;
        wl = weno_combine(q0, q1, q2, q3, K0 + h0 * h0, K1 + h1 * h1, K2 + h2 * h2, K3 + h3 * h3, eps);
    }
    {
        const T q0 = f * k3 - e * k76 + d * k116;
        const T q1 = d * k56 - e * k6 + c * k3;
        const T q2 = d * k3 + c * k56 - b * k6;
        const T q3 = c * k116 - b * k76 + a * k3;
        const T h0 = T(0.5) * f - T(2) * e + T(1.5) * d;
        const T h1 = T(0.5) * (c - e);
        const T h2 = T(2) * c - T(1.5) * d - T(0.5) * b;
        const T h3 = T(4) * b - T(2.5) * c - T(1.5) * a;
        wr = weno_combine(q0, q1, q2, q3, K3 + h0 * h0, K2 + h1 * h1, K1 + h2 * h2, K0 + h3 * h3, eps);
    }
}

// temperature with (gam - 1)/R given (one division)
template <typename T>
__host__ __device__ __forceinline__ T temp_gr(const T W[5], T gR)
{
    const T ir = g_rcp(W[0]);
    return gR * (W[4] - T(0.5) * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) * ir) * ir;
}

// temperature of a conservative state (update.cpp:624-629)
template <typename T>
__host__ __device__ __forceinline__ T temp_of(const T W[5], T gam, T R)
{
    T ir = T(1) / W[0];
    return (gam - T(1)) * (W[4] * ir - T(0.5) * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) * ir * ir) / R;
}

// ApplyWENO (update.cpp:556-680): S[s][q] for cells i-2..i+3, out Wl, Wr.
// Temperatures with (gam-1)/R formed once, and the energy rebuilt with
// R/(gam-1); both equal the reference's expressions up to rounding.
// weno5 with the six cells' temperatures given (Tc[s] = temp_gr(S[s], (gam - 1)/R)):
// two neighbouring faces of a line share five cells and their temperatures
template <typename T>
__host__ __device__ __forceinline__ void weno5_t(const T S[6][5], const T Tc[6], const Gas<T>& g, T Wl[5], T Wr[5])
{
#pragma unroll
    for (int q = 0; q < 4; q++) weno_pair(S[0][q], S[1][q], S[2][q], S[3][q], S[4][q], S[5][q], g.eps, Wl[q], Wr[q]);
    T Tl, Tr;
    weno_pair(Tc[0], Tc[1], Tc[2], Tc[3], Tc[4], Tc[5], g.eps, Tl, Tr);
    const T Rg = g.R / (g.gam - T(1));
    Wl[4] = Wl[0] * Rg * Tl + g_div(T(0.5) * (Wl[1] * Wl[1] + Wl[2] * Wl[2] + Wl[3] * Wl[3]), Wl[0]);
    Wr[4] = Wr[0] * Rg * Tr + g_div(T(0.5) * (Wr[1] * Wr[1] + Wr[2] * Wr[2] + Wr[3] * Wr[3]), Wr[0]);
}

template <typename T>
__host__ __device__ __forceinline__ void weno5(const T S[6][5], const Gas<T>& g, T Wl[5], T Wr[5])
{
#pragma unroll
    for (int q = 0; q < 4; q++) weno_pair(S[0][q], S[1][q], S[2][q], S[3][q], S[4][q], S[5][q], g.eps, Wl[q], Wr[q]);
    const T gR = (g.gam - T(1)) / g.R;
    T Tc[6];
#pragma unroll
    for (int s = 0; s < 6; s++) Tc[s] = temp_gr(S[s], gR);
    T Tl, Tr;
    weno_pair(Tc[0], Tc[1], Tc[2], Tc[3], Tc[4], Tc[5], g.eps, Tl, Tr);
    const T Rg = g.R / (g.gam - T(1));
    Wl[4] = Wl[0] * Rg * Tl + g_div(T(0.5) * (Wl[1] * Wl[1] + Wl[2] * Wl[2] + Wl[3] * Wl[3]), Wl[0]);
    Wr[4] = Wr[0] * Rg * Tr + g_div(T(0.5) * (Wr[1] * Wr[1] + Wr[2] * Wr[2] + Wr[3] * Wr[3]), Wr[0]);
}

// collision time mu(T)/p of the face's left cell, Sutherland law (update.cpp:276-278)
template <typename T>
__host__ __device__ __forceinline__ T collision_time(const T W[5], const Gas<T>& g)
{
    T ir = g_rcp(W[0]);
    T u = W[1] * ir, v = W[2] * ir, w = W[3] * ir;
    T P = (W[0] * (g.gam - T(1))) * (W[4] * ir - T(0.5) * (u * u + v * v + w * w));
    T Tt = g_div(P, g.R * W[0]);
    T x = g_div(Tt, g.T0);
    T mu = g_div(g.mu0 * (x * g_sqrt(x)) * (g.T0 + T(110.4)), Tt + T(110.4));
    return g_div(mu, P);
}

// ---------------------------------------------------------------------------
// Maxwellian moments.  For a state with mean (U,V,W) and lambda, the full
// moments <u^n> satisfy f_n = U f_{n-1} + (n-1)/(2 lambda) f_{n-2}, f_0 = 1,
// f_1 = U; the half-space moments over u>0 (u<0) start from
// 0.5 erfc(-+sqrt(lambda) U) and U h_0 +- exp(-lambda U^2)/(2 sqrt(pi lambda))
// and follow the same recursion (GKMflux.cpp:50-68).  Internal dof:
// <xi^2> = K/(2 lambda), <xi^4> = K(K+2)/(4 lambda^2) (GKMflux.cpp:45-48).
// ---------------------------------------------------------------------------
template <typename T>
struct Maxw {
    T den, u, v, w, lam, i2l, e2, e4;
};

template <typename T>
__host__ __device__ __forceinline__ void maxw_init(Maxw<T>& s, const T W[5], const Gas<T>& g)
{
    s.den = W[0];
    T ir = g_rcp(W[0]);
    s.u = W[1] * ir;
    s.v = W[2] * ir;
    s.w = W[3] * ir;
    T P = (s.den * (g.gam - T(1))) * (W[4] * ir - T(0.5) * (s.u * s.u + s.v * s.v + s.w * s.w));
    s.i2l = P * ir;  // 1/(2 lambda) = p/rho
    s.lam = T(0.5) * g_rcp(s.i2l);
    s.e2 = T(g.K) * s.i2l;
    s.e4 = T(g.K * (g.K + 2)) * s.i2l * s.i2l;
}

template <int N, typename T>
__host__ __device__ __forceinline__ void full_tab(T m, T i2l, T* f)
{
    f[0] = T(1);
    f[1] = m;
#pragma unroll
    for (int n = 2; n < N; n++) f[n] = m * f[n - 1] + T(n - 1) * i2l * f[n - 2];
}

template <int N, typename T>
__host__ __device__ __forceinline__ void half_tab(T m, T i2l, T h0, T h1, T* h)
{
    h[0] = h0;
    h[1] = h1;
#pragma unroll
    for (int n = 2; n < N; n++) h[n] = m * h[n - 1] + T(n - 1) * i2l * h[n - 2];
}

// <u^k v^l w^m psi>, psi = (1, u, v, w, (u^2+v^2+w^2+xi^2)/2)  (GKMflux.cpp:269-278)
// ===========================================================================
// Mixed precision.  FP64 where cancellation would cost digits: the stencil
// differences fed to WENO, primitive recovery (c2p, lambda = rho/(2p)), the
// equilibrium state We, the Euler part of the flux (the p0 and p3 terms,
// GKMflux.cpp:225-226, which carry dt times the inviscid flux) and the
// collision coefficients.  FP32 for WENO smoothness indicators and weights
// and for every slope-dependent moment term of the flux (the flux's split is
// gks_flux_tf<double, float>, gkm_flux.hpp).
// ===========================================================================

// WENO on FP64 stencils, FP32 arithmetic on differences from cell i, the base
// of both sides (the reconstruction is shift invariant: candidates are affine
// with unit coefficient sum and the indicators see differences only).
// weno5_mixed with the six cells' temperatures given (as weno5_t)
__host__ __device__ __forceinline__ void weno5_mixed_t(const double S[6][5], const double Tc[6], const Gas<double>& g,
                                                       double Wl[5], double Wr[5])
{
    const float eps = (float)g.eps;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double c = S[2][q];
        float l, r;
        weno_pair<float>((float)(S[0][q] - c), (float)(S[1][q] - c), 0.f, (float)(S[3][q] - c), (float)(S[4][q] - c),
                         (float)(S[5][q] - c), eps, l, r);
        Wl[q] = c + (double)l;
        Wr[q] = c + (double)r;
    }
    float l, r;
    weno_pair<float>((float)(Tc[0] - Tc[2]), (float)(Tc[1] - Tc[2]), 0.f, (float)(Tc[3] - Tc[2]),
                     (float)(Tc[4] - Tc[2]), (float)(Tc[5] - Tc[2]), eps, l, r);
    const double Tl = Tc[2] + (double)l, Tr = Tc[2] + (double)r;
    const double Rg = g.R / (g.gam - 1.0);
    Wl[4] = Wl[0] * Rg * Tl + g_div(0.5 * (Wl[1] * Wl[1] + Wl[2] * Wl[2] + Wl[3] * Wl[3]), Wl[0]);
    Wr[4] = Wr[0] * Rg * Tr + g_div(0.5 * (Wr[1] * Wr[1] + Wr[2] * Wr[2] + Wr[3] * Wr[3]), Wr[0]);
}

__host__ __device__ __forceinline__ void weno5_mixed(const double S[6][5], const Gas<double>& g, double Wl[5],
                                                     double Wr[5])
{
    const float eps = (float)g.eps;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double c = S[2][q];
        float l, r;
        weno_pair<float>((float)(S[0][q] - c), (float)(S[1][q] - c), 0.f, (float)(S[3][q] - c), (float)(S[4][q] - c),
                         (float)(S[5][q] - c), eps, l, r);
        Wl[q] = c + (double)l;
        Wr[q] = c + (double)r;
    }
    const double gR = (g.gam - 1.0) / g.R;
    double Tc[6];
#pragma unroll
    for (int s = 0; s < 6; s++) Tc[s] = temp_gr<double>(S[s], gR);
    float l, r;
    weno_pair<float>((float)(Tc[0] - Tc[2]), (float)(Tc[1] - Tc[2]), 0.f, (float)(Tc[3] - Tc[2]),
                     (float)(Tc[4] - Tc[2]), (float)(Tc[5] - Tc[2]), eps, l, r);
    const double Tl = Tc[2] + (double)l, Tr = Tc[2] + (double)r;
    const double Rg = g.R / (g.gam - 1.0);
    Wl[4] = Wl[0] * Rg * Tl + g_div(0.5 * (Wl[1] * Wl[1] + Wl[2] * Wl[2] + Wl[3] * Wl[3]), Wl[0]);
    Wr[4] = Wr[0] * Rg * Tr + g_div(0.5 * (Wr[1] * Wr[1] + Wr[2] * Wr[2] + Wr[3] * Wr[3]), Wr[0]);
}

__host__ __device__ __forceinline__ Gas<float> to_f32(const Gas<double>& g)
{
    return Gas<float>{(float)g.gam, (float)g.R, (float)g.inv_prn, (float)g.mu0, (float)g.T0, (float)g.eps, g.K};
}

}  // namespace gkm
