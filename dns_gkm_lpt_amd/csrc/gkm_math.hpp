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
//   GKS flux           GKMflux.cpp:11-229 with MCal GKMflux.cpp:238-278 and
//                      slopesolver GKMflux.cpp:283-297 (a[4] scaled by 1/Pr)
// The flux is reorganised by Maxwellian state: the interface terms that belong
// to the left state, the right state and the equilibrium state are each summed
// in one block while that state's moment tables are live, and the collision
// coefficients (which need only the densities and lambdas) are computed first.
// This changes the order of the floating-point sums, not the terms.
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
// wratio: the weights' normalising division; FP32 (mixed precision, device)
// multiplies by the hardware reciprocal (1 ulp) instead of dividing.
template <typename T>
__host__ __device__ __forceinline__ T wmax(T a, T b) { return a > b ? a : b; }
template <> __host__ __device__ __forceinline__ float wmax<float>(float a, float b) { return fmaxf(a, b); }
template <> __host__ __device__ __forceinline__ double wmax<double>(double a, double b) { return fmax(a, b); }
template <typename T>
__host__ __device__ __forceinline__ T wratio(T n, T d) { return n / d; }
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
    const T ir = T(1) / W[0];
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
    Wl[4] = Wl[0] * Rg * Tl + T(0.5) * (Wl[1] * Wl[1] + Wl[2] * Wl[2] + Wl[3] * Wl[3]) / Wl[0];
    Wr[4] = Wr[0] * Rg * Tr + T(0.5) * (Wr[1] * Wr[1] + Wr[2] * Wr[2] + Wr[3] * Wr[3]) / Wr[0];
}

// collision time mu(T)/p of the face's left cell, Sutherland law (update.cpp:276-278)
template <typename T>
__host__ __device__ __forceinline__ T collision_time(const T W[5], const Gas<T>& g)
{
    T ir = T(1) / W[0];
    T u = W[1] * ir, v = W[2] * ir, w = W[3] * ir;
    T P = (W[0] * (g.gam - T(1))) * (W[4] * ir - T(0.5) * (u * u + v * v + w * w));
    T Tt = P / (g.R * W[0]);
    T x = Tt / g.T0;
    T mu = g.mu0 * (x * g_sqrt(x)) * (g.T0 + T(110.4)) / (Tt + T(110.4));
    return mu / P;
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
    T ir = T(1) / W[0];
    s.u = W[1] * ir;
    s.v = W[2] * ir;
    s.w = W[3] * ir;
    T P = (s.den * (g.gam - T(1))) * (W[4] * ir - T(0.5) * (s.u * s.u + s.v * s.v + s.w * s.w));
    s.i2l = P * ir;  // 1/(2 lambda) = p/rho
    s.lam = T(0.5) / s.i2l;
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
template <int k, int l, int m, typename T>
__host__ __device__ __forceinline__ void mom(const T* tu, const T* tv, const T* tw, T e2, T M[5])
{
    T t = tu[k] * tv[l] * tw[m];
    M[0] = t;
    M[1] = tu[k + 1] * tv[l] * tw[m];
    M[2] = tu[k] * tv[l + 1] * tw[m];
    M[3] = tu[k] * tv[l] * tw[m + 1];
    M[4] = T(0.5) * (tu[k + 2] * tv[l] * tw[m] + tu[k] * tv[l + 2] * tw[m] + tu[k] * tv[l] * tw[m + 2] + t * e2);
}

// acc += c * <u^k v^l w^m psi (a . psi)>  (GKMflux.cpp:238-267)
template <int k, int l, int m, typename T>
__host__ __device__ __forceinline__ void mom_a(const T* tu, const T* tv, const T* tw, T e2, T e4, const T a[5], T c,
                                               T acc[5])
{
#define T3(p, q, r) (tu[p] * tv[q] * tw[r])
    T h0 = T(0.5) * (T3(2 + k, l, m) + T3(k, 2 + l, m) + T3(k, l, 2 + m) + T3(k, l, m) * e2);
    T h1 = T(0.5) * (T3(3 + k, l, m) + T3(1 + k, 2 + l, m) + T3(1 + k, l, 2 + m) + T3(1 + k, l, m) * e2);
    T h2 = T(0.5) * (T3(2 + k, 1 + l, m) + T3(k, 3 + l, m) + T3(k, 1 + l, 2 + m) + T3(k, 1 + l, m) * e2);
    T h3 = T(0.5) * (T3(2 + k, l, 1 + m) + T3(k, 2 + l, 1 + m) + T3(k, l, 3 + m) + T3(k, l, 1 + m) * e2);
    T m0 = a[0] * T3(k, l, m) + a[1] * T3(1 + k, l, m) + a[2] * T3(k, 1 + l, m) + a[3] * T3(k, l, 1 + m) + a[4] * h0;
    T m1 = a[0] * T3(1 + k, l, m) + a[1] * T3(2 + k, l, m) + a[2] * T3(1 + k, 1 + l, m) + a[3] * T3(1 + k, l, 1 + m) +
           a[4] * h1;
    T m2 = a[0] * T3(k, 1 + l, m) + a[1] * T3(1 + k, 1 + l, m) + a[2] * T3(k, 2 + l, m) + a[3] * T3(k, 1 + l, 1 + m) +
           a[4] * h2;
    T m3 = a[0] * T3(k, l, 1 + m) + a[1] * T3(1 + k, l, 1 + m) + a[2] * T3(k, 1 + l, 1 + m) + a[3] * T3(k, l, 2 + m) +
           a[4] * h3;
    T m4 = T(0.25) * a[4] *
               (T3(4 + k, l, m) + T3(k, 4 + l, m) + T3(k, l, 4 + m) + T3(k, l, m) * e4 +
                T(2) * (T3(2 + k, 2 + l, m) + T3(2 + k, l, 2 + m) + T3(k, 2 + l, 2 + m)) +
                T(2) * e2 * (T3(2 + k, l, m) + T3(k, 2 + l, m) + T3(k, l, 2 + m))) +
           a[0] * h0 + a[1] * h1 + a[2] * h2 + a[3] * h3;
#undef T3
    acc[0] += c * m0;
    acc[1] += c * m1;
    acc[2] += c * m2;
    acc[3] += c * m3;
    acc[4] += c * m4;
}

// micro slope a from macro slope b (GKMflux.cpp:283-297), a[4] scaled by 1/Pr
template <typename T>
__host__ __device__ __forceinline__ void slope(const T b[5], const Maxw<T>& s, const Gas<T>& g, T a[5])
{
    T R2 = b[1] - s.u * b[0];
    T R3 = b[2] - s.v * b[0];
    T R4 = b[3] - s.w * b[0];
    T q = s.u * s.u + s.v * s.v + s.w * s.w + T(g.K + 3) * s.i2l;
    T R5 = T(2) * b[4] - b[0] * q;
    T lam2 = T(2) * s.lam;
    // (4 lambda^2)/(K+3) = lam2^2 * 1/(K+3); K+3 is 5 for gamma = 1.4
    a[4] = g.inv_prn * (R5 - T(2) * (s.u * R2 + s.v * R3 + s.w * R4)) * (lam2 * lam2) * (T(1) / T(g.K + 3));
    a[3] = lam2 * R4 - s.w * a[4];
    a[2] = lam2 * R3 - s.v * a[4];
    a[1] = lam2 * R2 - s.u * a[4];
    a[0] = b[0] - a[1] * s.u - a[2] * s.v - a[3] * s.w - T(0.5) * a[4] * q;
}

// ---------------------------------------------------------------------------
// GKS time-integrated interface flux (GKMflux.cpp:11-229).
//   WL, WR : cell states either side of the face (rotated frame)
//   Wl, Wr : WENO states at the face
//   D[0..5]: DWxl, DWxr, DWyl, DWyr, DWzl, DWzr
//   tau    : collision time of the left cell (the pressure-jump term is added here)
// ---------------------------------------------------------------------------
template <typename T>
__host__ __device__ __forceinline__ void gks_flux(const T WL[5], const T WR[5], const T Wl[5], const T Wr[5],
                                                  const T D[6][5], T tau, T dt, T dx, const Gas<T>& g, T F[5])
{
    const T pi = T(3.141592653589793238462643383279502884);
    Maxw<T> L, R;
    maxw_init(L, Wl, g);
    maxw_init(R, Wr, g);

    // collision time with the artificial pressure-jump term (GKMflux.cpp:156-157)
    {
        T pl = L.den / L.lam, pr = R.den / R.lam;
        tau = tau + dt * (fabs(pl - pr) / fabs(pl + pr));
    }
    // time-integration coefficients (GKMflux.cpp:162-167, 197-204)
    const T ex = g_exp(-dt / tau);
    const T om = T(1) - ex;
    const T igm0 = T(1) / (dt - tau * om);
    const T gm1 = -om * igm0;
    const T gm2 = (-dt + T(2) * tau * om - dt * ex) * igm0;
    const T gm3 = -gm1;
    const T gm4 = (dt * ex - tau * om) * igm0;
    const T gm5 = tau * gm3;
    const T Pc = tau * om;
    const T Qc = -tau * dt * ex + tau * tau * om;
    const T p0 = dt - Pc;
    const T p1 = T(0.5) * dt * dt - tau * p0;
    const T p2 = -tau * p0 + Qc;
    const T p3 = Pc;
    const T p4 = -Qc - tau * Pc;
    const T p5 = -tau * Pc;

    // error-function / Gaussian terms of the half-space tables
    const T gl = T(0.5) * g_exp(-L.lam * L.u * L.u) / g_sqrt(L.lam * pi);
    const T hl0 = T(0.5) * g_erfc(-g_sqrt(L.lam) * L.u);
    const T gr = T(0.5) * g_exp(-R.lam * R.u * R.u) / g_sqrt(R.lam * pi);
    const T hr0 = T(0.5) * g_erfc(g_sqrt(R.lam) * R.u);

    // equilibrium state at the face (GKMflux.cpp:73-80)
    T We[5];
    {
        T hu[3], tv[3], tw[3], M[5];
        half_tab<3>(L.u, L.i2l, hl0, L.u * hl0 + gl, hu);
        full_tab<3>(L.v, L.i2l, tv);
        full_tab<3>(L.w, L.i2l, tw);
        mom<0, 0, 0>(hu, tv, tw, L.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) We[q] = L.den * M[q];
        half_tab<3>(R.u, R.i2l, hr0, R.u * hr0 - gr, hu);
        full_tab<3>(R.v, R.i2l, tv);
        full_tab<3>(R.w, R.i2l, tw);
        mom<0, 0, 0>(hu, tv, tw, R.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) We[q] += R.den * M[q];
    }
    Maxw<T> E;
    maxw_init(E, We, g);

    // Per-state sums.  Be accumulates den_e * B_ (GKMflux.cpp:183-190),
    // F accumulates the flux (GKMflux.cpp:223-228); bye/bze are the
    // phase-II right-hand sides of the equilibrium slopes (GKMflux.cpp:141-144).
    T Be[5] = {0, 0, 0, 0, 0}, bye[5] = {0, 0, 0, 0, 0}, bze[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 5; q++) F[q] = T(0);

    // ---- left (u>0 half of the left Maxwellian) and right (u<0 half of the right).
    // Each table set is live only inside its block: the full u-table gives the
    // time slope A of the state first, then the half table carries every term
    // that integrates over the half space.
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const Maxw<T>& S = side == 0 ? L : R;
        const int dxi = side == 0 ? 0 : 1;  // index of DWx{l,r} etc. in D
        T tv[6], tw[6];
        full_tab<6>(S.v, S.i2l, tv);
        full_tab<6>(S.w, S.i2l, tw);
        const T iden = T(1) / S.den;
        T b[5], ax[5], ay[5], az[5], A[5];
#pragma unroll
        for (int q = 0; q < 5; q++) b[q] = D[0 + dxi][q] * iden;
        slope(b, S, g, ax);
#pragma unroll
        for (int q = 0; q < 5; q++) b[q] = D[2 + dxi][q] * iden;
        slope(b, S, g, ay);
#pragma unroll
        for (int q = 0; q < 5; q++) b[q] = D[4 + dxi][q] * iden;
        slope(b, S, g, az);
        {   // phase II: A from the full-space moments (GKMflux.cpp:134-151)
            T fu[6];
            full_tab<6>(S.u, S.i2l, fu);
#pragma unroll
            for (int q = 0; q < 5; q++) b[q] = T(0);
            mom_a<1, 0, 0>(fu, tv, tw, S.e2, S.e4, ax, T(-1), b);
            mom_a<0, 1, 0>(fu, tv, tw, S.e2, S.e4, ay, T(-1), b);
            mom_a<0, 0, 1>(fu, tv, tw, S.e2, S.e4, az, T(-1), b);
            slope(b, S, g, A);
        }
        T hu[7];
        if (side == 0)
            half_tab<7>(S.u, S.i2l, hl0, S.u * hl0 + gl, hu);
        else
            half_tab<7>(S.u, S.i2l, hr0, S.u * hr0 - gr, hu);
        // phase II right-hand sides of the equilibrium slopes (GKMflux.cpp:130-133, 141-144)
        mom_a<0, 0, 0>(hu, tv, tw, S.e2, S.e4, ay, T(1), bye);
        mom_a<0, 0, 0>(hu, tv, tw, S.e2, S.e4, az, T(1), bze);
        // this state's share of den_e * B_ and of F
        const T cB = (gm4 + gm5) * S.den;
        T M[5];
        mom<0, 0, 0>(hu, tv, tw, S.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) Be[q] += gm3 * S.den * M[q];
        mom_a<1, 0, 0>(hu, tv, tw, S.e2, S.e4, ax, cB, Be);
        mom_a<0, 1, 0>(hu, tv, tw, S.e2, S.e4, ay, cB, Be);
        mom_a<0, 0, 1>(hu, tv, tw, S.e2, S.e4, az, cB, Be);
        mom_a<0, 0, 0>(hu, tv, tw, S.e2, S.e4, A, gm5 * S.den, Be);
        mom<1, 0, 0>(hu, tv, tw, S.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) F[q] += p3 * S.den * M[q];
        const T cF = p4 * S.den;
        mom_a<2, 0, 0>(hu, tv, tw, S.e2, S.e4, ax, cF, F);
        mom_a<1, 1, 0>(hu, tv, tw, S.e2, S.e4, ay, cF, F);
        mom_a<1, 0, 1>(hu, tv, tw, S.e2, S.e4, az, cF, F);
        mom_a<1, 0, 0>(hu, tv, tw, S.e2, S.e4, A, p5 * S.den, F);
    }

    // ---- equilibrium state: one half-space table at a time, then the full table
    {
        T tv[6], tw[6];
        full_tab<6>(E.v, E.i2l, tv);
        full_tab<6>(E.w, E.i2l, tw);
        const T ge = T(0.5) * g_exp(-E.lam * E.u * E.u) / g_sqrt(E.lam * pi);
        const T sl = g_sqrt(E.lam) * E.u;
        const T c = T(2) / (E.den * dx);
        const T cB = gm2 * E.den;
        const T cF = p2 * E.den;
        T b[5], a[5];
        {   // u>0 half with the left normal slope (GKMflux.cpp:115, 124, 170, 210)
            T h[7];
            const T h0 = T(0.5) * g_erfc(-sl);
            half_tab<7>(E.u, E.i2l, h0, E.u * h0 + ge, h);
#pragma unroll
            for (int q = 0; q < 5; q++) b[q] = (We[q] - WL[q]) * c;
            slope(b, E, g, a);
            mom_a<1, 0, 0>(h, tv, tw, E.e2, E.e4, a, cB, Be);
            mom_a<2, 0, 0>(h, tv, tw, E.e2, E.e4, a, cF, F);
        }
        {   // u<0 half with the right normal slope (GKMflux.cpp:116, 125, 171, 211)
            T h[7];
            const T h0 = T(0.5) * g_erfc(sl);
            half_tab<7>(E.u, E.i2l, h0, E.u * h0 - ge, h);
#pragma unroll
            for (int q = 0; q < 5; q++) b[q] = (WR[q] - We[q]) * c;
            slope(b, E, g, a);
            mom_a<1, 0, 0>(h, tv, tw, E.e2, E.e4, a, cB, Be);
            mom_a<2, 0, 0>(h, tv, tw, E.e2, E.e4, a, cF, F);
        }
        T fu[7];
        full_tab<7>(E.u, E.i2l, fu);
        // tangential slopes (GKMflux.cpp:148-149, 172-173, 212-213)
        slope(bye, E, g, a);
        mom_a<0, 1, 0>(fu, tv, tw, E.e2, E.e4, a, cB, Be);
        mom_a<1, 1, 0>(fu, tv, tw, E.e2, E.e4, a, cF, F);
        slope(bze, E, g, a);
        mom_a<0, 0, 1>(fu, tv, tw, E.e2, E.e4, a, cB, Be);
        mom_a<1, 0, 1>(fu, tv, tw, E.e2, E.e4, a, cF, F);
        // B_ and the time slope A_ (GKMflux.cpp:183-191)
        T M[5];
        mom<0, 0, 0>(fu, tv, tw, E.e2, M);
        const T iden = T(1) / E.den;
#pragma unroll
        for (int q = 0; q < 5; q++) b[q] = (Be[q] + gm1 * E.den * M[q]) * iden;
        slope(b, E, g, a);
        // equilibrium part of the flux (GKMflux.cpp:225)
        mom<1, 0, 0>(fu, tv, tw, E.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) F[q] += p0 * E.den * M[q];
        mom_a<1, 0, 0>(fu, tv, tw, E.e2, E.e4, a, p1 * E.den, F);
    }
}

// ===========================================================================
// Mixed precision.  FP64 where cancellation would cost digits: the stencil
// differences fed to WENO, primitive recovery (c2p, lambda = rho/(2p)), the
// equilibrium state We, the Euler part of the flux (the p0 and p3 terms,
// GKMflux.cpp:225-226, which carry dt times the inviscid flux) and the
// collision coefficients.  FP32 for WENO smoothness indicators and weights
// and for every slope-dependent moment term of the flux.
// ===========================================================================

// WENO on FP64 stencils, FP32 arithmetic on differences from cell i, the base
// of both sides (the reconstruction is shift invariant: candidates are affine
// with unit coefficient sum and the indicators see differences only).
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
    Wl[4] = Wl[0] * Rg * Tl + 0.5 * (Wl[1] * Wl[1] + Wl[2] * Wl[2] + Wl[3] * Wl[3]) / Wl[0];
    Wr[4] = Wr[0] * Rg * Tr + 0.5 * (Wr[1] * Wr[1] + Wr[2] * Wr[2] + Wr[3] * Wr[3]) / Wr[0];
}

template <typename T>
__host__ __device__ __forceinline__ Maxw<float> to_f32(const Maxw<T>& s)
{
    return Maxw<float>{(float)s.den, (float)s.u, (float)s.v, (float)s.w, (float)s.lam, (float)s.i2l, (float)s.e2,
                       (float)s.e4};
}

__host__ __device__ __forceinline__ Gas<float> to_f32(const Gas<double>& g)
{
    return Gas<float>{(float)g.gam, (float)g.R, (float)g.inv_prn, (float)g.mu0, (float)g.T0, (float)g.eps, g.K};
}

// Same inputs and result as gks_flux<double>; D (the ten slopes) arrives in FP64.
__host__ __device__ __forceinline__ void gks_flux_mixed(const double WL[5], const double WR[5], const double Wl[5],
                                                        const double Wr[5], const double D[6][5], double tau,
                                                        double dt, double dx, const Gas<double>& g, double F[5])
{
    const double pi = 3.141592653589793238462643383279502884;
    const Gas<float> gf = to_f32(g);
    Maxw<double> L, R;
    maxw_init(L, Wl, g);
    maxw_init(R, Wr, g);
    {
        const double pl = L.den / L.lam, pr = R.den / R.lam;
        tau = tau + dt * (fabs(pl - pr) / fabs(pl + pr));
    }
    const double ex = exp(-dt / tau);
    const double om = 1.0 - ex;
    const double igm0 = 1.0 / (dt - tau * om);
    const double gm1 = -om * igm0;
    const double gm2 = (-dt + 2.0 * tau * om - dt * ex) * igm0;
    const double gm3 = -gm1;
    const double gm4 = (dt * ex - tau * om) * igm0;
    const double gm5 = tau * gm3;
    const double Pc = tau * om;
    const double Qc = -tau * dt * ex + tau * tau * om;
    const double p0 = dt - Pc;
    const double p1 = 0.5 * dt * dt - tau * p0;
    const double p2 = -tau * p0 + Qc;
    const double p3 = Pc;
    const double p4 = -Qc - tau * Pc;
    const double p5 = -tau * Pc;

    const double gl = 0.5 * exp(-L.lam * L.u * L.u) / sqrt(L.lam * pi);
    const double hl0 = 0.5 * erfc(-sqrt(L.lam) * L.u);
    const double gr = 0.5 * exp(-R.lam * R.u * R.u) / sqrt(R.lam * pi);
    const double hr0 = 0.5 * erfc(sqrt(R.lam) * R.u);

    // FP64: equilibrium state, the Euler part p3 (l, r) and the 000 moments
    double We[5], F64[5], Bz[5];
    {
        double hu[4], tv[3], tw[3], M[5];
        half_tab<4>(L.u, L.i2l, hl0, L.u * hl0 + gl, hu);
        full_tab<3>(L.v, L.i2l, tv);
        full_tab<3>(L.w, L.i2l, tw);
        mom<0, 0, 0>(hu, tv, tw, L.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) We[q] = L.den * M[q];
        mom<1, 0, 0>(hu, tv, tw, L.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) F64[q] = p3 * L.den * M[q];
        half_tab<4>(R.u, R.i2l, hr0, R.u * hr0 - gr, hu);
        full_tab<3>(R.v, R.i2l, tv);
        full_tab<3>(R.w, R.i2l, tw);
        mom<0, 0, 0>(hu, tv, tw, R.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) We[q] += R.den * M[q];
        mom<1, 0, 0>(hu, tv, tw, R.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) F64[q] += p3 * R.den * M[q];
    }
    Maxw<double> E;
    maxw_init(E, We, g);
    {
        double fu[4], tv[3], tw[3], M[5];
        full_tab<4>(E.u, E.i2l, fu);
        full_tab<3>(E.v, E.i2l, tv);
        full_tab<3>(E.w, E.i2l, tw);
        mom<1, 0, 0>(fu, tv, tw, E.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) F64[q] += p0 * E.den * M[q];
        // gm1 den_e <psi>_e + gm3 We (GKMflux.cpp:185): zero up to rounding
        mom<0, 0, 0>(fu, tv, tw, E.e2, M);
#pragma unroll
        for (int q = 0; q < 5; q++) Bz[q] = gm1 * E.den * M[q] + gm3 * We[q];
    }

    // FP32: every slope-dependent term, state by state as in gks_flux
    const float fgm2 = (float)gm2, fgm45 = (float)(gm4 + gm5), fgm5 = (float)gm5;
    const float fp1 = (float)p1, fp2 = (float)p2, fp4 = (float)p4, fp5 = (float)p5;
    float Be[5], bye[5] = {0, 0, 0, 0, 0}, bze[5] = {0, 0, 0, 0, 0}, Ff[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 5; q++) Be[q] = (float)Bz[q];
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const Maxw<float> S = to_f32(side == 0 ? L : R);
        const int dxi = side;
        float tv[6], tw[6];
        full_tab<6>(S.v, S.i2l, tv);
        full_tab<6>(S.w, S.i2l, tw);
        const double iden = 1.0 / (side == 0 ? L.den : R.den);
        float b[5], ax[5], ay[5], az[5], A[5];
#pragma unroll
        for (int q = 0; q < 5; q++) b[q] = (float)(D[0 + dxi][q] * iden);
        slope(b, S, gf, ax);
#pragma unroll
        for (int q = 0; q < 5; q++) b[q] = (float)(D[2 + dxi][q] * iden);
        slope(b, S, gf, ay);
#pragma unroll
        for (int q = 0; q < 5; q++) b[q] = (float)(D[4 + dxi][q] * iden);
        slope(b, S, gf, az);
        {
            float fu[6];
            full_tab<6>(S.u, S.i2l, fu);
#pragma unroll
            for (int q = 0; q < 5; q++) b[q] = 0.f;
            mom_a<1, 0, 0>(fu, tv, tw, S.e2, S.e4, ax, -1.f, b);
            mom_a<0, 1, 0>(fu, tv, tw, S.e2, S.e4, ay, -1.f, b);
            mom_a<0, 0, 1>(fu, tv, tw, S.e2, S.e4, az, -1.f, b);
            slope(b, S, gf, A);
        }
        float hu[7];
        if (side == 0)
            half_tab<7>(S.u, S.i2l, (float)hl0, (float)(L.u * hl0 + gl), hu);
        else
            half_tab<7>(S.u, S.i2l, (float)hr0, (float)(R.u * hr0 - gr), hu);
        mom_a<0, 0, 0>(hu, tv, tw, S.e2, S.e4, ay, 1.f, bye);
        mom_a<0, 0, 0>(hu, tv, tw, S.e2, S.e4, az, 1.f, bze);
        const float cB = fgm45 * S.den;
        mom_a<1, 0, 0>(hu, tv, tw, S.e2, S.e4, ax, cB, Be);
        mom_a<0, 1, 0>(hu, tv, tw, S.e2, S.e4, ay, cB, Be);
        mom_a<0, 0, 1>(hu, tv, tw, S.e2, S.e4, az, cB, Be);
        mom_a<0, 0, 0>(hu, tv, tw, S.e2, S.e4, A, fgm5 * S.den, Be);
        const float cF = fp4 * S.den;
        mom_a<2, 0, 0>(hu, tv, tw, S.e2, S.e4, ax, cF, Ff);
        mom_a<1, 1, 0>(hu, tv, tw, S.e2, S.e4, ay, cF, Ff);
        mom_a<1, 0, 1>(hu, tv, tw, S.e2, S.e4, az, cF, Ff);
        mom_a<1, 0, 0>(hu, tv, tw, S.e2, S.e4, A, fp5 * S.den, Ff);
    }
    {
        const Maxw<float> Ef = to_f32(E);
        const double ge = 0.5 * exp(-E.lam * E.u * E.u) / sqrt(E.lam * pi);
        const double sl = sqrt(E.lam) * E.u;
        float tv[6], tw[6];
        full_tab<6>(Ef.v, Ef.i2l, tv);
        full_tab<6>(Ef.w, Ef.i2l, tw);
        const double c = 2.0 / (E.den * dx);
        const float cB = fgm2 * Ef.den;
        const float cF = fp2 * Ef.den;
        float b[5], a[5];
        {
            float h[7];
            const double h0 = 0.5 * erfc(-sl);
            half_tab<7>(Ef.u, Ef.i2l, (float)h0, (float)(E.u * h0 + ge), h);
#pragma unroll
            for (int q = 0; q < 5; q++) b[q] = (float)((We[q] - WL[q]) * c);
            slope(b, Ef, gf, a);
            mom_a<1, 0, 0>(h, tv, tw, Ef.e2, Ef.e4, a, cB, Be);
            mom_a<2, 0, 0>(h, tv, tw, Ef.e2, Ef.e4, a, cF, Ff);
        }
        {
            float h[7];
            const double h0 = 0.5 * erfc(sl);
            half_tab<7>(Ef.u, Ef.i2l, (float)h0, (float)(E.u * h0 - ge), h);
#pragma unroll
            for (int q = 0; q < 5; q++) b[q] = (float)((WR[q] - We[q]) * c);
            slope(b, Ef, gf, a);
            mom_a<1, 0, 0>(h, tv, tw, Ef.e2, Ef.e4, a, cB, Be);
            mom_a<2, 0, 0>(h, tv, tw, Ef.e2, Ef.e4, a, cF, Ff);
        }
        float fu[7];
        full_tab<7>(Ef.u, Ef.i2l, fu);
        slope(bye, Ef, gf, a);
        mom_a<0, 1, 0>(fu, tv, tw, Ef.e2, Ef.e4, a, cB, Be);
        mom_a<1, 1, 0>(fu, tv, tw, Ef.e2, Ef.e4, a, cF, Ff);
        slope(bze, Ef, gf, a);
        mom_a<0, 0, 1>(fu, tv, tw, Ef.e2, Ef.e4, a, cB, Be);
        mom_a<1, 0, 1>(fu, tv, tw, Ef.e2, Ef.e4, a, cF, Ff);
        const float iden = (float)(1.0 / E.den);
#pragma unroll
        for (int q = 0; q < 5; q++) Be[q] *= iden;
        slope(Be, Ef, gf, a);
        mom_a<1, 0, 0>(fu, tv, tw, Ef.e2, Ef.e4, a, fp1 * Ef.den, Ff);
    }
#pragma unroll
    for (int q = 0; q < 5; q++) F[q] = F64[q] + (double)Ff[q];
}

}  // namespace gkm
