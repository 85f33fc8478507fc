// adj_flux.hpp -- reverse-mode derivative of the GKS face flux with respect to
// all 50 of its inputs (no reference counterpart).
//
// face_flux_vjp gives xbar[k] = sum_q Fbar[q] dF_q/dx_k for face_flux (the FP64
// flux gks_flux_tf<double, double> of gkm_flux.hpp, which restates
// GKMflux.cpp:11-297) in one forward and one reverse pass: the forward pass
// keeps the states, the time coefficients, the equilibrium state and the four
// accumulators (den_e B_, the slope flux, the right-hand sides of E's
// tangential slopes); the reverse pass walks the terms backwards, recomputing
// each side block's and each equilibrium half's intermediates where it needs
// them.  Every building block has its vector-Jacobian product here with the
// adjoints of its coefficients as well as of its vector argument: the moment
// tables, the slope solver and the tangential shifts depend on the Maxwellian
// parameters (den, u, v, w, theta = 1/(2 lambda)), and those on the WENO
// states (c2p), the equilibrium state on both through its half-space moments.
// The derivative of |.| at 0 is taken as 0 and the time coefficients are
// differentiated in the collision time with one dual-number evaluation, as in
// the forward-mode derivative (gkm_dual.hpp) this replaces.
#pragma once
#include <hip/hip_runtime.h>

#include "adj_math.hpp"
#include "gkm_dual.hpp"

namespace gkm {
namespace vjp {

// adjoints of a Maxwellian's parameters; those of lambda (lam) and of 2 lambda
// (lam2) are kept apart until the end, where they join theta's (lambda = 1/(2 theta))
struct MaxwBar {
    double den, u, v, w, th, lam, lam2;
};
__host__ __device__ __forceinline__ MaxwBar maxw_bar_zero() { return MaxwBar{0, 0, 0, 0, 0, 0, 0}; }

// theta's total adjoint: lambda = 1/(2 theta), 2 lambda = 1/theta
__host__ __device__ __forceinline__ double theta_bar(const MaxwBar& b, const Maxw<double>& s)
{
    const double l2 = 2.0 * s.lam;
    return b.th - 2.0 * s.lam * s.lam * b.lam - l2 * l2 * b.lam2;
}

// c2p (maxw_init): adjoint of the conservative state from the parameters'
__host__ __device__ __forceinline__ void maxw_vjp(const double W[5], const Maxw<double>& s, double gam,
                                                  const MaxwBar& b, double Wb[5])
{
    // theta = (gam - 1) (W4 / W0 - |W1..3|^2 / (2 W0^2)), u = W1 / W0, ...
    const double ir = g_rcp(W[0]);
    const double t = (gam - 1.0) * theta_bar(b, s);
    Wb[0] += b.den + (t * (s.u * s.u + s.v * s.v + s.w * s.w - W[4] * ir) - (s.u * b.u + s.v * b.v + s.w * b.w)) * ir;
    Wb[1] += (b.u - t * s.u) * ir;
    Wb[2] += (b.v - t * s.v) * ir;
    Wb[3] += (b.w - t * s.w) * ir;
    Wb[4] += t * ir;
}

// m' = tshift(m, dV, dW) with m the forward input: on entry mb = d/dm', on exit
// d/dm; the shifts' adjoints are added
__host__ __device__ __forceinline__ void tshift_vjp(const double m[5], double dV, double dW, double mb[5], double& dVb,
                                                    double& dWb)
{
    dVb += mb[4] * (m[2] + dV * m[0]) + mb[2] * m[0];
    dWb += mb[4] * (m[3] + dW * m[0]) + mb[3] * m[0];
    tshift_T(mb, dV, dW);
}
// m' = tunshift(m, dV, dW), likewise
__host__ __device__ __forceinline__ void tunshift_vjp(const double m[5], double dV, double dW, double mb[5],
                                                      double& dVb, double& dWb)
{
    dVb += mb[4] * (dV * m[0] - m[2]) - mb[2] * m[0];
    dWb += mb[4] * (dW * m[0] - m[3]) - mb[3] * m[0];
    tunshift_T(mb, dV, dW);
}

// a = slope_t(b, s): bb = d/db (assigned), state adjoints added (u, theta, 2 lambda)
__host__ __device__ __forceinline__ void slope_t_vjp(const double b[5], const TState<double>& s, double inv_prn,
                                                     double kp3, double ikp3, const double ab[5], double bb[5],
                                                     double& ub, double& thb, double& l2b)
{
    const double u = s.u, l2 = s.lam2;
    const double R2 = b[1] - u * b[0];
    const double q = u * u + kp3 * s.th;
    const double R5 = 2.0 * b[4] - b[0] * q;
    const double k4 = inv_prn * (l2 * l2) * ikp3;
    const double X = R5 - 2.0 * u * R2;
    const double a4 = k4 * X;
    const double a1 = l2 * R2 - u * a4;
    // a0 = b0 - a1 u - a4 q / 2
    double b0b = ab[0];
    const double a1b = ab[1] - u * ab[0];
    double a4b = ab[4] - 0.5 * q * ab[0];
    double qb = -0.5 * a4 * ab[0];
    double u_b = -a1 * ab[0];
    // a1 = 2 lambda R2 - u a4; a2, a3 = 2 lambda b2, b3
    double l2_b = R2 * a1b + b[2] * ab[2] + b[3] * ab[3];
    double R2b = l2 * a1b;
    u_b -= a4 * a1b;
    a4b -= u * a1b;
    // a4 = k4 X, k4 = inv_prn (2 lambda)^2 / (K + 3)
    l2_b += 2.0 * inv_prn * l2 * ikp3 * X * a4b;
    const double Xb = k4 * a4b;
    // X = R5 - 2 u R2
    u_b -= 2.0 * R2 * Xb;
    R2b -= 2.0 * u * Xb;
    // R5 = 2 b4 - b0 q
    bb[4] = 2.0 * Xb;
    b0b -= q * Xb;
    qb -= b[0] * Xb;
    // q = u^2 + (K + 3) theta
    u_b += 2.0 * u * qb;
    thb += kp3 * qb;
    // R2 = b1 - u b0
    bb[1] = R2b;
    b0b -= u * R2b;
    u_b -= b[0] * R2b;
    bb[0] = b0b;
    bb[2] = l2 * ab[2];
    bb[3] = l2 * ab[3];
    ub += u_b;
    l2b += l2_b;
}

// acc += mom_u(H, th, a): ab, Hb[0..4], thb added
__host__ __device__ __forceinline__ void mom_u_vjp(const double* H, double th, double kp2, double kp4,
                                                   const double a[5], const double accb[5], double ab[5], double* Hb,
                                                   double& thb)
{
    mom_u_T(H, th, kp2, kp4, accb, ab);
    const double kth = kp2 * th;
    const double s1b = a[4] * accb[0] + a[0] * accb[4];
    const double s2b = a[4] * accb[1] + a[1] * accb[4];
    const double s4b = a[4] * accb[4];
    const double t23 = a[2] * accb[2] + a[3] * accb[3];
    Hb[0] += a[0] * accb[0] + th * t23 + 0.5 * kth * s1b + 0.25 * kth * kp4 * th * s4b;
    Hb[1] += a[1] * accb[0] + a[0] * accb[1] + 0.5 * kth * s2b;
    Hb[2] += a[1] * accb[1] + 0.5 * s1b + 0.5 * kth * s4b;
    Hb[3] += 0.5 * s2b;
    Hb[4] += 0.25 * s4b;
    thb += H[0] * t23 + 0.5 * kp2 * (H[0] * s1b + H[1] * s2b + (H[2] + kp4 * th * H[0]) * s4b);
}
// acc += mom_uvw(H, th, a, c): ab, Hb[0..2], thb added
__host__ __device__ __forceinline__ void mom_uvw_vjp(const double* H, double th, double kp4, const double a[5],
                                                     const double accb[5], double ab[5], double* Hb, double& thb,
                                                     int c)
{
    mom_uvw_T(H, th, kp4, accb, ab, c);
    const double ac = a[c];
    const double tH0b = ac * accb[0] + a[0] * accb[c];
    const double tH1b = ac * accb[1] + a[1] * accb[c];
    const double tb = a[4] * accb[c] + ac * accb[4];  // t = th H2 / 2 + kp4 th^2 H0 / 2
    Hb[0] += th * tH0b + 0.5 * kp4 * th * th * tb;
    Hb[1] += th * tH1b;
    Hb[2] += 0.5 * th * tb;
    thb += H[0] * tH0b + H[1] * tH1b + (0.5 * H[2] + kp4 * th * H[0]) * tb;
}

// f = full_tab / half_tab (f_n = m f_(n-1) + (n-1) th f_(n-2)): fb is taken back
// to entries 0 and 1; m's and theta's adjoints added
template <int N>
__host__ __device__ __forceinline__ void tab_vjp(double m, double th, const double* f, double* fb, double& mb,
                                                 double& thb)
{
#pragma unroll
    for (int n = N - 1; n >= 2; n--) {
        mb += f[n - 1] * fb[n];
        thb += double(n - 1) * f[n - 2] * fb[n];
        fb[n - 1] += m * fb[n];
        fb[n - 2] += double(n - 1) * th * fb[n];
    }
}

// first two half-space moments, h0 = erfc(-sgn sqrt(lam) u)/2 and
// h1 = u h0 + sgn exp(-lam u^2)/(2 sqrt(pi lam)) (sgn = +1: u > 0, -1: u < 0):
// adjoints of u and lambda added from h0's and h1's
__host__ __device__ __forceinline__ void half_start_vjp(double u, double lam, double sgn, double h0, double h0b,
                                                        double h1b, double& ub, double& lamb)
{
    const double pi = 3.141592653589793238462643383279502884;
    const double isqpi = 0.5641895835477562869480794515607726;  // 1/sqrt(pi)
    const double sl = g_sqrt(lam), e = g_exp(-lam * u * u);
    const double gsg = g_div(sgn * 0.5 * e, g_sqrt(lam * pi));  // the sgn-signed exponential term of h1
    const double h0t = h0b + u * h1b;                     // h0's total adjoint
    const double zb = h0t * sgn * e * isqpi;              // of sqrt(lam) u
    const double gb = gsg * h1b;
    ub += h0 * h1b + sl * zb - 2.0 * lam * u * gb;
    lamb += g_div(0.5 * u * zb, sl) - (u * u + 0.5 * g_rcp(lam)) * gb;
}

__host__ __device__ __forceinline__ double dot5(const double a[5], const double b[5])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4];
}

// time-integration coefficients of the flux in the collision time (the forms of
// gks_flux_tf, GKMflux.cpp:162-167, 197-204)
template <typename T>
struct TimeCoefs {
    T gm1, gm2, gm3, gm4, gm5, p0, p1, p2, p3, p4, p5;
};
template <typename T>
__host__ __device__ __forceinline__ TimeCoefs<T> time_coefs(T tau, T dt)
{
    TimeCoefs<T> c;
    const T ex = g_exp(-g_div(dt, tau));
    const T om = T(1) - ex;
    const T igm0 = g_rcp(dt - tau * om);
    c.gm1 = -om * igm0;
    c.gm2 = (-dt + T(2) * tau * om - dt * ex) * igm0;
    c.gm3 = -c.gm1;
    c.gm4 = (dt * ex - tau * om) * igm0;
    c.gm5 = tau * c.gm3;
    const T Pc = tau * om;
    const T Qc = -tau * dt * ex + tau * tau * om;
    c.p0 = dt - Pc;
    c.p1 = T(0.5) * dt * dt - tau * c.p0;
    c.p2 = -tau * c.p0 + Qc;
    c.p3 = Pc;
    c.p4 = -Qc - tau * Pc;
    c.p5 = -tau * Pc;
    return c;
}

// constants of one flux evaluation
struct FluxK {
    double kp2, kp3, kp4, ikp3, inv_prn;
};

// One side's slope-dependent terms (gks_flux_tf's side block, FP64 form), in
// the side's tangential rest frame before the shift into E's: sy, sz (E's
// tangential right-hand sides), b (den_e B_), f (slope flux); and the
// intermediates the reverse pass reads.
struct SideFwd {
    TState<double> st;
    double iden, dV, dW;
    double m[3][5];  // tunshifted lab slopes D iden (x, y, z)
    double ax[5], ay[5], az[5], tvw[5], nb[5], A[5];
    double fu[6], hu[7];
    double sy[5], sz[5], b[5], f[5];
};

template <int side>
__host__ __device__ __forceinline__ void side_fwd(const Maxw<double>& S, double h0, double h1, const double* x,
                                                  const Maxw<double>& E, const TimeCoefs<double>& tc, const FluxK& k,
                                                  SideFwd& s)
{
    s.st = TState<double>{S.u, S.i2l, 2.0 * S.lam};
    s.iden = g_rcp(S.den);
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int q = 0; q < 5; q++) s.m[d][q] = x[20 + 5 * (2 * d + side) + q] * s.iden;
        tunshift(s.m[d], S.v, S.w);
        slope_t(s.m[d], s.st, k.inv_prn, k.kp3, k.ikp3, d == 0 ? s.ax : (d == 1 ? s.ay : s.az));
    }
    for (int q = 0; q < 5; q++) s.tvw[q] = S.v * s.ay[q] + S.w * s.az[q];
    full_tab<6>(s.st.u, s.st.th, s.fu);
    double bt[5] = {0, 0, 0, 0, 0};
    mom_u(s.fu + 1, s.st.th, k.kp2, k.kp4, s.ax, bt);
    mom_uvw(s.fu, s.st.th, k.kp4, s.ay, bt, 2);
    mom_uvw(s.fu, s.st.th, k.kp4, s.az, bt, 3);
    mom_u(s.fu, s.st.th, k.kp2, k.kp4, s.tvw, bt);
    for (int q = 0; q < 5; q++) s.nb[q] = -bt[q];
    slope_t(s.nb, s.st, k.inv_prn, k.kp3, k.ikp3, s.A);
    half_tab<7>(s.st.u, s.st.th, h0, h1, s.hu);
    for (int q = 0; q < 5; q++) s.sy[q] = s.sz[q] = s.b[q] = s.f[q] = 0.0;
    mom_u(s.hu, s.st.th, k.kp2, k.kp4, s.ay, s.sy);
    mom_u(s.hu, s.st.th, k.kp2, k.kp4, s.az, s.sz);
    const double cB = (tc.gm4 + tc.gm5) * S.den, cF = tc.p4 * S.den, cA = tc.gm5 * S.den, cA2 = tc.p5 * S.den;
    double c[5];
    for (int q = 0; q < 5; q++) c[q] = cB * s.ax[q];
    mom_u(s.hu + 1, s.st.th, k.kp2, k.kp4, c, s.b);
    for (int q = 0; q < 5; q++) c[q] = cF * s.ax[q];
    mom_u(s.hu + 2, s.st.th, k.kp2, k.kp4, c, s.f);
    for (int q = 0; q < 5; q++) c[q] = cB * s.ay[q];
    mom_uvw(s.hu, s.st.th, k.kp4, c, s.b, 2);
    for (int q = 0; q < 5; q++) c[q] = cF * s.ay[q];
    mom_uvw(s.hu + 1, s.st.th, k.kp4, c, s.f, 2);
    for (int q = 0; q < 5; q++) c[q] = cB * s.az[q];
    mom_uvw(s.hu, s.st.th, k.kp4, c, s.b, 3);
    for (int q = 0; q < 5; q++) c[q] = cF * s.az[q];
    mom_uvw(s.hu + 1, s.st.th, k.kp4, c, s.f, 3);
    for (int q = 0; q < 5; q++) c[q] = cB * s.tvw[q] + cA * s.A[q];
    mom_u(s.hu, s.st.th, k.kp2, k.kp4, c, s.b);
    for (int q = 0; q < 5; q++) c[q] = cF * s.tvw[q] + cA2 * s.A[q];
    mom_u(s.hu + 1, s.st.th, k.kp2, k.kp4, c, s.f);
    s.dV = S.v - E.v;
    s.dW = S.w - E.w;
}

// adjoints of the four coefficients' time coefficients, in collision-time order
struct CoefBar {
    double gm1, gm2, gm3, gm4, gm5, p0, p1, p2, p3, p4, p5;
};

// Reverse of one side block: from the adjoints of E's-frame accumulators
// (Beb, Fsb, byeb, bzeb) to the side's parameters (Sb), E's tangential
// velocity, the side's two half-space moments h0, h1, the six-slope inputs of
// this side (xbar[20 + 5 (2 d + side) + q]) and the time coefficients.
template <int side>
__host__ __device__ __noinline__ void side_rev(const Maxw<double>& S, double h0, double h1, const double* x,
                                         const Maxw<double>& E, const TimeCoefs<double>& tc, const FluxK& k,
                                         const double Beb[5], const double Fsb[5], const double byeb[5],
                                         const double bzeb[5], MaxwBar& Sb, MaxwBar& Eb, double& h0b, double& h1b,
                                         CoefBar& cb, double* xbar)
{
    SideFwd s;
    side_fwd<side>(S, h0, h1, x, E, tc, k, s);
    const double th = s.st.th;
    double dVb = 0.0, dWb = 0.0;
    double syb[5], szb[5], bb[5], fb[5];
    for (int q = 0; q < 5; q++) {
        syb[q] = byeb[q];
        szb[q] = bzeb[q];
        bb[q] = Beb[q];
        fb[q] = Fsb[q];
    }
    // the shift of each pre-shift sum into E's rest frame (tshift is applied to the sums)
    tshift_vjp(s.sy, s.dV, s.dW, syb, dVb, dWb);
    tshift_vjp(s.sz, s.dV, s.dW, szb, dVb, dWb);
    tshift_vjp(s.b, s.dV, s.dW, bb, dVb, dWb);
    tshift_vjp(s.f, s.dV, s.dW, fb, dVb, dWb);
    Sb.v += dVb;
    Sb.w += dWb;
    Eb.v -= dVb;
    Eb.w -= dWb;
    double axb[5] = {0, 0, 0, 0, 0}, ayb[5] = {0, 0, 0, 0, 0}, azb[5] = {0, 0, 0, 0, 0};
    double tvwb[5] = {0, 0, 0, 0, 0}, Ab[5] = {0, 0, 0, 0, 0};
    double hub[7] = {0, 0, 0, 0, 0, 0, 0}, fub[6] = {0, 0, 0, 0, 0, 0};
    double thb = 0.0;
    const double cB = (tc.gm4 + tc.gm5) * S.den, cF = tc.p4 * S.den, cA = tc.gm5 * S.den, cA2 = tc.p5 * S.den;
    double cBb = 0.0, cFb = 0.0, cAb = 0.0, cA2b = 0.0;
    double c[5], t[5];
    auto zero = [&](double* v) {
        for (int q = 0; q < 5; q++) v[q] = 0.0;
    };
    // f = mom_u(hu+2, cF ax) + mom_uvw(hu+1, cF ay, 2) + mom_uvw(hu+1, cF az, 3) + mom_u(hu+1, cF tvw + cA2 A)
    for (int q = 0; q < 5; q++) c[q] = cF * s.tvw[q] + cA2 * s.A[q];
    zero(t);
    mom_u_vjp(s.hu + 1, th, k.kp2, k.kp4, c, fb, t, hub + 1, thb);
    for (int q = 0; q < 5; q++) {
        tvwb[q] += cF * t[q];
        Ab[q] += cA2 * t[q];
    }
    cFb += dot5(s.tvw, t);
    cA2b += dot5(s.A, t);
    for (int q = 0; q < 5; q++) c[q] = cF * s.az[q];
    zero(t);
    mom_uvw_vjp(s.hu + 1, th, k.kp4, c, fb, t, hub + 1, thb, 3);
    for (int q = 0; q < 5; q++) azb[q] += cF * t[q];
    cFb += dot5(s.az, t);
    for (int q = 0; q < 5; q++) c[q] = cF * s.ay[q];
    zero(t);
    mom_uvw_vjp(s.hu + 1, th, k.kp4, c, fb, t, hub + 1, thb, 2);
    for (int q = 0; q < 5; q++) ayb[q] += cF * t[q];
    cFb += dot5(s.ay, t);
    for (int q = 0; q < 5; q++) c[q] = cF * s.ax[q];
    zero(t);
    mom_u_vjp(s.hu + 2, th, k.kp2, k.kp4, c, fb, t, hub + 2, thb);
    for (int q = 0; q < 5; q++) axb[q] += cF * t[q];
    cFb += dot5(s.ax, t);
    // b = mom_u(hu+1, cB ax) + mom_uvw(hu, cB ay, 2) + mom_uvw(hu, cB az, 3) + mom_u(hu, cB tvw + cA A)
    for (int q = 0; q < 5; q++) c[q] = cB * s.tvw[q] + cA * s.A[q];
    zero(t);
    mom_u_vjp(s.hu, th, k.kp2, k.kp4, c, bb, t, hub, thb);
    for (int q = 0; q < 5; q++) {
        tvwb[q] += cB * t[q];
        Ab[q] += cA * t[q];
    }
    cBb += dot5(s.tvw, t);
    cAb += dot5(s.A, t);
    for (int q = 0; q < 5; q++) c[q] = cB * s.az[q];
    zero(t);
    mom_uvw_vjp(s.hu, th, k.kp4, c, bb, t, hub, thb, 3);
    for (int q = 0; q < 5; q++) azb[q] += cB * t[q];
    cBb += dot5(s.az, t);
    for (int q = 0; q < 5; q++) c[q] = cB * s.ay[q];
    zero(t);
    mom_uvw_vjp(s.hu, th, k.kp4, c, bb, t, hub, thb, 2);
    for (int q = 0; q < 5; q++) ayb[q] += cB * t[q];
    cBb += dot5(s.ay, t);
    for (int q = 0; q < 5; q++) c[q] = cB * s.ax[q];
    zero(t);
    mom_u_vjp(s.hu + 1, th, k.kp2, k.kp4, c, bb, t, hub + 1, thb);
    for (int q = 0; q < 5; q++) axb[q] += cB * t[q];
    cBb += dot5(s.ax, t);
    // sy = mom_u(hu, ay), sz = mom_u(hu, az)
    mom_u_vjp(s.hu, th, k.kp2, k.kp4, s.az, szb, azb, hub, thb);
    mom_u_vjp(s.hu, th, k.kp2, k.kp4, s.ay, syb, ayb, hub, thb);
    // the coefficients
    cb.gm4 += S.den * cBb;
    cb.gm5 += S.den * (cBb + cAb);
    cb.p4 += S.den * cFb;
    cb.p5 += S.den * cA2b;
    Sb.den += (tc.gm4 + tc.gm5) * cBb + tc.p4 * cFb + tc.gm5 * cAb + tc.p5 * cA2b;
    // A = slope_t(-bt)
    double nbb[5];
    slope_t_vjp(s.nb, s.st, k.inv_prn, k.kp3, k.ikp3, Ab, nbb, Sb.u, thb, Sb.lam2);
    for (int q = 0; q < 5; q++) nbb[q] = -nbb[q];
    // bt = mom_u(fu+1, ax) + mom_uvw(fu, ay, 2) + mom_uvw(fu, az, 3) + mom_u(fu, tvw)
    mom_u_vjp(s.fu, th, k.kp2, k.kp4, s.tvw, nbb, tvwb, fub, thb);
    mom_uvw_vjp(s.fu, th, k.kp4, s.az, nbb, azb, fub, thb, 3);
    mom_uvw_vjp(s.fu, th, k.kp4, s.ay, nbb, ayb, fub, thb, 2);
    mom_u_vjp(s.fu + 1, th, k.kp2, k.kp4, s.ax, nbb, axb, fub + 1, thb);
    // tvw = V ay + W az
    for (int q = 0; q < 5; q++) {
        ayb[q] += S.v * tvwb[q];
        azb[q] += S.w * tvwb[q];
    }
    Sb.v += dot5(s.ay, tvwb);
    Sb.w += dot5(s.az, tvwb);
    // micro slopes a_d = slope_t(tunshift(D_d iden)), lab-frame slope inputs D_d
    double idenb = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double mb[5];
        slope_t_vjp(s.m[d], s.st, k.inv_prn, k.kp3, k.ikp3, d == 0 ? axb : (d == 1 ? ayb : azb), mb, Sb.u, thb,
                    Sb.lam2);
        double Dl[5];
        for (int q = 0; q < 5; q++) Dl[q] = x[20 + 5 * (2 * d + side) + q] * s.iden;
        tunshift_vjp(Dl, S.v, S.w, mb, Sb.v, Sb.w);
        for (int q = 0; q < 5; q++) {
            xbar[20 + 5 * (2 * d + side) + q] = s.iden * mb[q];
            idenb += x[20 + 5 * (2 * d + side) + q] * mb[q];
        }
    }
    Sb.den -= idenb * s.iden * s.iden;
    // the moment tables
    tab_vjp<6>(s.st.u, th, s.fu, fub, Sb.u, thb);
    Sb.u += fub[1];
    tab_vjp<7>(s.st.u, th, s.hu, hub, Sb.u, thb);
    h0b += hub[0];
    h1b += hub[1];
    Sb.th += thb;
}

}  // namespace vjp

// xbar[k] = sum_q Fbar[q] dF_q/dx_k for face_flux (gkm_flux.hpp's FP64 flux with
// the collision time of the left cell), x = (WL, WR, Wl, Wr, D[6]) in the
// rotated frame; xbar is assigned.
__host__ __device__ inline void face_flux_vjp(const double x[50], double dt, double dx, const Gas<double>& g,
                                              const double Fbar[5], double xbar[50])
{
    using namespace vjp;
    const FluxK k{double(g.K + 2), double(g.K + 3), double(g.K + 4), 1.0 / double(g.K + 3), g.inv_prn};
    const double* WL = x;
    const double* WR = x + 5;
    const double* Wl = x + 10;
    const double* Wr = x + 15;
    for (int i = 0; i < 50; i++) xbar[i] = 0.0;

    // ---------------- forward
    const double tau0 = collision_time<double>(WL, g);
    Maxw<double> L, R;
    maxw_init(L, Wl, g);
    maxw_init(R, Wr, g);
    const double pl = 2.0 * L.den * L.i2l, pr = 2.0 * R.den * R.i2l;  // rho / lambda, as the flux forms it
    const double taue = tau0 + dt * g_div(::fabs(pl - pr), ::fabs(pl + pr));
    const TimeCoefs<Dual> tcd = time_coefs<Dual>(Dual(taue, 1.0), Dual(dt));
    const TimeCoefs<double> tc{tcd.gm1.v, tcd.gm2.v, tcd.gm3.v, tcd.gm4.v, tcd.gm5.v, tcd.p0.v,
                               tcd.p1.v,  tcd.p2.v,  tcd.p3.v,  tcd.p4.v,  tcd.p5.v};
    // the half-space starts as gks_flux_tf forms them
    const double isqpi = 0.5641895835477562869480794515607726;  // 1/sqrt(pi)
    double hl[4], hr[4];
    {
        const double sl = g_sqrt(L.lam), zl = sl * L.u, el = g_exp(-zl * zl);
        const double gl = g_div(0.5 * isqpi * el, sl);
        const double h0 = half_erfc(-zl, el);
        half_tab<4>(L.u, L.i2l, h0, L.u * h0 + gl, hl);
        const double sr = g_sqrt(R.lam), zr = sr * R.u, er = g_exp(-zr * zr);
        const double gr = g_div(0.5 * isqpi * er, sr);
        const double r0 = half_erfc(zr, er);
        half_tab<4>(R.u, R.i2l, r0, R.u * r0 - gr, hr);
    }
    // the half Maxwellians' <psi> (equilibrium) and <u psi> (p3 flux) before the
    // tangential shift, and after it
    auto half_moms = [&](const Maxw<double>& S, const double* h, double m[5], double n[5]) {
        m[0] = h[0];
        m[1] = h[1];
        m[2] = m[3] = 0.0;
        m[4] = 0.5 * (h[2] + k.kp2 * S.i2l * h[0]);
        n[0] = h[1];
        n[1] = h[2];
        n[2] = n[3] = 0.0;
        n[4] = 0.5 * (h[3] + k.kp2 * S.i2l * h[1]);
    };
    double We[5];
    {
        double m[5], n[5];
        half_moms(L, hl, m, n);
        tshift(m, L.v, L.w);
        for (int q = 0; q < 5; q++) We[q] = L.den * m[q];
        half_moms(R, hr, m, n);
        tshift(m, R.v, R.w);
        for (int q = 0; q < 5; q++) We[q] += R.den * m[q];
    }
    Maxw<double> E;
    maxw_init(E, We, g);
    const TState<double> sE{E.u, E.i2l, 2.0 * E.lam};
    const double f2 = E.u * E.u + E.i2l, f3 = E.u * f2 + 2.0 * E.i2l * E.u;
    // den_e B_'s start: gm1 den_e <psi>_e + gm3 We, in E's rest frame
    double m1[5] = {1.0, E.u, 0.0, 0.0, 0.5 * (f2 + k.kp2 * E.i2l)}, m1s[5], Bz[5], Be[5];
    for (int q = 0; q < 5; q++) m1s[q] = m1[q];
    tshift(m1s, E.v, E.w);
    for (int q = 0; q < 5; q++) Bz[q] = tc.gm1 * E.den * m1s[q] + tc.gm3 * We[q];
    for (int q = 0; q < 5; q++) Be[q] = Bz[q];
    tunshift(Be, E.v, E.w);
    double Fs[5] = {0, 0, 0, 0, 0}, bye[5] = {0, 0, 0, 0, 0}, bze[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int side = 0; side < 2; side++) {
        SideFwd s;
        if (side == 0) side_fwd<0>(L, hl[0], hl[1], x, E, tc, k, s);
        else side_fwd<1>(R, hr[0], hr[1], x, E, tc, k, s);
        tshift(s.sy, s.dV, s.dW);
        tshift(s.sz, s.dV, s.dW);
        tshift(s.b, s.dV, s.dW);
        tshift(s.f, s.dV, s.dW);
        for (int q = 0; q < 5; q++) {
            bye[q] += s.sy[q];
            bze[q] += s.sz[q];
            Be[q] += s.b[q];
            Fs[q] += s.f[q];
        }
    }
    // equilibrium halves with the cells' normal slopes
    const double se = g_sqrt(E.lam), ze = se * E.u, ee = g_exp(-ze * ze);
    const double ge = g_div(0.5 * isqpi * ee, se);
    const double hp0 = half_erfc(-ze, ee);
    const double hm0 = 1.0 - hp0;
    const double cn = 2.0 * g_rcp(E.den * dx);
    const double cBE = tc.gm2 * E.den, cFE = tc.p2 * E.den;
    auto normal_fwd = [&](int half, double hh[7], double dn[5], double m[5], double a[5]) {
        if (half == 0) half_tab<7>(sE.u, sE.th, hp0, E.u * hp0 + ge, hh);
        else half_tab<7>(sE.u, sE.th, hm0, E.u * hm0 - ge, hh);
        for (int q = 0; q < 5; q++) {
            dn[q] = half == 0 ? We[q] - WL[q] : WR[q] - We[q];
            m[q] = dn[q] * cn;
        }
        tunshift(m, E.v, E.w);
        slope_t(m, sE, k.inv_prn, k.kp3, k.ikp3, a);
    };
#pragma unroll
    for (int half = 0; half < 2; half++) {
        double hh[7], dn[5], m[5], a[5], t[5];
        normal_fwd(half, hh, dn, m, a);
        for (int q = 0; q < 5; q++) t[q] = cBE * a[q];
        mom_u(hh + 1, sE.th, k.kp2, k.kp4, t, Be);
        for (int q = 0; q < 5; q++) t[q] = cFE * a[q];
        mom_u(hh + 2, sE.th, k.kp2, k.kp4, t, Fs);
    }
    // E's tangential and time-slope terms
    double fuE[6], ayE[5], azE[5], tvwE[5], aE[5], bE[5];
    full_tab<6>(sE.u, sE.th, fuE);
    slope_t(bye, sE, k.inv_prn, k.kp3, k.ikp3, ayE);
    slope_t(bze, sE, k.inv_prn, k.kp3, k.ikp3, azE);
    for (int q = 0; q < 5; q++) tvwE[q] = E.v * ayE[q] + E.w * azE[q];
    {
        double t[5];
        for (int q = 0; q < 5; q++) t[q] = cBE * ayE[q];
        mom_uvw(fuE, sE.th, k.kp4, t, Be, 2);
        for (int q = 0; q < 5; q++) t[q] = cBE * azE[q];
        mom_uvw(fuE, sE.th, k.kp4, t, Be, 3);
        for (int q = 0; q < 5; q++) t[q] = cBE * tvwE[q];
        mom_u(fuE, sE.th, k.kp2, k.kp4, t, Be);
        for (int q = 0; q < 5; q++) t[q] = cFE * ayE[q];
        mom_uvw(fuE + 1, sE.th, k.kp4, t, Fs, 2);
        for (int q = 0; q < 5; q++) t[q] = cFE * azE[q];
        mom_uvw(fuE + 1, sE.th, k.kp4, t, Fs, 3);
    }
    const double idenE = g_rcp(E.den);
    for (int q = 0; q < 5; q++) bE[q] = Be[q] * idenE;
    slope_t(bE, sE, k.inv_prn, k.kp3, k.ikp3, aE);
    const double c1 = tc.p1 * E.den;
    double tF[5];
    for (int q = 0; q < 5; q++) tF[q] = cFE * tvwE[q] + c1 * aE[q];
    mom_u(fuE + 1, sE.th, k.kp2, k.kp4, tF, Fs);

    // ---------------- reverse
    MaxwBar Lb = maxw_bar_zero(), Rb = maxw_bar_zero(), Eb = maxw_bar_zero();
    CoefBar cb{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double Web[5] = {0, 0, 0, 0, 0};
    double thEb = 0.0;
    // F = F64 + tshift(Fs, E.v, E.w)
    double Fsb[5];
    for (int q = 0; q < 5; q++) Fsb[q] = Fbar[q];
    tshift_vjp(Fs, E.v, E.w, Fsb, Eb.v, Eb.w);
    // Fs += mom_u(fuE + 1, cF tvw + c1 a), a = slope_t(Be / den)
    double fuEb[6] = {0, 0, 0, 0, 0, 0}, tvwEb[5], ayEb[5] = {0, 0, 0, 0, 0}, azEb[5] = {0, 0, 0, 0, 0};
    double cBEb = 0.0, cFEb = 0.0, Beb[5];
    {
        double tb[5] = {0, 0, 0, 0, 0}, aEb[5], bEb[5];
        mom_u_vjp(fuE + 1, sE.th, k.kp2, k.kp4, tF, Fsb, tb, fuEb + 1, thEb);
        for (int q = 0; q < 5; q++) {
            tvwEb[q] = cFE * tb[q];
            aEb[q] = c1 * tb[q];
        }
        cFEb += dot5(tvwE, tb);
        const double c1b = dot5(aE, tb);
        cb.p1 += E.den * c1b;
        Eb.den += tc.p1 * c1b;
        slope_t_vjp(bE, sE, k.inv_prn, k.kp3, k.ikp3, aEb, bEb, Eb.u, thEb, Eb.lam2);
        for (int q = 0; q < 5; q++) Beb[q] = bEb[q] * idenE;
        Eb.den -= dot5(Be, bEb) * idenE * idenE;
    }
    {
        double t[5], tb[5];
        auto zero = [&](double* v) {
            for (int q = 0; q < 5; q++) v[q] = 0.0;
        };
        for (int q = 0; q < 5; q++) t[q] = cFE * azE[q];
        zero(tb);
        mom_uvw_vjp(fuE + 1, sE.th, k.kp4, t, Fsb, tb, fuEb + 1, thEb, 3);
        for (int q = 0; q < 5; q++) azEb[q] += cFE * tb[q];
        cFEb += dot5(azE, tb);
        for (int q = 0; q < 5; q++) t[q] = cFE * ayE[q];
        zero(tb);
        mom_uvw_vjp(fuE + 1, sE.th, k.kp4, t, Fsb, tb, fuEb + 1, thEb, 2);
        for (int q = 0; q < 5; q++) ayEb[q] += cFE * tb[q];
        cFEb += dot5(ayE, tb);
        for (int q = 0; q < 5; q++) t[q] = cBE * tvwE[q];
        zero(tb);
        mom_u_vjp(fuE, sE.th, k.kp2, k.kp4, t, Beb, tb, fuEb, thEb);
        for (int q = 0; q < 5; q++) tvwEb[q] += cBE * tb[q];
        cBEb += dot5(tvwE, tb);
        for (int q = 0; q < 5; q++) t[q] = cBE * azE[q];
        zero(tb);
        mom_uvw_vjp(fuE, sE.th, k.kp4, t, Beb, tb, fuEb, thEb, 3);
        for (int q = 0; q < 5; q++) azEb[q] += cBE * tb[q];
        cBEb += dot5(azE, tb);
        for (int q = 0; q < 5; q++) t[q] = cBE * ayE[q];
        zero(tb);
        mom_uvw_vjp(fuE, sE.th, k.kp4, t, Beb, tb, fuEb, thEb, 2);
        for (int q = 0; q < 5; q++) ayEb[q] += cBE * tb[q];
        cBEb += dot5(ayE, tb);
    }
    for (int q = 0; q < 5; q++) {  // tvw = V ay + W az
        ayEb[q] += E.v * tvwEb[q];
        azEb[q] += E.w * tvwEb[q];
    }
    Eb.v += dot5(ayE, tvwEb);
    Eb.w += dot5(azE, tvwEb);
    double byeb[5], bzeb[5];
    slope_t_vjp(bye, sE, k.inv_prn, k.kp3, k.ikp3, ayEb, byeb, Eb.u, thEb, Eb.lam2);
    slope_t_vjp(bze, sE, k.inv_prn, k.kp3, k.ikp3, azEb, bzeb, Eb.u, thEb, Eb.lam2);
    tab_vjp<6>(sE.u, sE.th, fuE, fuEb, Eb.u, thEb);
    Eb.u += fuEb[1];
    // the equilibrium halves with the normal slopes
    {
        double geb_u = 0.0, geb_l = 0.0, cnb = 0.0;
#pragma unroll
        for (int half = 0; half < 2; half++) {
            double hh[7], dn[5], m[5], a[5], t[5], tb[5] = {0, 0, 0, 0, 0}, ab[5], hhb[7] = {0, 0, 0, 0, 0, 0, 0};
            normal_fwd(half, hh, dn, m, a);
            for (int q = 0; q < 5; q++) t[q] = cFE * a[q];
            mom_u_vjp(hh + 2, sE.th, k.kp2, k.kp4, t, Fsb, tb, hhb + 2, thEb);
            for (int q = 0; q < 5; q++) ab[q] = cFE * tb[q];
            cFEb += dot5(a, tb);
            for (int q = 0; q < 5; q++) {
                t[q] = cBE * a[q];
                tb[q] = 0.0;
            }
            mom_u_vjp(hh + 1, sE.th, k.kp2, k.kp4, t, Beb, tb, hhb + 1, thEb);
            for (int q = 0; q < 5; q++) ab[q] += cBE * tb[q];
            cBEb += dot5(a, tb);
            double mb[5], mc[5];
            // m was tunshifted in place from dn cn: recover the shift's input
            for (int q = 0; q < 5; q++) mc[q] = dn[q] * cn;
            slope_t_vjp(m, sE, k.inv_prn, k.kp3, k.ikp3, ab, mb, Eb.u, thEb, Eb.lam2);
            tunshift_vjp(mc, E.v, E.w, mb, Eb.v, Eb.w);
            cnb += dot5(dn, mb);
            for (int q = 0; q < 5; q++) {
                if (half == 0) {
                    Web[q] += cn * mb[q];
                    xbar[q] -= cn * mb[q];
                } else {
                    xbar[5 + q] += cn * mb[q];
                    Web[q] -= cn * mb[q];
                }
            }
            tab_vjp<7>(sE.u, sE.th, hh, hhb, Eb.u, thEb);
            // hh0 = hp0 or 1 - hp0; hh1 = E.u hh0 +- ge
            const double sgn = half == 0 ? 1.0 : -1.0;
            half_start_vjp(E.u, E.lam, sgn, hh[0], hhb[0], hhb[1], geb_u, geb_l);
        }
        Eb.u += geb_u;
        Eb.lam += geb_l;
        Eb.den -= cnb * cn * idenE;  // cn = 2/(den dx)
    }
    cb.gm2 += E.den * cBEb;
    cb.p2 += E.den * cFEb;
    Eb.den += tc.gm2 * cBEb + tc.p2 * cFEb;
    // the side blocks
    double hlb[4] = {0, 0, 0, 0}, hrb[4] = {0, 0, 0, 0};
    side_rev<0>(L, hl[0], hl[1], x, E, tc, k, Beb, Fsb, byeb, bzeb, Lb, Eb, hlb[0], hlb[1], cb, xbar);
    side_rev<1>(R, hr[0], hr[1], x, E, tc, k, Beb, Fsb, byeb, bzeb, Rb, Eb, hrb[0], hrb[1], cb, xbar);
    // Be's start: tunshift(Bz), Bz = gm1 den_e tshift(m1) + gm3 We
    {
        double Bzb[5];
        for (int q = 0; q < 5; q++) Bzb[q] = Beb[q];
        tunshift_vjp(Bz, E.v, E.w, Bzb, Eb.v, Eb.w);
        const double c1z = tc.gm1 * E.den;
        double m1b[5];
        for (int q = 0; q < 5; q++) {
            m1b[q] = c1z * Bzb[q];
            Web[q] += tc.gm3 * Bzb[q];
        }
        const double c1zb = dot5(m1s, Bzb);
        cb.gm3 += dot5(We, Bzb);
        cb.gm1 += E.den * c1zb;
        Eb.den += tc.gm1 * c1zb;
        tshift_vjp(m1, E.v, E.w, m1b, Eb.v, Eb.w);
        // m1 = (1, U, 0, 0, (f2 + (K+2) th)/2); F64's E part adds to f2b, f3b below
        Eb.u += m1b[1];
        double f2b = 0.5 * m1b[4], f3b = 0.0;
        thEb += 0.5 * k.kp2 * m1b[4];
        // F64 += p0 den_e tshift(m0), m0 = (U, f2, 0, 0, (f3 + (K+2) th U)/2)
        double m0[5] = {E.u, f2, 0.0, 0.0, 0.5 * (f3 + k.kp2 * E.i2l * E.u)}, m0s[5], m0b[5];
        for (int q = 0; q < 5; q++) m0s[q] = m0[q];
        tshift(m0s, E.v, E.w);
        const double ce = tc.p0 * E.den;
        for (int q = 0; q < 5; q++) m0b[q] = ce * Fbar[q];
        const double ceb = dot5(m0s, Fbar);
        cb.p0 += E.den * ceb;
        Eb.den += tc.p0 * ceb;
        tshift_vjp(m0, E.v, E.w, m0b, Eb.v, Eb.w);
        Eb.u += m0b[0] + 0.5 * k.kp2 * E.i2l * m0b[4];
        f2b += m0b[1];
        f3b += 0.5 * m0b[4];
        thEb += 0.5 * k.kp2 * E.u * m0b[4];
        // f3 = U f2 + 2 th U, f2 = U^2 + th
        Eb.u += (f2 + 2.0 * E.i2l) * f3b;
        f2b += E.u * f3b;
        thEb += 2.0 * E.u * f3b;
        Eb.u += 2.0 * E.u * f2b;
        thEb += f2b;
    }
    Eb.th += thEb;
    maxw_vjp(We, E, g.gam, Eb, Web);
    // We = den_l tshift(m_l) + den_r tshift(m_r); F64 += p3 (den_l tshift(n_l) + den_r tshift(n_r))
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const Maxw<double>& S = side == 0 ? L : R;
        const double* h = side == 0 ? hl : hr;
        double* hb = side == 0 ? hlb : hrb;
        MaxwBar& Sb = side == 0 ? Lb : Rb;
        double m[5], n[5], ms[5], ns[5], mb[5], nb[5];
        half_moms(S, h, m, n);
        for (int q = 0; q < 5; q++) {
            ms[q] = m[q];
            ns[q] = n[q];
        }
        tshift(ms, S.v, S.w);
        tshift(ns, S.v, S.w);
        const double cS = tc.p3 * S.den;
        for (int q = 0; q < 5; q++) {
            mb[q] = S.den * Web[q];
            nb[q] = cS * Fbar[q];
        }
        const double cSb = dot5(ns, Fbar);
        Sb.den += dot5(ms, Web) + tc.p3 * cSb;
        cb.p3 += S.den * cSb;
        tshift_vjp(m, S.v, S.w, mb, Sb.v, Sb.w);
        tshift_vjp(n, S.v, S.w, nb, Sb.v, Sb.w);
        hb[0] += mb[0] + 0.5 * k.kp2 * S.i2l * mb[4];
        hb[1] += mb[1] + nb[0] + 0.5 * k.kp2 * S.i2l * nb[4];
        hb[2] += 0.5 * mb[4] + nb[1];
        hb[3] += 0.5 * nb[4];
        Sb.th += 0.5 * k.kp2 * (h[0] * mb[4] + h[1] * nb[4]);
        tab_vjp<4>(S.u, S.i2l, h, hb, Sb.u, Sb.th);
        half_start_vjp(S.u, S.lam, side == 0 ? 1.0 : -1.0, h[0], hb[0], hb[1], Sb.u, Sb.lam);
    }
    // the collision time: taue = tau0 + dt |pl - pr| / |pl + pr|, pl = den_l / lam_l
    const double taueb = cb.gm1 * tcd.gm1.d + cb.gm2 * tcd.gm2.d + cb.gm3 * tcd.gm3.d + cb.gm4 * tcd.gm4.d +
                         cb.gm5 * tcd.gm5.d + cb.p0 * tcd.p0.d + cb.p1 * tcd.p1.d + cb.p2 * tcd.p2.d +
                         cb.p3 * tcd.p3.d + cb.p4 * tcd.p4.d + cb.p5 * tcd.p5.d;
    {
        const double A = pl - pr, B = pl + pr;
        const double sA = A > 0 ? 1.0 : (A < 0 ? -1.0 : 0.0), sB = B > 0 ? 1.0 : (B < 0 ? -1.0 : 0.0);
        const double aB = ::fabs(B);
        const double fb = dt * taueb;
        const double iaB = g_rcp(aB);
        const double Ab = fb * sA * iaB, Bb = -fb * (::fabs(A) * iaB) * sB * iaB;
        const double plb = Ab + Bb, prb = Bb - Ab;
        const double ill = g_rcp(L.lam), ilr = g_rcp(R.lam);
        Lb.den += plb * ill;
        Lb.lam -= plb * pl * ill;
        Rb.den += prb * ilr;
        Rb.lam -= prb * pr * ilr;
    }
    maxw_vjp(Wl, L, g.gam, Lb, xbar + 10);
    maxw_vjp(Wr, R, g.gam, Rb, xbar + 15);
    // tau0 = collision_time(WL): five dual passes of the short closed form
    const Gas<Dual> gd = dual_gas(g);
    for (int i = 0; i < 5; i++) {
        Dual w[5];
        for (int j = 0; j < 5; j++) w[j] = Dual(WL[j], i == j ? 1.0 : 0.0);
        xbar[i] += taueb * collision_time<Dual>(w, gd).d;
    }
}

}  // namespace gkm
