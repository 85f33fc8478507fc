// adj_math.hpp -- reverse-mode derivatives of device math used by the adjoint
// (no reference counterpart).
#pragma once
#include <hip/hip_runtime.h>

#include "gkm_flux.hpp"
#include "gkm_math.hpp"
// after gkm_math.hpp: its double-only functions must not see the Dual overloads
#include "gkm_dual.hpp"

namespace gkm {

// Gradient of weno_side(a..f, eps) (WENO-SYMOO, update.cpp:561-621) with
// respect to its six cells, scaled by wbar: g[k] += wbar * d weno_side / d x_k.
// Reverse mode over the same operations: the candidates q_k are linear, the
// smoothness indicators quadratic, and the maximum that replaces IS_3 follows
// the branch the forward evaluation takes (as the dual-number derivative does).
__host__ __device__ __forceinline__ void weno_side_grad(double a, double b, double c, double d, double e, double f,
                                                        double eps, double wbar, double g[6])
{
    const double r3 = 1.0 / 3.0, r6 = 1.0 / 6.0, r12 = 1.0 / 12.0;
    const double q0 = a * r3 - 7.0 * b * r6 + 11.0 * c * r6;
    const double q1 = -b * r6 + 5.0 * c * r6 + d * r3;
    const double q2 = c * r3 + 5.0 * d * r6 - e * r6;
    const double q3 = 11.0 * d * r6 - 7.0 * e * r6 + f * r3;
    const double u0 = a - 2.0 * b + c, v0 = a - 4.0 * b + 3.0 * c;
    const double u1 = b - 2.0 * c + d, v1 = -b + d;
    const double u2 = c - 2.0 * d + e, v2 = -3.0 * c + 4.0 * d - e;
    const double u3 = d - 2.0 * e + f, v3 = -5.0 * d + 8.0 * e - 3.0 * f;
    const double s0 = 13.0 * sq(u0) * r12 + sq(v0) * 0.25;
    const double s1 = 13.0 * sq(u1) * r12 + sq(v1) * 0.25;
    const double s2 = 13.0 * sq(u2) * r12 + sq(v2) * 0.25;
    const double s3r = 13.0 * sq(u3) * r12 + sq(v3) * 0.25;
    // smax = s0 > s1 ? s0 : s1; smax = smax > s2 ? smax : s2; s3 = smax > s3r ? smax : s3r
    const int i01 = s0 > s1 ? 0 : 1;
    const double m01 = i01 == 0 ? s0 : s1;
    const int imax = m01 > s2 ? i01 : 2;
    const double smax = m01 > s2 ? m01 : s2;
    const bool s3_is_max = smax > s3r;
    const double s3 = s3_is_max ? smax : s3r;
    const double D0 = eps + s0, D1 = eps + s1, D2 = eps + s2, D3 = eps + s3;
    const double p01 = D0 * D1, p23 = D2 * D3;
    const double n0 = 0.05 * D1 * p23, n1 = 0.45 * D0 * p23, n2 = 0.45 * p01 * D3, n3 = 0.05 * p01 * D2;
    const double Dn = n0 + n1 + n2 + n3;
    const double out = (n0 * q0 + n1 * q1 + n2 * q2 + n3 * q3) / Dn;
    // out = N / Dn
    const double iDn = wbar / Dn;
    const double q0b = n0 * iDn, q1b = n1 * iDn, q2b = n2 * iDn, q3b = n3 * iDn;
    const double n0b = (q0 - out) * iDn, n1b = (q1 - out) * iDn, n2b = (q2 - out) * iDn, n3b = (q3 - out) * iDn;
    double D0b = 0.45 * (n1b * p23 + n2b * D1 * D3) + 0.05 * n3b * D1 * D2;
    double D1b = 0.05 * (n0b * p23 + n3b * D0 * D2) + 0.45 * n2b * D0 * D3;
    double D2b = 0.05 * (n0b * D1 * D3 + n3b * p01) + 0.45 * n1b * D0 * D3;
    const double D3b = 0.05 * n0b * D1 * D2 + 0.45 * (n1b * D0 * D2 + n2b * p01);
    // s3 -> the indicator it was taken from
    double s3rb = 0.0;
    if (s3_is_max) {
        if (imax == 0) D0b += D3b;
        else if (imax == 1) D1b += D3b;
        else D2b += D3b;
    } else {
        s3rb = D3b;
    }
    // s = 13/12 u^2 + 1/4 v^2
    const double k13 = 13.0 * r12 * 2.0;
    const double u0b = k13 * u0 * D0b, v0b = 0.5 * v0 * D0b;
    const double u1b = k13 * u1 * D1b, v1b = 0.5 * v1 * D1b;
    const double u2b = k13 * u2 * D2b, v2b = 0.5 * v2 * D2b;
    const double u3b = k13 * u3 * s3rb, v3b = 0.5 * v3 * s3rb;
    g[0] += q0b * r3 + u0b + v0b;
    g[1] += -7.0 * q0b * r6 - q1b * r6 - 2.0 * u0b - 4.0 * v0b + u1b - v1b;
    g[2] += 11.0 * q0b * r6 + 5.0 * q1b * r6 + q2b * r3 + u0b + 3.0 * v0b - 2.0 * u1b + u2b - 3.0 * v2b;
    g[3] += q1b * r3 + 5.0 * q2b * r6 + 11.0 * q3b * r6 + u1b + v1b - 2.0 * u2b + 4.0 * v2b + u3b - 5.0 * v3b;
    g[4] += -q2b * r6 - 7.0 * q3b * r6 + u2b - v2b - 2.0 * u3b + 8.0 * v3b;
    g[5] += q3b * r3 + u3b - 3.0 * v3b;
}

// ---- transposes of the linear building blocks of gks_flux_tf (gkm_flux.hpp),
// each applied as y -> M^T y for the map the forward code applies
// tshift: in place
__host__ __device__ __forceinline__ void tshift_T(double m[5], double dV, double dW)
{
    m[0] += dV * m[2] + dW * m[3] + 0.5 * (dV * dV + dW * dW) * m[4];
    m[2] += dV * m[4];
    m[3] += dW * m[4];
}
// tunshift (the forward subtracts with the already-shifted m2, m3): in place
__host__ __device__ __forceinline__ void tunshift_T(double m[5], double dV, double dW)
{
    m[0] += -dV * m[2] - dW * m[3] + 0.5 * (dV * dV + dW * dW) * m[4];
    m[2] -= dV * m[4];
    m[3] -= dW * m[4];
}
// slope_t: a = S b; here b_bar = S^T a_bar
__host__ __device__ __forceinline__ void slope_t_T(const double ab[5], const TState<double>& s, double inv_prn,
                                                   double kp3, double ikp3, double bb[5])
{
    const double q = s.u * s.u + kp3 * s.th;
    const double k4 = inv_prn * (s.lam2 * s.lam2) * ikp3;
    const double a1 = ab[1] - s.u * ab[0];
    const double a4 = ab[4] - 0.5 * q * ab[0] - s.u * a1;
    double R2 = s.lam2 * a1 - 2.0 * s.u * k4 * a4;
    const double R5 = k4 * a4;
    bb[0] = ab[0] - q * R5 - s.u * R2;
    bb[1] = R2;
    bb[2] = s.lam2 * ab[2];
    bb[3] = s.lam2 * ab[3];
    bb[4] = 2.0 * R5;
}
// mom_u: acc += M a; here a_bar += M^T acc_bar
__host__ __device__ __forceinline__ void mom_u_T(const double* H, double th, double kp2, double kp4, const double accb[5],
                                                 double ab[5])
{
    const double kth = kp2 * th;
    const double s1 = 0.5 * (H[2] + kth * H[0]);
    const double s2 = 0.5 * (H[3] + kth * H[1]);
    const double s4 = 0.25 * (H[4] + 2.0 * kth * H[2] + kth * (kp4 * th) * H[0]);
    ab[0] += H[0] * accb[0] + H[1] * accb[1] + s1 * accb[4];
    ab[1] += H[1] * accb[0] + H[2] * accb[1] + s2 * accb[4];
    ab[2] += th * H[0] * accb[2];
    ab[3] += th * H[0] * accb[3];
    ab[4] += s1 * accb[0] + s2 * accb[1] + s4 * accb[4];
}
// mom_uvw (weight v for c = 2, w for c = 3)
__host__ __device__ __forceinline__ void mom_uvw_T(const double* H, double th, double kp4, const double accb[5],
                                                   double ab[5], int c)
{
    const double tH0 = th * H[0];
    const double t = 0.5 * th * (H[2] + kp4 * tH0);
    ab[c] += tH0 * accb[0] + th * H[1] * accb[1] + t * accb[4];
    ab[0] += tH0 * accb[c];
    ab[1] += th * H[1] * accb[c];
    ab[4] += t * accb[c];
}

// Adjoint of the face flux with respect to the inputs in which it is affine:
// for gks_flux_tf<double, double> with the WENO states Wl, Wr and the collision
// time tau (of the left cell) held fixed, the flux is affine in the six slope
// vectors D and in the cell states WL, WR (these enter only through the
// equilibrium state's normal slopes (We - WL) 2/dx and (WR - We) 2/dx).  Gives
// Dbar[d][q] = sum_k Fbar[k] dF_k/dD_d[q] and WLbar, WRbar likewise, at fixed tau.
// The prelude repeats gks_flux_tf's (states, time coefficients, half-space
// tables, equilibrium state); the affine terms are then walked backwards,
// building block by building block, through their transposes.
__host__ __device__ inline void flux_affine_adjoint(const double Wl[5], const double Wr[5], double tau, double dt,
                                                    double dx, const Gas<double>& g, const double Fbar[5],
                                                    double WLbar[5], double WRbar[5], double Dbar[6][5])
{
    const double pi = 3.141592653589793238462643383279502884;
    const double kp2 = double(g.K + 2), kp3 = double(g.K + 3), kp4 = double(g.K + 4);
    const double ikp3 = 1.0 / kp3, inv_prn = g.inv_prn;
    Maxw<double> L, R;
    maxw_init(L, Wl, g);
    maxw_init(R, Wr, g);
    {
        const double pl = L.den / L.lam, pr = R.den / R.lam;
        tau = tau + dt * (::fabs(pl - pr) / ::fabs(pl + pr));
    }
    const double ex = g_exp(-dt / tau);
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
    const double p4 = -Qc - tau * Pc;
    const double p5 = -tau * Pc;
    double hl[4], hr[4];
    {
        const double gl = 0.5 * g_exp(-L.lam * L.u * L.u) / g_sqrt(L.lam * pi);
        const double h0 = 0.5 * g_erfc(-g_sqrt(L.lam) * L.u);
        half_tab<4>(L.u, L.i2l, h0, L.u * h0 + gl, hl);
        const double gr = 0.5 * g_exp(-R.lam * R.u * R.u) / g_sqrt(R.lam * pi);
        const double r0 = 0.5 * g_erfc(g_sqrt(R.lam) * R.u);
        half_tab<4>(R.u, R.i2l, r0, R.u * r0 - gr, hr);
    }
    double We[5];
    {
        double m[5] = {hl[0], hl[1], 0.0, 0.0, 0.5 * (hl[2] + kp2 * L.i2l * hl[0])};
        tshift(m, L.v, L.w);
        for (int q = 0; q < 5; q++) We[q] = L.den * m[q];
        double n[5] = {hr[0], hr[1], 0.0, 0.0, 0.5 * (hr[2] + kp2 * R.i2l * hr[0])};
        tshift(n, R.v, R.w);
        for (int q = 0; q < 5; q++) We[q] += R.den * n[q];
    }
    Maxw<double> E;
    maxw_init(E, We, g);

    // ---- reverse: F = F64 + tshift(Fs, E.v, E.w)
    double Fs_b[5];
    for (int q = 0; q < 5; q++) Fs_b[q] = Fbar[q];
    tshift_T(Fs_b, E.v, E.w);
    // equilibrium tangential and time-slope terms
    const TState<double> sE{E.u, E.i2l, 2.0 * E.lam};
    double fuE[6];
    full_tab<6>(sE.u, sE.th, fuE);
    const double cBE = gm2 * E.den, cFE = p2 * E.den, c1 = p1 * E.den, idenE = 1.0 / E.den;
    double t_b[5] = {0, 0, 0, 0, 0}, tvwE_b[5], aE_b[5], Be_b[5], ayE_b[5] = {0, 0, 0, 0, 0},
           azE_b[5] = {0, 0, 0, 0, 0};
    mom_u_T(fuE + 1, sE.th, kp2, kp4, Fs_b, t_b);  // Fs += mom_u(fu + 1, cF tvw + c1 a)
    for (int q = 0; q < 5; q++) {
        tvwE_b[q] = cFE * t_b[q];
        aE_b[q] = c1 * t_b[q];
    }
    slope_t_T(aE_b, sE, inv_prn, kp3, ikp3, Be_b);  // a = slope_t(Be / den)
    for (int q = 0; q < 5; q++) Be_b[q] *= idenE;
    {
        double tmp[5] = {0, 0, 0, 0, 0};
        mom_uvw_T(fuE + 1, sE.th, kp4, Fs_b, tmp, 3);  // Fs += mom_uvw(fu + 1, cF az, 3)
        for (int q = 0; q < 5; q++) azE_b[q] += cFE * tmp[q];
        for (int q = 0; q < 5; q++) tmp[q] = 0.0;
        mom_uvw_T(fuE + 1, sE.th, kp4, Fs_b, tmp, 2);  // Fs += mom_uvw(fu + 1, cF ay, 2)
        for (int q = 0; q < 5; q++) ayE_b[q] += cFE * tmp[q];
        for (int q = 0; q < 5; q++) tmp[q] = 0.0;
        mom_u_T(fuE, sE.th, kp2, kp4, Be_b, tmp);  // Be += mom_u(fu, cB tvw)
        for (int q = 0; q < 5; q++) tvwE_b[q] += cBE * tmp[q];
        for (int q = 0; q < 5; q++) tmp[q] = 0.0;
        mom_uvw_T(fuE, sE.th, kp4, Be_b, tmp, 3);  // Be += mom_uvw(fu, cB az, 3)
        for (int q = 0; q < 5; q++) azE_b[q] += cBE * tmp[q];
        for (int q = 0; q < 5; q++) tmp[q] = 0.0;
        mom_uvw_T(fuE, sE.th, kp4, Be_b, tmp, 2);  // Be += mom_uvw(fu, cB ay, 2)
        for (int q = 0; q < 5; q++) ayE_b[q] += cBE * tmp[q];
    }
    for (int q = 0; q < 5; q++) {  // tvw = V ay + W az
        ayE_b[q] += E.v * tvwE_b[q];
        azE_b[q] += E.w * tvwE_b[q];
    }
    double bye_b[5], bze_b[5];
    slope_t_T(ayE_b, sE, inv_prn, kp3, ikp3, bye_b);  // ay = slope_t(bye)
    slope_t_T(azE_b, sE, inv_prn, kp3, ikp3, bze_b);

    // ---- equilibrium state's normal slopes: a = slope_t(tunshift((We - WL) c)) for
    // the u > 0 half, (WR - We) c for the u < 0 half, entering Be (cB, hp+1) and Fs (cF, hp+2)
    {
        const double ge = 0.5 * g_exp(-E.lam * E.u * E.u) / g_sqrt(E.lam * pi);
        const double hp0 = 0.5 * g_erfc(-g_sqrt(E.lam) * E.u);
        const double hm0 = 1.0 - hp0;
        const double c = 2.0 / (E.den * dx);
        for (int half = 0; half < 2; half++) {
            double hh[7];
            if (half == 0) half_tab<7>(sE.u, sE.th, hp0, E.u * hp0 + ge, hh);
            else half_tab<7>(sE.u, sE.th, hm0, E.u * hm0 - ge, hh);
            double a_b[5] = {0, 0, 0, 0, 0}, tmp[5] = {0, 0, 0, 0, 0}, bb[5];
            mom_u_T(hh + 1, sE.th, kp2, kp4, Be_b, tmp);
            for (int q = 0; q < 5; q++) a_b[q] = cBE * tmp[q];
            for (int q = 0; q < 5; q++) tmp[q] = 0.0;
            mom_u_T(hh + 2, sE.th, kp2, kp4, Fs_b, tmp);
            for (int q = 0; q < 5; q++) a_b[q] += cFE * tmp[q];
            slope_t_T(a_b, sE, inv_prn, kp3, ikp3, bb);
            tunshift_T(bb, E.v, E.w);
            double* out = half == 0 ? WLbar : WRbar;
            const double sgn = half == 0 ? -c : c;
            for (int q = 0; q < 5; q++) out[q] = sgn * bb[q];
        }
    }

    // ---- the two sides (their sums entered bye, bze, Be, Fs after tshift to E's frame)
    for (int side = 0; side < 2; side++) {
        const Maxw<double>& S = side == 0 ? L : R;
        const double* h = side == 0 ? hl : hr;
        const TState<double> st{S.u, S.i2l, 2.0 * S.lam};
        const double iden = 1.0 / S.den, dV = S.v - E.v, dW = S.w - E.w;
        const double cB = (gm4 + gm5) * S.den, cA = gm5 * S.den, cF = p4 * S.den, cA2 = p5 * S.den;
        double fu[6], hu[7];
        full_tab<6>(st.u, st.th, fu);
        half_tab<7>(st.u, st.th, h[0], h[1], hu);
        double sy_b[5], sz_b[5], sB_b[5], sF_b[5];
        for (int q = 0; q < 5; q++) {
            sy_b[q] = bye_b[q];
            sz_b[q] = bze_b[q];
            sB_b[q] = Be_b[q];
            sF_b[q] = Fs_b[q];
        }
        tshift_T(sy_b, dV, dW);
        tshift_T(sz_b, dV, dW);
        tshift_T(sB_b, dV, dW);
        tshift_T(sF_b, dV, dW);
        double ax_b[5] = {0, 0, 0, 0, 0}, ay_b[5] = {0, 0, 0, 0, 0}, az_b[5] = {0, 0, 0, 0, 0};
        double tvw_b[5] = {0, 0, 0, 0, 0}, A_b[5] = {0, 0, 0, 0, 0}, tmp[5];
        auto clear = [&] {
            for (int q = 0; q < 5; q++) tmp[q] = 0.0;
        };
        // flux terms: sF = mom_u(hu+2, cF ax) + mom_uvw(hu+1, cF ay, 2) + mom_uvw(hu+1, cF az, 3)
        //                  + mom_u(hu+1, cF tvw + cA2 A)
        clear();
        mom_u_T(hu + 2, st.th, kp2, kp4, sF_b, tmp);
        for (int q = 0; q < 5; q++) ax_b[q] += cF * tmp[q];
        clear();
        mom_uvw_T(hu + 1, st.th, kp4, sF_b, tmp, 2);
        for (int q = 0; q < 5; q++) ay_b[q] += cF * tmp[q];
        clear();
        mom_uvw_T(hu + 1, st.th, kp4, sF_b, tmp, 3);
        for (int q = 0; q < 5; q++) az_b[q] += cF * tmp[q];
        clear();
        mom_u_T(hu + 1, st.th, kp2, kp4, sF_b, tmp);
        for (int q = 0; q < 5; q++) {
            tvw_b[q] += cF * tmp[q];
            A_b[q] += cA2 * tmp[q];
        }
        // den_e B_ terms: sB = mom_u(hu+1, cB ax) + mom_uvw(hu, cB ay, 2) + mom_uvw(hu, cB az, 3)
        //                      + mom_u(hu, cB tvw + cA A)
        clear();
        mom_u_T(hu + 1, st.th, kp2, kp4, sB_b, tmp);
        for (int q = 0; q < 5; q++) ax_b[q] += cB * tmp[q];
        clear();
        mom_uvw_T(hu, st.th, kp4, sB_b, tmp, 2);
        for (int q = 0; q < 5; q++) ay_b[q] += cB * tmp[q];
        clear();
        mom_uvw_T(hu, st.th, kp4, sB_b, tmp, 3);
        for (int q = 0; q < 5; q++) az_b[q] += cB * tmp[q];
        clear();
        mom_u_T(hu, st.th, kp2, kp4, sB_b, tmp);
        for (int q = 0; q < 5; q++) {
            tvw_b[q] += cB * tmp[q];
            A_b[q] += cA * tmp[q];
        }
        // equilibrium tangential right-hand sides: sy = mom_u(hu, ay), sz = mom_u(hu, az)
        mom_u_T(hu, st.th, kp2, kp4, sy_b, ay_b);
        mom_u_T(hu, st.th, kp2, kp4, sz_b, az_b);
        // time slope: A = slope_t(-(mom_u(fu+1, ax) + mom_uvw(fu, ay, 2) + mom_uvw(fu, az, 3) + mom_u(fu, tvw)))
        double nb[5];
        slope_t_T(A_b, st, inv_prn, kp3, ikp3, nb);
        for (int q = 0; q < 5; q++) nb[q] = -nb[q];
        mom_u_T(fu + 1, st.th, kp2, kp4, nb, ax_b);
        mom_uvw_T(fu, st.th, kp4, nb, ay_b, 2);
        mom_uvw_T(fu, st.th, kp4, nb, az_b, 3);
        mom_u_T(fu, st.th, kp2, kp4, nb, tvw_b);
        for (int q = 0; q < 5; q++) {  // tvw = V ay + W az
            ay_b[q] += S.v * tvw_b[q];
            az_b[q] += S.w * tvw_b[q];
        }
        // micro slopes from the lab-frame macro slopes D iden (slope_from_lab: tunshift, then slope_t)
        double* outs[3] = {Dbar[side], Dbar[2 + side], Dbar[4 + side]};
        double* ins[3] = {ax_b, ay_b, az_b};
        for (int k = 0; k < 3; k++) {
            double b[5];
            slope_t_T(ins[k], st, inv_prn, kp3, ikp3, b);
            tunshift_T(b, S.v, S.w);
            for (int q = 0; q < 5; q++) outs[k][q] = iden * b[q];
        }
    }
}

// flux of one face in the rotated frame from the 50 inputs x = (WL, WR, Wl, Wr,
// D[6]) (each 5 values, rotated frame), collision time from WL
template <typename T>
__host__ __device__ __forceinline__ void face_flux(const T x[50], T dt, T dx, const Gas<T>& g, T F[5])
{
    const T* WL = x;
    const T tau = collision_time<T>(WL, g);
    T D[6][5];
#pragma unroll
    for (int d = 0; d < 6; d++)
#pragma unroll
        for (int q = 0; q < 5; q++) D[d][q] = x[20 + 5 * d + q];
    gks_flux_tf<T, T>(x, x + 5, x + 10, x + 15, D, tau, dt, dx, g, F);
}

__host__ __device__ inline Gas<Dual> dual_gas(const Gas<double>& g)
{
    return Gas<Dual>{Dual(g.gam), Dual(g.R), Dual(g.inv_prn), Dual(g.mu0), Dual(g.T0), Dual(g.eps), g.K};
}

// xbar[k] = sum_q Fbar[q] dF_q/dx_k for the 50 inputs of face_flux:
//   slopes D and the cell states' normal-slope part: flux_affine_adjoint (one
//   reverse pass);
//   the collision time: one dual pass in tau, times dtau/dWL (five dual passes
//   of collision_time);
//   the WENO states Wl, Wr (in which the flux is not affine): one dual pass each.
__host__ __device__ inline void face_flux_adjoint(const double x[50], double dt, double dx, const Gas<double>& g,
                                                  const double Fbar[5], double xbar[50])
{
    const double tau = collision_time<double>(x, g);
    {
        double Dbar[6][5];
        flux_affine_adjoint(x + 10, x + 15, tau, dt, dx, g, Fbar, xbar, xbar + 5, Dbar);
        for (int d = 0; d < 6; d++)
            for (int q = 0; q < 5; q++) xbar[20 + 5 * d + q] = Dbar[d][q];
    }
    const Gas<Dual> gd = dual_gas(g);
    auto flux_dual = [&](int k, Dual taud, Dual F[5]) {
        Dual xd[50];
        for (int i = 0; i < 50; i++) xd[i] = Dual(x[i], i == k ? 1.0 : 0.0);
        Dual D[6][5];
        for (int d = 0; d < 6; d++)
            for (int q = 0; q < 5; q++) D[d][q] = xd[20 + 5 * d + q];
        gks_flux_tf<Dual, Dual>(xd, xd + 5, xd + 10, xd + 15, D, taud, Dual(dt), Dual(dx), gd, F);
    };
    {
        Dual F[5];
        flux_dual(-1, Dual(tau, 1.0), F);
        double taub = 0.0;
        for (int q = 0; q < 5; q++) taub += Fbar[q] * F[q].d;
#pragma unroll 1
        for (int k = 0; k < 5; k++) {
            Dual w[5];
            for (int i = 0; i < 5; i++) w[i] = Dual(x[i], i == k ? 1.0 : 0.0);
            xbar[k] += taub * collision_time<Dual>(w, gd).d;
        }
    }
#pragma unroll 1
    for (int k = 10; k < 20; k++) {
        Dual F[5];
        flux_dual(k, Dual(tau), F);
        double s = 0.0;
        for (int q = 0; q < 5; q++) s += Fbar[q] * F[q].d;
        xbar[k] = s;
    }
}

}  // namespace gkm
