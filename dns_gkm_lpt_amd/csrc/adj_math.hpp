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
    const double out = g_div(n0 * q0 + n1 * q1 + n2 * q2 + n3 * q3, Dn);  // as wratio
    // out = N / Dn
    const double iDn = g_div(wbar, Dn);
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

}  // namespace gkm
