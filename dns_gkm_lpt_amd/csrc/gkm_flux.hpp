// gkm_flux.hpp -- GKS interface flux evaluated in tangential rest frames.
//
// Computes the time-integrated flux of the reference's flux() (GKMflux.cpp:11-229,
// moments MCal GKMflux.cpp:238-278, slopes slopesolver GKMflux.cpp:283-297 with
// a[4] scaled by 1/Pr) with the same terms, reorganised for fewer operations:
//
//  * Every Maxwellian (left, right, equilibrium) is full-space in v and w.  In the
//    frame that moves with the state's tangential velocity (V, W) its odd v and w
//    moments vanish, so a moment <u^k v^l w^m psi (a.psi)> of that state needs
//    only the u table H and theta = 1/(2 lambda):  ~15 operations instead of the
//    ~115 of a product-table evaluation.
//  * Micro slopes are solved in that frame.  The slope solver commutes with the
//    Galilean shift (b -> T^-1 b, a -> T^T a), including the 1/Pr factor on a[4],
//    so solving for the shifted macro slope gives the shifted micro slope.
//  * The moment map is linear in the micro slope, so terms that share a state, a
//    weight u^k v^l w^m and a destination are merged before their moment is taken.
//  * The slope-dependent sums of all three states are accumulated in the frame
//    of the equilibrium state and shifted back to the lab frame once.
//
// A tangential shift by (dV, dW) acts on a moment-like vector m = <psi f> as
//   m2 += dV m0,  m3 += dW m0,  m4 += dV m2 + dW m3 + (dV^2 + dW^2)/2 m0
// (with m2, m3 before the update), because psi4 = psi4' + dV v' + dW w' + (dV^2+dW^2)/2.
//
// Types: TE carries the prelude (primitive states, collision coefficients), the
// equilibrium state, and the Euler part of the flux (the p0 and p3 terms, which
// carry dt times the inviscid flux, and the gm1/gm3 terms of the time slope, which
// cancel to rounding); TS carries every slope-dependent term.  TE = TS = double is
// the FP64 path, TE = double / TS = float the mixed path, TE = TS = float the
// all-FP32 diagnostic path.
#pragma once
#include "gkm_math.hpp"

namespace gkm {

// tangential shift of a moment-like vector (see above), and its inverse
template <typename T>
__host__ __device__ __forceinline__ void tshift(T m[5], T dV, T dW)
{
    const T m2 = m[2], m3 = m[3];
    m[4] += dV * m2 + dW * m3 + T(0.5) * (dV * dV + dW * dW) * m[0];
    m[2] = m2 + dV * m[0];
    m[3] = m3 + dW * m[0];
}
template <typename T>
__host__ __device__ __forceinline__ void tunshift(T m[5], T dV, T dW)
{
    m[2] -= dV * m[0];
    m[3] -= dW * m[0];
    m[4] -= dV * m[2] + dW * m[3] + T(0.5) * (dV * dV + dW * dW) * m[0];
}

// A state seen in its tangential rest frame: normal velocity u, theta = 1/(2 lambda).
template <typename T>
struct TState {
    T u, th, lam2;  // lam2 = 2 lambda = 1/theta
};

// micro slope from macro slope in the state's tangential rest frame
// (slopesolver with V = W = 0; GKMflux.cpp:283-297, a[4] scaled by 1/Pr)
template <typename T>
__host__ __device__ __forceinline__ void slope_t(const T b[5], const TState<T>& s, T inv_prn, T kp3, T ikp3,
                                                 T a[5])
{
    const T R2 = b[1] - s.u * b[0];
    const T q = s.u * s.u + kp3 * s.th;
    const T R5 = T(2) * b[4] - b[0] * q;
    a[4] = inv_prn * (R5 - T(2) * s.u * R2) * (s.lam2 * s.lam2) * ikp3;
    a[3] = s.lam2 * b[3];
    a[2] = s.lam2 * b[2];
    a[1] = s.lam2 * R2 - s.u * a[4];
    a[0] = b[0] - a[1] * s.u - T(0.5) * a[4] * q;
}

// Moments in a tangential rest frame.  H points at u-table entry k (full or half
// space), so H[n] = <u^(k+n)>; th = theta; kp2 = K+2, kp4 = K+4.
//   mom_u            acc += <u^k psi (a.psi)>
//   mom_uvw(.., 2)   acc += <u^k v psi (a.psi)>
//   mom_uvw(.., 3)   acc += <u^k w psi (a.psi)>
template <typename T>
__host__ __device__ __forceinline__ void mom_u(const T* H, T th, T kp2, T kp4, const T a[5], T acc[5])
{
    const T kth = kp2 * th;
    const T s1 = T(0.5) * (H[2] + kth * H[0]);
    const T s2 = T(0.5) * (H[3] + kth * H[1]);
    const T s4 = T(0.25) * (H[4] + T(2) * kth * H[2] + kth * (kp4 * th) * H[0]);
    acc[0] += a[0] * H[0] + a[1] * H[1] + a[4] * s1;
    acc[1] += a[0] * H[1] + a[1] * H[2] + a[4] * s2;
    acc[2] += a[2] * (th * H[0]);
    acc[3] += a[3] * (th * H[0]);
    acc[4] += a[0] * s1 + a[1] * s2 + a[4] * s4;
}
template <typename T>
__host__ __device__ __forceinline__ void mom_uvw(const T* H, T th, T kp4, const T a[5], T acc[5], int c)
{
    // c = 2: weight v (uses a[2], fills slot 2); c = 3: weight w (a[3], slot 3)
    const T tH0 = th * H[0];
    const T t = T(0.5) * th * (H[2] + kp4 * tH0);
    const T ac = a[c];
    acc[0] += ac * tH0;
    acc[1] += ac * (th * H[1]);
    acc[c] += a[0] * tH0 + a[1] * (th * H[1]) + a[4] * t;
    acc[4] += ac * t;
}

// macro slope b (lab frame, TE) of a state with tangential velocity (V, W) ->
// micro slope in the state's rest frame (TS)
template <typename TE, typename TS>
__host__ __device__ __forceinline__ void slope_from_lab(TE b0, TE b1, TE b2, TE b3, TE b4, TE V, TE W,
                                                        const TState<TS>& s, TS inv_prn, TS kp3, TS ikp3, TS a[5])
{
    TE m[5] = {b0, b1, b2, b3, b4};
    tunshift(m, V, W);
    TS bs[5];
#pragma unroll
    for (int q = 0; q < 5; q++) bs[q] = (TS)m[q];
    slope_t(bs, s, inv_prn, kp3, ikp3, a);
}

// Inputs come through an accessor `in` (all in the rotated frame, normal momentum in slot 1):
//   in.WL(q), in.WR(q) : cell states either side of the face
//   in.Wl(q), in.Wr(q) : WENO states at the face
//   in.D(d, q)         : d = 0..5 -> DWxl, DWxr, DWyl, DWyr, DWzl, DWzr (the derivsX/Y/Z slopes)
// Each value is fetched where it is used, so a caller that keeps them in LDS or
// cache does not hold them in registers across the whole flux.
// tau: collision time of the left cell; the pressure-jump term is added here.
template <typename TE, typename TS, typename IN>
__host__ __device__ __forceinline__ void gks_flux_tf(const IN& in, TE tau, TE dt, TE dx, const Gas<TE>& g, TE F[5])
{
    const TE isqpi = TE(0.5641895835477562869480794515607726);  // 1/sqrt(pi)
    const TE kp2e = TE(g.K + 2);
    const TS kp2 = TS(g.K + 2), kp3 = TS(g.K + 3), kp4 = TS(g.K + 4);
    const TS ikp3 = TS(1) / kp3;
    const TS inv_prn = (TS)g.inv_prn;

    Maxw<TE> L, R;
    {
        TE Wl[5], Wr[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            Wl[q] = in.Wl(q);
            Wr[q] = in.Wr(q);
        }
        maxw_init(L, Wl, g);
        maxw_init(R, Wr, g);
    }
    {   // pressure-jump term of the collision time (GKMflux.cpp:156-157)
        const TE pl = TE(2) * L.den * L.i2l, pr = TE(2) * R.den * R.i2l;  // rho / lambda
        tau = tau + dt * g_div(fabs(pl - pr), fabs(pl + pr));
    }
    // time-integration coefficients (GKMflux.cpp:162-167, 197-204)
    const TE ex = g_exp(-g_div(dt, tau));
    const TE om = TE(1) - ex;
    const TE igm0 = g_rcp(dt - tau * om);
    const TE gm1 = -om * igm0;
    const TE gm2 = (-dt + TE(2) * tau * om - dt * ex) * igm0;
    const TE gm3 = -gm1;
    const TE gm4 = (dt * ex - tau * om) * igm0;
    const TE gm5 = tau * gm3;
    const TE Pc = tau * om;
    const TE Qc = -tau * dt * ex + tau * tau * om;
    const TE p0 = dt - Pc;
    const TE p1 = TE(0.5) * dt * dt - tau * p0;
    const TE p2 = -tau * p0 + Qc;
    const TE p3 = Pc;
    const TE p4 = -Qc - tau * Pc;
    const TE p5 = -tau * Pc;

    // half-space u tables of the left (u > 0) and right (u < 0) states, entries 0..3 in TE
    // (0.5 erfc and the exponential share exp(-lambda u^2) = exp(-z^2), z = sqrt(lambda) u)
    TE hl[4], hr[4];
    {
        const TE sl = g_sqrt(L.lam), zl = sl * L.u, el = g_exp(-zl * zl);
        const TE gl = g_div(TE(0.5) * isqpi * el, sl);
        const TE h0 = half_erfc(-zl, el);
        half_tab<4>(L.u, L.i2l, h0, L.u * h0 + gl, hl);
        const TE sr = g_sqrt(R.lam), zr = sr * R.u, er = g_exp(-zr * zr);
        const TE gr = g_div(TE(0.5) * isqpi * er, sr);
        const TE r0 = half_erfc(zr, er);
        half_tab<4>(R.u, R.i2l, r0, R.u * r0 - gr, hr);
    }

    // ---- TE part: equilibrium state, Euler part of the flux, gm1/gm3 part of den_e B_
    TE We[5], F64[5], Bz[5];
    {
        TE m[5];
        m[0] = hl[0]; m[1] = hl[1]; m[2] = 0; m[3] = 0; m[4] = TE(0.5) * (hl[2] + kp2e * L.i2l * hl[0]);
        tshift(m, L.v, L.w);
#pragma unroll
        for (int q = 0; q < 5; q++) We[q] = L.den * m[q];
        m[0] = hr[0]; m[1] = hr[1]; m[2] = 0; m[3] = 0; m[4] = TE(0.5) * (hr[2] + kp2e * R.i2l * hr[0]);
        tshift(m, R.v, R.w);
#pragma unroll
        for (int q = 0; q < 5; q++) We[q] += R.den * m[q];
        // p3 terms: <u psi> of the two half Maxwellians (GKMflux.cpp:226)
        const TE cl = p3 * L.den, cr = p3 * R.den;
        m[0] = hl[1]; m[1] = hl[2]; m[2] = 0; m[3] = 0; m[4] = TE(0.5) * (hl[3] + kp2e * L.i2l * hl[1]);
        tshift(m, L.v, L.w);
#pragma unroll
        for (int q = 0; q < 5; q++) F64[q] = cl * m[q];
        m[0] = hr[1]; m[1] = hr[2]; m[2] = 0; m[3] = 0; m[4] = TE(0.5) * (hr[3] + kp2e * R.i2l * hr[1]);
        tshift(m, R.v, R.w);
#pragma unroll
        for (int q = 0; q < 5; q++) F64[q] += cr * m[q];
    }
    Maxw<TE> E;
    maxw_init(E, We, g);
    {
        // full u table of E: 1, U, U^2 + th, U^3 + 3 U th
        const TE f2 = E.u * E.u + E.i2l, f3 = E.u * f2 + TE(2) * E.i2l * E.u;
        TE m[5];
        // p0 term: <u psi>_e (GKMflux.cpp:225)
        m[0] = E.u; m[1] = f2; m[2] = 0; m[3] = 0; m[4] = TE(0.5) * (f3 + kp2e * E.i2l * E.u);
        tshift(m, E.v, E.w);
        const TE ce = p0 * E.den;
#pragma unroll
        for (int q = 0; q < 5; q++) F64[q] += ce * m[q];
        // gm1 den_e <psi>_e + gm3 We (GKMflux.cpp:185): zero up to rounding
        m[0] = 1; m[1] = E.u; m[2] = 0; m[3] = 0; m[4] = TE(0.5) * (f2 + kp2e * E.i2l);
        tshift(m, E.v, E.w);
        const TE c1 = gm1 * E.den;
#pragma unroll
        for (int q = 0; q < 5; q++) Bz[q] = c1 * m[q] + gm3 * We[q];
    }

    // ---- TS part: every slope-dependent term, accumulated in E's tangential rest frame
    TS Be[5], Fs[5] = {0, 0, 0, 0, 0}, bye[5] = {0, 0, 0, 0, 0}, bze[5] = {0, 0, 0, 0, 0};
    {
        TE m[5] = {Bz[0], Bz[1], Bz[2], Bz[3], Bz[4]};
        tunshift(m, E.v, E.w);
#pragma unroll
        for (int q = 0; q < 5; q++) Be[q] = (TS)m[q];
    }
    const TS gm45 = (TS)(gm4 + gm5), tgm5 = (TS)gm5, tp4 = (TS)p4, tp5 = (TS)p5;

    // ---- equilibrium state, normal-slope terms (E's rest frame, GKMflux.cpp:115-116,
    // 124-125, 170-171, 210-211).  With FP32 slope terms they go first, so that We and
    // the cell states are dead before the side blocks; with FP64 slope terms the
    // compiler allocates registers better when they follow the side blocks.
    auto normal_terms = [&]() {
        const TState<TS> st{(TS)E.u, (TS)E.i2l, (TS)(TE(2) * E.lam)};
        const TS cB = (TS)(gm2 * E.den), cF = (TS)(p2 * E.den);
        const TE se = g_sqrt(E.lam), ze = se * E.u, ee = g_exp(-ze * ze);
        const TE ge = g_div(TE(0.5) * isqpi * ee, se);
        // erfc(x) + erfc(-x) = 2: the u<0 half's zeroth moment is 1 - that of the u>0 half
        const TE hp0 = half_erfc(-ze, ee);
        const TE hm0 = TE(1) - hp0;
        const TE c = TE(2) * g_rcp(E.den * dx);
        TS a[5], t[5];
        {   // u > 0 half with the left normal slope (GKMflux.cpp:115, 124, 170, 210)
            TS hp[7];
            half_tab<7>(st.u, st.th, (TS)hp0, (TS)(E.u * hp0 + ge), hp);
            slope_from_lab((We[0] - in.WL(0)) * c, (We[1] - in.WL(1)) * c, (We[2] - in.WL(2)) * c,
                           (We[3] - in.WL(3)) * c, (We[4] - in.WL(4)) * c, E.v, E.w, st, inv_prn, kp3, ikp3, a);
#pragma unroll
            for (int q = 0; q < 5; q++) t[q] = cB * a[q];
            mom_u(hp + 1, st.th, kp2, kp4, t, Be);
#pragma unroll
            for (int q = 0; q < 5; q++) t[q] = cF * a[q];
            mom_u(hp + 2, st.th, kp2, kp4, t, Fs);
        }
        {   // u < 0 half with the right normal slope (GKMflux.cpp:116, 125, 171, 211)
            TS hm[7];
            half_tab<7>(st.u, st.th, (TS)hm0, (TS)(E.u * hm0 - ge), hm);
            slope_from_lab((in.WR(0) - We[0]) * c, (in.WR(1) - We[1]) * c, (in.WR(2) - We[2]) * c,
                           (in.WR(3) - We[3]) * c, (in.WR(4) - We[4]) * c, E.v, E.w, st, inv_prn, kp3, ikp3, a);
#pragma unroll
            for (int q = 0; q < 5; q++) t[q] = cB * a[q];
            mom_u(hm + 1, st.th, kp2, kp4, t, Be);
#pragma unroll
            for (int q = 0; q < 5; q++) t[q] = cF * a[q];
            mom_u(hm + 2, st.th, kp2, kp4, t, Fs);
        }
    };
    constexpr bool kNormalFirst = sizeof(TS) == 4;
    if (kNormalFirst) normal_terms();

#pragma unroll
    for (int side = 0; side < 2; side++) {
        const Maxw<TE>& S = side == 0 ? L : R;
        const TE* h = side == 0 ? hl : hr;
        const TS V = (TS)S.v, W = (TS)S.w, den = (TS)S.den;
        const TState<TS> st{(TS)S.u, (TS)S.i2l, (TS)(TE(2) * S.lam)};
        const TE iden = g_rcp(S.den);
        TS ax[5], ay[5], az[5];
        slope_from_lab(in.D(side, 0) * iden, in.D(side, 1) * iden, in.D(side, 2) * iden, in.D(side, 3) * iden,
                       in.D(side, 4) * iden, S.v, S.w, st, inv_prn, kp3, ikp3, ax);
        slope_from_lab(in.D(2 + side, 0) * iden, in.D(2 + side, 1) * iden, in.D(2 + side, 2) * iden,
                       in.D(2 + side, 3) * iden, in.D(2 + side, 4) * iden, S.v, S.w, st, inv_prn, kp3, ikp3, ay);
        slope_from_lab(in.D(4 + side, 0) * iden, in.D(4 + side, 1) * iden, in.D(4 + side, 2) * iden,
                       in.D(4 + side, 3) * iden, in.D(4 + side, 4) * iden, S.v, S.w, st, inv_prn, kp3, ikp3, az);
        // V ay + W az: the lab-frame v and w weights split into rest-frame weight + (V, W)
        TS tvw[5];
#pragma unroll
        for (int q = 0; q < 5; q++) tvw[q] = V * ay[q] + W * az[q];

        // time slope A of the state: <(u ax + v ay + w az).psi psi> over the full space
        // (GKMflux.cpp:134-151), then the slope solver
        TS A[5];
        {
            TS fu[6];
            full_tab<6>(st.u, st.th, fu);
            TS b[5] = {0, 0, 0, 0, 0};
            mom_u(fu + 1, st.th, kp2, kp4, ax, b);
            mom_uvw(fu, st.th, kp4, ay, b, 2);
            mom_uvw(fu, st.th, kp4, az, b, 3);
            mom_u(fu, st.th, kp2, kp4, tvw, b);
#pragma unroll
            for (int q = 0; q < 5; q++) b[q] = -b[q];
            slope_t(b, st, inv_prn, kp3, ikp3, A);
        }
        TS hu[7];
        half_tab<7>(st.u, st.th, (TS)h[0], (TS)h[1], hu);

        if constexpr (sizeof(TS) == 8) {
            // FP64: each term goes through the shift into E's tangential rest
            // frame and straight into the accumulators (the shift is linear),
            // so no per-side sums stay live: fewer registers spilled.
            const TS dV = (TS)(S.v - E.v), dW = (TS)(S.w - E.w);
            auto put = [&](TS t[5], TS acc[5]) {
                tshift(t, dV, dW);
#pragma unroll
                for (int q = 0; q < 5; q++) acc[q] += t[q];
            };
            {   // equilibrium tangential slopes' right-hand sides (GKMflux.cpp:130-133, 141-144)
                TS t[5] = {0, 0, 0, 0, 0};
                mom_u(hu, st.th, kp2, kp4, ay, t);
                put(t, bye);
                TS u[5] = {0, 0, 0, 0, 0};
                mom_u(hu, st.th, kp2, kp4, az, u);
                put(u, bze);
            }
            {   // den_e B_ terms (GKMflux.cpp:183-190) and flux terms (GKMflux.cpp:223-228), by input
                const TS cB = gm45 * den, cF = tp4 * den;
                TS b[5] = {0, 0, 0, 0, 0}, f[5] = {0, 0, 0, 0, 0}, c[5];
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cB * ax[q];
                mom_u(hu + 1, st.th, kp2, kp4, c, b);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cF * ax[q];
                mom_u(hu + 2, st.th, kp2, kp4, c, f);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cB * ay[q];
                mom_uvw(hu, st.th, kp4, c, b, 2);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cF * ay[q];
                mom_uvw(hu + 1, st.th, kp4, c, f, 2);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cB * az[q];
                mom_uvw(hu, st.th, kp4, c, b, 3);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cF * az[q];
                mom_uvw(hu + 1, st.th, kp4, c, f, 3);
                const TS cA = tgm5 * den, cA2 = tp5 * den;
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cB * tvw[q] + cA * A[q];
                mom_u(hu, st.th, kp2, kp4, c, b);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cF * tvw[q] + cA2 * A[q];
                mom_u(hu + 1, st.th, kp2, kp4, c, f);
                put(b, Be);
                put(f, Fs);
            }
        } else {
            TS sy[5] = {0, 0, 0, 0, 0}, sz[5] = {0, 0, 0, 0, 0}, sB[5] = {0, 0, 0, 0, 0}, sF[5] = {0, 0, 0, 0, 0};
            // equilibrium tangential slopes' right-hand sides (GKMflux.cpp:130-133, 141-144)
            mom_u(hu, st.th, kp2, kp4, ay, sy);
            mom_u(hu, st.th, kp2, kp4, az, sz);
            // den_e B_ terms (GKMflux.cpp:183-190)
            {
                const TS cB = gm45 * den;
                TS c[5];
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cB * ax[q];
                mom_u(hu + 1, st.th, kp2, kp4, c, sB);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cB * ay[q];
                mom_uvw(hu, st.th, kp4, c, sB, 2);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cB * az[q];
                mom_uvw(hu, st.th, kp4, c, sB, 3);
                const TS cA = tgm5 * den;
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cB * tvw[q] + cA * A[q];
                mom_u(hu, st.th, kp2, kp4, c, sB);
            }
            // flux terms (GKMflux.cpp:223-228)
            {
                const TS cF = tp4 * den;
                TS c[5];
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cF * ax[q];
                mom_u(hu + 2, st.th, kp2, kp4, c, sF);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cF * ay[q];
                mom_uvw(hu + 1, st.th, kp4, c, sF, 2);
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cF * az[q];
                mom_uvw(hu + 1, st.th, kp4, c, sF, 3);
                const TS cA = tp5 * den;
#pragma unroll
                for (int q = 0; q < 5; q++) c[q] = cF * tvw[q] + cA * A[q];
                mom_u(hu + 1, st.th, kp2, kp4, c, sF);
            }
            // into E's tangential rest frame
            const TS dV = (TS)(S.v - E.v), dW = (TS)(S.w - E.w);
            tshift(sy, dV, dW);
            tshift(sz, dV, dW);
            tshift(sB, dV, dW);
            tshift(sF, dV, dW);
#pragma unroll
            for (int q = 0; q < 5; q++) {
                bye[q] += sy[q];
                bze[q] += sz[q];
                Be[q] += sB[q];
                Fs[q] += sF[q];
            }
        }
    }

    if (!kNormalFirst) normal_terms();

    // ---- equilibrium state, tangential and time-slope terms (its own tangential rest frame)
    {
        const TState<TS> st{(TS)E.u, (TS)E.i2l, (TS)(TE(2) * E.lam)};
        const TS V = (TS)E.v, W = (TS)E.w;
        const TS cB = (TS)(gm2 * E.den), cF = (TS)(p2 * E.den);
        TS a[5], t[5];
        TS fu[6];
        full_tab<6>(st.u, st.th, fu);
        // tangential slopes of E (GKMflux.cpp:148-149, 172-173, 212-213); bye, bze are
        // already in E's rest frame
        TS ay[5], az[5], tvw[5];
        slope_t(bye, st, inv_prn, kp3, ikp3, ay);
        slope_t(bze, st, inv_prn, kp3, ikp3, az);
#pragma unroll
        for (int q = 0; q < 5; q++) tvw[q] = V * ay[q] + W * az[q];
#pragma unroll
        for (int q = 0; q < 5; q++) t[q] = cB * ay[q];
        mom_uvw(fu, st.th, kp4, t, Be, 2);
#pragma unroll
        for (int q = 0; q < 5; q++) t[q] = cB * az[q];
        mom_uvw(fu, st.th, kp4, t, Be, 3);
#pragma unroll
        for (int q = 0; q < 5; q++) t[q] = cB * tvw[q];
        mom_u(fu, st.th, kp2, kp4, t, Be);
#pragma unroll
        for (int q = 0; q < 5; q++) t[q] = cF * ay[q];
        mom_uvw(fu + 1, st.th, kp4, t, Fs, 2);
#pragma unroll
        for (int q = 0; q < 5; q++) t[q] = cF * az[q];
        mom_uvw(fu + 1, st.th, kp4, t, Fs, 3);
        // time slope of E from den_e B_ (GKMflux.cpp:183-191), then the p1 term
        // (GKMflux.cpp:225) merged with the (V, W) part of the tangential terms
        const TS iden = (TS)g_rcp(E.den);
        TS b[5];
#pragma unroll
        for (int q = 0; q < 5; q++) b[q] = Be[q] * iden;
        slope_t(b, st, inv_prn, kp3, ikp3, a);
        const TS c1 = (TS)(p1 * E.den);
#pragma unroll
        for (int q = 0; q < 5; q++) t[q] = cF * tvw[q] + c1 * a[q];
        mom_u(fu + 1, st.th, kp2, kp4, t, Fs);
    }
    // back to the lab frame and add the Euler part
    TE Fl[5];
#pragma unroll
    for (int q = 0; q < 5; q++) Fl[q] = (TE)Fs[q];
    tshift(Fl, E.v, E.w);
#pragma unroll
    for (int q = 0; q < 5; q++) F[q] = F64[q] + Fl[q];
}

// accessor over plain arrays (pointwise kernels, host checks)
template <typename T>
struct ArrIn {
    const T *wl, *wr, *fl, *fr;
    const T (*d)[5];
    __host__ __device__ T WL(int q) const { return wl[q]; }
    __host__ __device__ T WR(int q) const { return wr[q]; }
    __host__ __device__ T Wl(int q) const { return fl[q]; }
    __host__ __device__ T Wr(int q) const { return fr[q]; }
    __host__ __device__ T D(int k, int q) const { return d[k][q]; }
};

template <typename TE, typename TS>
__host__ __device__ __forceinline__ void gks_flux_tf(const TE WL[5], const TE WR[5], const TE Wl[5],
                                                     const TE Wr[5], const TE D[6][5], TE tau, TE dt, TE dx,
                                                     const Gas<TE>& g, TE F[5])
{
    const ArrIn<TE> in{WL, WR, Wl, Wr, D};
    gks_flux_tf<TE, TS>(in, tau, dt, dx, g, F);
}

}  // namespace gkm
