"""Pins of the oracle's full particle properties (getParticleProp with the
rho, p and Laplacian splines of generateSplines) that do not retype its code:
scipy's natural cubic spline, exact reproduction of linear data, and the
already-pinned velocity sample and spline evaluation."""
import numpy as np
from scipy.interpolate import CubicSpline

import pyoracle as po
from cases import make_case

NC = 16


def knots(nc, dx):
    return (np.arange(nc + 6) - 3) * dx + 0.5 * dx


def test_lap_field_matches_scipy_natural_second_derivatives():
    rng = np.random.default_rng(0)
    nt = NC + 6
    dx = 2 * np.pi / NC
    f = rng.standard_normal((nt, nt, nt))
    lap = po.lap_field(po.spline_coefs(f), dx)
    t = knots(NC, dx)
    # at knots in the other two axes the tensor spline is the 1-D natural spline of that line
    ref = np.zeros_like(f)
    for axis in range(3):
        ref += np.apply_along_axis(lambda v: CubicSpline(t, v, bc_type="natural")(t, 2), axis, f)
    assert np.max(np.abs(lap - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_lap_field_of_linear_data_is_zero():
    nt = NC + 6
    dx = 2 * np.pi / NC
    t = knots(NC, dx)
    f = 3.0 * t[:, None, None] - 2.0 * t[None, :, None] + 0.5 * t[None, None, :] + 7.0
    lap = po.lap_field(po.spline_coefs(np.ascontiguousarray(f)), dx)
    assert np.max(np.abs(lap)) < 1e-11


def test_lap_field_agrees_with_spline_hessian_trace():
    rng = np.random.default_rng(1)
    nt = NC + 6
    dx = 2 * np.pi / NC
    c = po.spline_coefs(rng.standard_normal((nt, nt, nt)))
    lap = po.lap_field(c, dx)
    t = knots(NC, dx)
    for (j, i, k) in [(0, 0, 0), (5, 7, 3), (nt - 2, 1, nt - 2), (3, nt - 2, 9)]:
        _, _, h = po.spline_eval(c, dx, t[j], t[i], t[k])
        assert abs(np.trace(h) - lap[j, i, k]) <= 1e-10 * max(1.0, abs(lap[j, i, k]))


def test_props_velocity_columns_equal_the_pinned_sample():
    W, P, gas = make_case(NC, seed=2)
    pos = np.random.default_rng(3).uniform(0, 2 * np.pi, (200, 3))
    out = po.particle_props(W, P["gam"], gas.dx, pos)
    U, gU, gV, gW = po.particle_sample(po.velocity_coefs(W), gas.dx, pos)
    assert np.array_equal(out[:, 1:4], U)
    assert np.array_equal(out[:, 8:11], gU)
    assert np.array_equal(out[:, 11:14], gV)
    assert np.array_equal(out[:, 14:17], gW)


def test_props_density_pressure_and_hessian():
    W, P, gas = make_case(NC, seed=4)
    gam = P["gam"]
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 2 * np.pi, (40, 3))
    out = po.particle_props(W, gam, gas.dx, pos)
    den = W[0]
    u, v, w = W[1] / den, W[2] / den, W[3] / den
    p = (den * (gam - 1)) * (W[4] / den - 0.5 * (u * u + v * v + w * w))
    cd, cp = po.spline_coefs(den), po.spline_coefs(np.ascontiguousarray(p))
    for n in range(len(pos)):
        vd, gd, _ = po.spline_eval(cd, gas.dx, *pos[n])
        vp, gp, hp = po.spline_eval(cp, gas.dx, *pos[n])
        assert out[n, 0] == vd and np.array_equal(out[n, 5:8], gd)
        assert out[n, 4] == vp and np.array_equal(out[n, 17:20], gp)
        H = out[n, 23:32].reshape(3, 3)
        assert np.array_equal(H, hp) and np.array_equal(H, H.T)
    # d^2 p / dy^2 at knot x, z and arbitrary y is the 1-D natural spline's
    t = knots(NC, gas.dx)
    i, k = 7, 11
    ys = rng.uniform(t[0], t[-1], 10)
    pts = np.stack([ys, np.full(10, t[i]), np.full(10, t[k])], 1)
    out = po.particle_props(W, gam, gas.dx, pts)
    ref = CubicSpline(t, p[:, i, k], bc_type="natural")(ys, 2)
    assert np.max(np.abs(out[:, 23] - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_props_laplacian_columns_follow_lap_splines():
    W, P, gas = make_case(NC, seed=6)
    pos = np.random.default_rng(7).uniform(0, 2 * np.pi, (30, 3))
    out = po.particle_props(W, P["gam"], gas.dx, pos)
    cs = po.velocity_coefs(W)
    for q in range(3):
        cl = po.spline_coefs(po.lap_field(cs[q], gas.dx))
        for n in range(len(pos)):
            vl, gl, _ = po.spline_eval(cl, gas.dx, *pos[n])
            assert out[n, 20 + q] == vl
            assert np.array_equal(out[n, 32 + 3 * q:35 + 3 * q], gl)
