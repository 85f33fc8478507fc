"""Closed-form and invariant pins of the oracle, independent of the reference
build: each check is chosen so that a plausible slip (a dropped term, a sign,
a swapped axis, a wrong boundary row) fails it.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import pyoracle as po
from cases import make_case, random_faces, random_stencils
from dns_gkm_lpt_amd import ic as icm

INNER = (slice(None),) + (slice(3, -3),) * 3


# ---------------------------------------------------------------- flux / WENO
def test_flux_mirror_symmetry():
    """x -> -x: left/right swap, normal momentum and normal slopes change sign;
    the flux must come out as F' = -M F with M = diag(1,-1,1,1,1)."""
    W, P, gas = make_case(16)
    X = random_faces(200, seed=9, jump=0.3)
    n = X.shape[0]
    V = X.reshape(n, 10, 5)
    M = np.array([1, -1, 1, 1, 1.0])
    Y = np.empty_like(V)
    Y[:, 0], Y[:, 1] = V[:, 1] * M, V[:, 0] * M            # WL' = M WR, WR' = M WL
    Y[:, 2], Y[:, 3] = V[:, 3] * M, V[:, 2] * M            # Wl' = M Wr, Wr' = M Wl
    Y[:, 4], Y[:, 5] = -V[:, 5] * M, -V[:, 4] * M          # normal slopes flip sign
    Y[:, 6], Y[:, 7] = V[:, 7] * M, V[:, 6] * M            # tangential slopes swap sides
    Y[:, 8], Y[:, 9] = V[:, 9] * M, V[:, 8] * M
    tau = np.full(n, 4e-6)
    F = po.flux(X, tau, 2e-4, gas)
    G = po.flux(Y.reshape(n, 50), tau, 2e-4, gas)
    scale = np.max(np.abs(F), axis=0)
    assert np.all(np.max(np.abs(G + F * M), axis=0) <= 1e-11 * scale)


def test_flux_uniform_state_is_time_integrated_euler_flux():
    W, P, gas = make_case(16)
    rng = np.random.default_rng(3)
    for _ in range(20):
        rho, p = 1 + 0.2 * rng.random(), 86100 * (1 + 0.2 * rng.random())
        u = 80 * rng.standard_normal(3)
        E = p / 0.4 + 0.5 * rho * u @ u
        w = np.array([rho, *(rho * u), E])
        X = np.concatenate([w, w, w, w, np.zeros(30)])
        dt = 1e-4 * (1 + rng.random())
        F = po.flux(X, [1e-5], dt, gas)[0]
        euler = np.array([rho * u[0], rho * u[0] ** 2 + p, rho * u[0] * u[1], rho * u[0] * u[2], u[0] * (E + p)])
        assert np.allclose(F, dt * euler, rtol=1e-11, atol=1e-11 * dt * p)


def test_weno_exact_for_quadratic_cell_averages():
    """Every candidate stencil is exact for quadratics, so the reconstruction
    equals p(x_{i+1/2}) whatever the nonlinear weights are."""
    gas = po.make_gas(16)
    rng = np.random.default_rng(4)
    for _ in range(50):
        co = rng.standard_normal((4, 3)) * [1.0, 0.05, 0.01]
        co[0, 0] = 1.2  # density level
        co[1:, 0] *= 50

        def avg(c, i):  # cell average over [i-1/2, i+1/2]
            return c[0] + c[1] * i + c[2] * (i * i + 1.0 / 12)

        def point(c, x):
            return c[0] + c[1] * x + c[2] * x * x

        S = np.zeros((6, 5))
        Tco = np.array([300.0, 2.0, -0.3])
        for s, i in enumerate(range(-2, 4)):
            for q in range(4):
                S[s, q] = avg(co[q], i)
            rho, m = S[s, 0], S[s, 1:4]
            T = avg(Tco, i)
            S[s, 4] = rho * 287 * T / 0.4 + 0.5 * m @ m / rho
        out = po.weno(S.reshape(1, 30), gas)[0]
        for q in range(4):
            assert out[q] == pytest.approx(point(co[q], 0.5), rel=1e-12, abs=1e-12)
            assert out[5 + q] == pytest.approx(point(co[q], 0.5), rel=1e-12, abs=1e-12)
        for off in (0, 5):
            rho, m = out[off], out[off + 1:off + 4]
            T = 0.4 * (out[off + 4] / rho - 0.5 * m @ m / rho ** 2) / 287
            assert T == pytest.approx(point(Tco, 0.5), rel=1e-12)


def test_weno_mirror_and_constant():
    gas = po.make_gas(16)
    S = random_stencils(100, seed=5, jump=0.5).reshape(100, 6, 5)
    out = po.weno(S.reshape(100, 30), gas)
    rev = po.weno(S[:, ::-1].reshape(100, 30).copy(), gas)
    assert np.array_equal(out[:, 5:], rev[:, :5]) and np.array_equal(out[:, :5], rev[:, 5:])
    C = np.tile(S[:, 2:3], (1, 6, 1))
    o = po.weno(C.reshape(100, 30), gas)
    assert np.allclose(o[:, :5], S[:, 2], rtol=1e-13) and np.allclose(o[:, 5:], S[:, 2], rtol=1e-13)


def test_weno_nonlinear_weights_pick_smooth_side():
    """Across a jump between cells i and i+1 the left state follows the smooth
    left data (the near-discontinuity candidate weights are suppressed)."""
    gas = po.make_gas(16)
    S = np.zeros((6, 5))
    S[:, 0] = [1.0, 1.0, 1.0, 2.0, 2.0, 2.0]
    S[:, 4] = S[:, 0] * 287 * 300 / 0.4
    out = po.weno(S.reshape(1, 30), gas)[0]
    assert abs(out[0] - 1.0) < 1e-6 and abs(out[5] - 2.0) < 1e-6


# ------------------------------------------------------------- whole step
def test_uniform_state_is_steady():
    nc = 12
    ic = icm.make_ic(nc, seed=0)
    z = np.zeros((nc,) * 3)
    W = icm.prim_to_cons(z + 30.0, z - 10.0, z + 5.0, z + 1.1, z + 90000.0)
    gas = po.make_gas(nc, mu0=ic["params"]["mu0"])
    Wn = po.eulerian_step(W, 1e-4, gas)
    assert np.allclose(Wn, W, rtol=1e-13, atol=0)


def test_step_conserves_interior_sums():
    W, P, gas = make_case(16, seed=2, perturb=0.1)
    Wn = po.eulerian_step(W, po.cfl_dt(W, 0.5, gas), gas, nthreads=4)
    s0, s1 = W[INNER].sum(axis=(1, 2, 3)), Wn[INNER].sum(axis=(1, 2, 3))
    assert np.all(np.abs(s1 - s0) <= 1e-13 * np.abs(W[INNER]).sum(axis=(1, 2, 3)))


def _swap_xy(W):
    V = W.transpose(0, 2, 1, 3)[[0, 2, 1, 3, 4]]
    return np.ascontiguousarray(V)


def _swap_xz(W):
    V = W.transpose(0, 1, 3, 2)[[0, 3, 2, 1, 4]]
    return np.ascontiguousarray(V)


def test_step_commutes_with_xy_transposition():
    """x<->y (with u<->v) is a symmetry of the scheme: the x and y sweeps map
    onto each other and the z sweep is symmetric in its two tangential axes."""
    W, P, gas = make_case(16, seed=3, perturb=0.1)
    dt = po.cfl_dt(W, 0.5, gas)
    a = _swap_xy(po.eulerian_step(W, dt, gas, nthreads=4))
    b = po.eulerian_step(_swap_xy(W), dt, gas, nthreads=4)
    d = np.max(np.abs(a - b), axis=(1, 2, 3)) / np.max(np.abs(W - a), axis=(1, 2, 3))
    assert np.all(d < 1e-9), d


def test_step_breaks_xz_symmetry_by_derivsZ_pairing():
    """derivsZ pairs the x-derivative with v and the y-derivative with u
    (update.cpp:412-438), so x<->z is not a symmetry of the reference step."""
    W, P, gas = make_case(16, seed=3, perturb=0.1)
    dt = po.cfl_dt(W, 0.5, gas)
    a = _swap_xz(po.eulerian_step(W, dt, gas, nthreads=4))
    b = po.eulerian_step(_swap_xz(W), dt, gas, nthreads=4)
    d = np.max(np.abs(a - b), axis=(1, 2, 3)) / np.max(np.abs(W - a), axis=(1, 2, 3))
    assert np.max(d[1:4]) > 1e-6, d


def test_halo_closed_form():
    rng = np.random.default_rng(0)
    nc = 9
    W = rng.standard_normal((5, nc + 6, nc + 6, nc + 6))
    H = po.halo_fill(W.copy())
    idx = (np.arange(nc + 6) - 3) % nc + 3
    assert np.array_equal(H, W[:, idx][:, :, idx][:, :, :, idx])


def test_cfl_closed_form():
    W, P, gas = make_case(16, seed=1)
    w = W[INNER]
    u = w[1:4] / w[0]
    p = 0.4 * (w[4] - 0.5 * w[0] * (u * u).sum(0))
    vmax = np.max(np.sqrt((u * u).sum(0)) + np.sqrt(1.4 * p / w[0]))
    assert po.cfl_dt(W, 0.5, gas) == pytest.approx(0.5 * gas.dx / vmax, rel=1e-15)


# ------------------------------------------------------------------ spline
def _natural_tensor_eval(f, start, h, pt):
    """Natural tricubic interpolant at pt via scipy, axis by axis."""
    n = f.shape[0]
    x = start + h * np.arange(n)
    a = CubicSpline(x, f, axis=2, bc_type="natural")(pt[2])
    b = CubicSpline(x, a, axis=1, bc_type="natural")(pt[1])
    return CubicSpline(x, b, axis=0, bc_type="natural")(pt[0])


def test_spline_matches_scipy_natural_interpolant():
    nc = 10
    nt = nc + 6
    rng = np.random.default_rng(7)
    f = rng.standard_normal((nt, nt, nt))
    c = po.spline_coefs(f)
    dx = 2 * np.pi / nc
    start = -3 * dx + 0.5 * dx
    for _ in range(30):
        pt = rng.uniform(-2.4 * dx, 2 * np.pi + 2.4 * dx, 3)
        v, g, h = po.spline_eval(c, dx, *pt)
        assert v == pytest.approx(_natural_tensor_eval(f, start, dx, pt), abs=1e-11)
        eps = 1e-5
        for a in range(3):
            e = np.zeros(3)
            e[a] = eps
            fd = (po.spline_eval(c, dx, *(pt + e))[0] - po.spline_eval(c, dx, *(pt - e))[0]) / (2 * eps)
            assert g[a] == pytest.approx(fd, rel=1e-6, abs=1e-6)


def test_spline_interpolates_knots_and_reproduces_linear():
    nc = 8
    nt = nc + 6
    dx = 2 * np.pi / nc
    start = -3 * dx + 0.5 * dx
    rng = np.random.default_rng(8)
    f = rng.standard_normal((nt, nt, nt))
    c = po.spline_coefs(f)
    for _ in range(20):
        j, i, k = rng.integers(0, nt, 3)
        v, _, _ = po.spline_eval(c, dx, start + j * dx, start + i * dx, start + k * dx)
        assert v == pytest.approx(f[j, i, k], abs=1e-12)
    y, x, z = np.meshgrid(*(start + dx * np.arange(nt),) * 3, indexing="ij")
    lin = 0.3 + 1.5 * y - 0.7 * x + 2.0 * z
    c = po.spline_coefs(lin)
    for _ in range(20):
        pt = rng.uniform(-2 * dx, 2 * np.pi + 2 * dx, 3)
        v, g, h = po.spline_eval(c, dx, *pt)
        assert v == pytest.approx(0.3 + 1.5 * pt[0] - 0.7 * pt[1] + 2.0 * pt[2], abs=1e-11)
        assert np.allclose(g, [1.5, -0.7, 2.0], atol=1e-10) and np.allclose(h, 0, atol=1e-8)


# --------------------------------------------------------------- particles
def test_heun_uniform_velocity_translates_and_wraps():
    nc = 16
    z = np.zeros((nc,) * 3)
    W = icm.prim_to_cons(z + 40.0, z - 25.0, z + 10.0, z + 1.0, z + 86100.0)
    cs = po.velocity_coefs(W)
    dx = 2 * np.pi / nc
    pos = np.array([[0.1, 6.25, 3.0], [6.28, 0.01, 0.5], [3.0, 3.0, 6.2]])  # (y, x, z)
    U, _, _, _ = po.particle_sample(cs, dx, pos)
    assert np.allclose(U, [[40.0, -25.0, 10.0]] * 3, rtol=1e-13)
    dt = 2e-3
    new = po.particle_heun(cs, dx, pos, U, dt)
    expect = (pos + dt * np.array([-25.0, 40.0, 10.0])) % (2 * np.pi)
    assert np.allclose(new, expect, rtol=0, atol=1e-13)


def test_heun_linear_field_closed_form():
    """u = A x (reproduced exactly by the spline): one Heun step is
    x + (A x + A(x + A x dt)) dt / 2 = (I + A dt + A^2 dt^2 / 2) x."""
    nc = 16
    nt = nc + 6
    dx = 2 * np.pi / nc
    start = -3 * dx + 0.5 * dx
    g = start + dx * np.arange(nt)
    Y, X, Z = np.meshgrid(g, g, g, indexing="ij")
    A = np.array([[0.0, 0.4, -0.2], [-0.3, 0.1, 0.5], [0.2, -0.1, 0.0]])  # u = A [x, y, z]
    fields = [A[r, 0] * X + A[r, 1] * Y + A[r, 2] * Z for r in range(3)]
    cs = [po.spline_coefs(np.ascontiguousarray(f)) for f in fields]
    rng = np.random.default_rng(2)
    xyz = rng.uniform(2.0, 4.0, (50, 3))
    pos = np.ascontiguousarray(xyz[:, [1, 0, 2]])
    U = xyz @ A.T
    dt = 0.05
    new = po.particle_heun(cs, dx, pos, U, dt)
    M = np.eye(3) + A * dt + A @ A * dt * dt / 2
    expect = xyz @ M.T
    assert np.allclose(new[:, [1, 0, 2]], expect, atol=1e-11)


# -------------------------------------------------------------- statistics
def test_field_stats_taylor_green():
    """u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 at uniform rho, p:
    <rho|u|^2> = rho/4 (x2 components), <(du/dx)^2> = 1/8, skewness 0,
    enstrophy <|curl u|^2> = 3/4, S_ij S_ij = 3/8 (spline gradients, 64^3)."""
    nc = 64
    dx = 2 * np.pi / nc
    c = (np.arange(nc) + 0.5) * dx
    y, x, z = np.meshgrid(c, c, c, indexing="ij")
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    W = icm.prim_to_cons(u, v, w, np.full_like(u, 1.2), np.full_like(u, 86100.0))
    gas = po.make_gas(nc, mu0=0.0)
    cs = po.velocity_coefs(W)
    s = po.field_stats(W, cs, gas, 86100.0, 0.3, 1.0)
    assert s.tke == pytest.approx(1.2 * 0.25, rel=1e-12)
    assert s.ens == pytest.approx(0.75, rel=2e-3)
    assert s.sijsij == pytest.approx(0.375, rel=2e-3)
    assert abs(s.skew[0]) < 1e-10 and abs(s.skew[1]) < 1e-10
    assert s.kurt[0] == pytest.approx(8 * (3 / 8) ** 3 / (1 / 8) ** 2 / 8, rel=5e-3)
    assert s.prms == 0.0 and s.diss == 0.0


def test_particle_stats_closed_form():
    rng = np.random.default_rng(1)
    n = 1000
    U = rng.standard_normal((n, 3))
    G = [rng.standard_normal((n, 3)) for _ in range(3)]
    s = po.particle_stats(U, *G, n_total=2 * n)
    assert s.tke == pytest.approx((U ** 2).sum() / (2 * n), rel=1e-13)
    dudx, dwdz = G[0][:, 1], G[2][:, 2]
    m2, m3 = (dudx ** 2).sum() / (2 * n), (dudx ** 3).sum() / (2 * n)
    assert s.skew_u == pytest.approx(m3 / m2 ** 1.5, rel=1e-12)
    m2, m3 = (dwdz ** 2).sum() / (2 * n), (dwdz ** 3).sum() / (2 * n)
    assert s.skew_w == pytest.approx(m3 / m2 ** 1.5, rel=1e-12)
