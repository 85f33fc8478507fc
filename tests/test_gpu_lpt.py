"""Parity of the gfx950 spline / particle / statistics kernels with the oracle.

Tolerances: sampled velocities and gradients 1e-11 relative to the field's
max; particle positions 1e-11 (minimum-image distance on the 2 pi torus);
statistics 1e-10 relative.
"""
import numpy as np
import pytest

import pyoracle as po
from cases import make_case, rel_err, torus_dist
from dns_gkm_lpt_amd import _lib as gl
from dns_gkm_lpt_amd import ic as icm

pytestmark = pytest.mark.gpu


def particles(n, seed):
    xyz = icm.uniform_particles(n, seed)
    return np.ascontiguousarray(xyz[:, [1, 0, 2]])  # reference order (y, x, z)


def test_sample_matches_oracle():
    nc = 32
    W, P, gas = make_case(nc, seed=1)
    pos = particles(2000, 1)
    cs = po.velocity_coefs(W)
    U, gU, gV, gW = po.particle_sample(cs, gas.dx, pos)
    with gl.Context(nc, P) as c:
        c.upload(W)
        c.lpt_upload(pos)
        U2, gU2, gV2, gW2 = c.lpt_sample()
    assert np.max(np.abs(U2 - U)) <= 1e-11 * np.max(np.abs(U))
    for a, b in ((gU2, gU), (gV2, gV), (gW2, gW)):
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))


def test_coupled_step_matches_oracle():
    nc, nsteps = 32, 3
    W, P, gas = make_case(nc, seed=2)
    pos = particles(1000, 2)
    Wr, xr = W.copy(), pos.copy()
    dts = []
    for _ in range(nsteps):
        dt = po.cfl_dt(Wr, 0.5, gas)
        dts.append(dt)
        cs_old = po.velocity_coefs(Wr)
        U, _, _, _ = po.particle_sample(cs_old, gas.dx, xr)
        Wr = po.eulerian_step(Wr, dt, gas, nthreads=8)
        cs_new = po.velocity_coefs(Wr)
        xr = po.particle_heun(cs_new, gas.dx, xr, U, dt)
    with gl.Context(nc, P) as c:
        c.upload(W)
        c.lpt_upload(pos)
        for dt in dts:
            c.lpt_coupled_step(dt)
        x, _ = c.lpt_download()
        out = c.download()
    assert np.all(rel_err(out, Wr) < 1e-11)
    assert np.max(torus_dist(x, xr)) < 1e-11


def test_field_stats_match_oracle():
    nc = 32
    W, P, gas = make_case(nc, seed=3)
    cs = po.velocity_coefs(W)
    ref = po.field_stats(W, cs, gas, P["p0"], P["Mt"], P["rho0"])
    with gl.Context(nc, P) as c:
        c.upload(W)
        s = c.stats()
    pairs = [(s.tke, ref.tke), (s.diss, ref.diss), (s.sijsij, ref.sijsij), (s.enstrophy, ref.ens),
             (s.prms, ref.prms), (s.trms, ref.trms), (s.vrms, ref.vrms)]
    pairs += [(s.skew[a], ref.skew[a]) for a in range(3)] + [(s.kurt[a], ref.kurt[a]) for a in range(3)]
    for a, b in pairs:
        assert abs(a - b) <= 1e-10 * abs(b) + 1e-300, (a, b)


def test_particle_stats_match_oracle():
    nc = 32
    W, P, gas = make_case(nc, seed=4)
    pos = particles(3000, 4)
    cs = po.velocity_coefs(W)
    U, gU, gV, gW = po.particle_sample(cs, gas.dx, pos)
    ref = po.particle_stats(U, gU, gV, gW)
    with gl.Context(nc, P) as c:
        c.upload(W)
        c.lpt_upload(pos)
        c.lpt_sample()
        s = c.lpt_stats()
    for a, b in ((s.tke, ref.tke), (s.sijsij, ref.sijsij), (s.skew_u, ref.skew_u), (s.skew_w, ref.skew_w)):
        assert abs(a - b) <= 1e-10 * abs(b), (a, b)


def test_particle_properties_match_oracle():
    """lpt_properties (getParticleProp with all eight splines) against the oracle,
    column by column: within 1e-10 of the column's largest magnitude."""
    nc = 32
    W, P, gas = make_case(nc, seed=5)
    pos = particles(2000, 5)
    ref = po.particle_props(W, P["gam"], gas.dx, pos)
    with gl.Context(nc, P) as c:
        c.upload(W)
        c.lpt_upload(pos)
        out = c.lpt_properties()
    assert out.shape == (2000, gl.PROP_WIDTH) == ref.shape
    err = np.max(np.abs(out - ref), axis=0) / np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
    worst = int(np.argmax(err))
    print("props worst column", gl.PROP_NAMES[worst], err[worst], flush=True)
    assert np.all(err < 1e-10), dict(zip(gl.PROP_NAMES, err))


def test_field_stats_closed_forms():
    """The GPU statistics on the closed-form fields that pin the oracle's
    dissipation and rms normalisations (test_oracle_stats_pins.py): Taylor-Green
    at T(z) = 380 (1 + 0.3 cos 2z) K with Sutherland viscosity, and a fluid at
    rest with p = p0 (1 + 0.07 sin x), rho = 1.3 (1 + 0.12 sin y)."""
    R = 287.0
    nc = 64
    dx = 2 * np.pi / nc
    c = (np.arange(nc) + 0.5) * dx
    y, x, z = np.meshgrid(c, c, c, indexing="ij")
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    T = 380.0 * (1 + 0.3 * np.cos(2 * z))
    W = icm.prim_to_cons(u, v, np.zeros_like(u), np.full_like(u, 0.9), 0.9 * R * T)
    P = {"gam": 1.4, "rho0": 1.0, "Mt": 0.3, "Re": 30.0, "mu0": 2.0e-3, "p0": 86100.0, "t0": 1.0}
    with gl.Context(nc, P) as cx:
        cx.upload(W)
        s = cx.stats()
    Tz = 380.0 * (1 + 0.3 * np.cos(2 * c))
    mu = 2.0e-3 * (Tz / 300.0) ** 1.5 * (300.0 + 110.4) / (Tz + 110.4)
    expect = np.mean(mu * (np.cos(c) ** 2 + 0.5 * np.sin(c) ** 2))
    assert s.diss == pytest.approx(expect, rel=2e-3)

    nc = 32
    c = (np.arange(nc) + 0.5) * 2 * np.pi / nc
    y, x, _ = np.meshgrid(c, c, c, indexing="ij")
    eps, dlt, p0, r0, Mt = 0.07, 0.12, 86100.0, 1.3, 0.45
    zero = np.zeros_like(x)
    W = icm.prim_to_cons(zero, zero, zero, r0 * (1 + dlt * np.sin(y)), p0 * (1 + eps * np.sin(x)))
    P = {"gam": 1.4, "rho0": r0, "Mt": Mt, "Re": 30.0, "mu0": 1e-3, "p0": p0, "t0": 1.0}
    with gl.Context(nc, P) as cx:
        cx.upload(W)
        s = cx.stats()
    a1, a2 = (1 - dlt ** 2) ** -0.5, (1 - dlt ** 2) ** -1.5
    T0 = p0 / (R * r0)
    assert s.prms == pytest.approx(eps / np.sqrt(2) / (1.4 * Mt * Mt), rel=1e-11)
    assert s.trms == pytest.approx(T0 * np.sqrt((1 + eps ** 2 / 2) * a2 - a1 ** 2) / (0.4 * T0 * Mt * Mt), rel=1e-10)
    assert s.vrms == pytest.approx(np.sqrt(a2 - a1 ** 2) / r0 / (Mt * Mt / r0), rel=1e-10)
