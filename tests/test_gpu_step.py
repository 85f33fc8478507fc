"""Parity of the gfx950 GKM-WENO step with the FP64 oracle (through the C ABI).

Tolerances (FP64 path): per-step conserved fields within 1e-12 normwise
relative (max |dW| / max |W_ref| per variable) from identical input; the
step increments W^{n+1} - W^n within 1e-9 normwise relative, which is the
stricter check since the increment is ~1e-3 of W.
"""
import numpy as np
import pytest

import pyoracle as po
from cases import make_case, random_faces, random_stencils, rel_err
from dns_gkm_lpt_amd import _lib as gl

pytestmark = pytest.mark.gpu


def ctx_for(nc, P, **kw):
    return gl.Context(nc, P, **kw)


def test_flux_batch_matches_oracle():
    W, P, gas = make_case(32)
    X = random_faces(512, seed=3)
    tau = np.array([po.tau_of(X[i, :5], gas) for i in range(X.shape[0])])
    dt = 2e-4
    ref = po.flux(X, tau, dt, gas)
    with ctx_for(32, P) as c:
        F = c.flux_batch(X, tau, dt)
    err = np.max(np.abs(F - ref), axis=0) / np.max(np.abs(ref), axis=0)
    assert np.all(err < 1e-12), err


def test_weno_batch_matches_oracle():
    W, P, gas = make_case(32)
    S = random_stencils(512, seed=5)
    ref = po.weno(S, gas)
    with ctx_for(32, P) as c:
        out = c.weno_batch(S)
    err = np.max(np.abs(out - ref), axis=0) / np.max(np.abs(ref), axis=0)
    assert np.all(err < 1e-12), err


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_sweep_flux_differences_match_oracle(direction):
    """Teacher-forced per-direction check of derivsX/Y/Z: F(i+1/2) - F(i-1/2)."""
    nc = 32
    W, P, gas = make_case(nc, seed=0)
    dt = po.cfl_dt(W, 0.5, gas)
    F = po.fluxes(W, dt, gas, nthreads=8)[direction]
    shift = [(0, 0, 1, 0), (0, 1, 0, 0), (0, 0, 0, 1)][direction]  # axis of (q, y, x, z)
    Fm = np.roll(F, 1, axis=shift.index(1))
    ref = (F - Fm)[:, 3:-3, 3:-3, 3:-3]
    with ctx_for(nc, P) as c:
        c.upload(W)
        out = c.sweep_flux_diff(direction, dt)
    e = rel_err(out, ref)
    print("direction", direction, "rel err", e, flush=True)
    assert np.all(e < 1e-11), e


@pytest.mark.parametrize("nc,seed", [(32, 0), (48, 1)])
def test_one_step_matches_oracle(nc, seed):
    W, P, gas = make_case(nc, seed=seed)
    dt = po.cfl_dt(W, 0.5, gas)
    ref = po.eulerian_step(W, dt, gas, nthreads=8)
    with ctx_for(nc, P) as c:
        c.upload(W)
        c.step(dt)
        out = c.download()
    e = rel_err(out, ref)
    assert np.all(e < 1e-12), e
    inc = rel_err(out - W, ref - W)
    assert np.all(inc < 1e-9), inc


def test_cfl_matches_oracle():
    W, P, gas = make_case(32, seed=2)
    ref = po.cfl_dt(W, 0.5, gas)
    with ctx_for(32, P, cfl=0.5) as c:
        c.upload(W)
        dt = c.cfl_dt()
    assert abs(dt - ref) <= 1e-14 * ref


def test_run_matches_oracle_trajectory():
    """Free-running 6 steps with device-side CFL dt vs the oracle's loop."""
    nc, nsteps = 32, 6
    W, P, gas = make_case(nc, seed=4)
    Wr = W.copy()
    dts = []
    for _ in range(nsteps):
        dt = po.cfl_dt(Wr, 0.5, gas)
        dts.append(dt)
        Wr = po.eulerian_step(Wr, dt, gas, nthreads=8)
    with ctx_for(nc, P, cfl=0.5) as c:
        c.upload(W)
        hist = c.run(nsteps, record_dt=True)
        out = c.download()
    assert np.allclose(hist, dts, rtol=1e-12, atol=0)
    e = rel_err(out, Wr)
    assert np.all(e < 1e-11), e


def test_download_fills_periodic_ghosts():
    W, P, gas = make_case(32, seed=6)
    with ctx_for(32, P) as c:
        c.upload(W)
        out = c.download()
    assert np.array_equal(out, W)


def test_conservation_to_roundoff():
    W, P, gas = make_case(32, seed=7)
    inner = (slice(None),) + (slice(3, -3),) * 3
    with ctx_for(32, P) as c:
        c.upload(W)
        c.run(3)
        out = c.download()
    s0 = W[inner].sum(axis=(1, 2, 3))
    s1 = out[inner].sum(axis=(1, 2, 3))
    scale = np.abs(W[inner]).sum(axis=(1, 2, 3))
    assert np.all(np.abs(s1 - s0) <= 1e-13 * scale), (s1 - s0) / scale


def test_invalid_config_rejected():
    W, P, gas = make_case(32)
    for bad in (8, 15, 1025, 2048):  # any nc in [16, 1024] is accepted
        with pytest.raises(gl.GkmError) as ei:
            gl.Context(bad, P)
        assert ei.value.code == 1
    with pytest.raises(gl.GkmError) as ei:
        gl.Context(32, dict(P, gam=1.0))
    assert ei.value.code == 1


def test_nonphysical_state_reported():
    W, P, gas = make_case(32, seed=8)
    W = W.copy()
    W[0, 10, 10, 10] = -1.0
    with ctx_for(32, P) as c:
        c.upload(W)
        with pytest.raises(gl.GkmError) as ei:
            c.cfl_dt()
    assert ei.value.code == 4


@pytest.mark.parametrize("nc", [32, 48, 64])
def test_tile_variants_agree(nc):
    """Every sweep tile shape evaluates the same per-face formulas; each
    instantiation is compiled on its own (the compiler may fuse multiply-adds
    differently), so the shapes agree to rounding: 1e-14 normwise per
    variable, and each matches the oracle to 1e-12."""
    W, P, gas = make_case(nc, seed=9)
    dt = po.cfl_dt(W, 0.5, gas)
    ref = po.eulerian_step(W, dt, gas, nthreads=8)
    outs = []
    for tiles in ((0, 0, 0), (1, 1, 1), (2, 2, 2), (1, 2, 0)):
        with ctx_for(nc, P) as c:
            c.set_tiles(tiles)
            assert c.tiles() == tiles
            c.upload(W)
            c.step(dt)
            outs.append(c.download())
        e = rel_err(outs[-1], ref)
        print("tiles", tiles, "rel err", e, flush=True)
        assert np.all(e < 1e-12), (tiles, e)
    for o in outs[1:]:
        d = rel_err(o, outs[0])
        print("max difference between tile variants", d, flush=True)
        assert np.all(d < 1e-14), d


def test_autotune_keeps_state_and_picks_valid_tiles():
    nc = 48
    W, P, gas = make_case(nc, seed=10)
    dt = po.cfl_dt(W, 0.5, gas)
    ref = po.eulerian_step(W, dt, gas, nthreads=8)
    with ctx_for(nc, P) as c:
        c.upload(W)
        tiles, ms = c.autotune(reps=1)
        print("autotune", tiles, ms.tolist(), flush=True)
        assert all(0 <= t < 3 for t in tiles) and c.tiles() == tiles
        assert np.all(ms > 0)
        assert np.array_equal(c.download(), W)
        c.step(dt)
        out = c.download()
        with pytest.raises(gl.GkmError) as ei:
            c.set_tiles((3, 0, 0))
        assert ei.value.code == 1
    assert np.all(rel_err(out, ref) < 1e-12)
