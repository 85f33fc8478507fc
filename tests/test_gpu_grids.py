"""Grid sizes that are not multiples of the sweep tiles.

The reference reads any nc from its params file (FlowField.cpp:9-58); the
sweeps accept any nc in [16, 1024]: the last pencils and segments of a grid
that the tile does not divide reach past nc, and those lines and faces are
computed from wrapped cells and not written.  Checked against the FP64
oracle at 16^3, 24^3, 40^3 and 100^3 (tolerances as test_gpu_step.py):
  * per-direction flux differences (derivsX/Y/Z, update.cpp:214-503) from the
    same input within 1e-11 normwise, for every tile variant that fits;
  * one step within 1e-12 normwise (increments 1e-9), a 4-step run with dt
    from the device CFL within 1e-11 (dt 1e-12);
  * the mixed-precision step within its one-step tolerance (1e-7);
  * a coupled particle step at 24^3 (tracer positions 1e-11);
  * an ensemble of 24^3 and 40^3 members equal to their single runs bit for
    bit;
  * nc = 15 and 1025 refused, and a tile larger than the grid refused.
"""
import numpy as np
import pytest

import pyoracle as po
from cases import make_case, rel_err, torus_dist
from dns_gkm_lpt_amd import _lib as gl
from dns_gkm_lpt_amd import ic as icm

pytestmark = pytest.mark.gpu

GRIDS = [16, 24, 40, 100]


def _flux_diff_ref(W, dt, gas, direction):
    F = po.fluxes(W, dt, gas, nthreads=8)[direction]
    axis = [2, 1, 3][direction]  # axis of (q, y, x, z)
    return (F - np.roll(F, 1, axis=axis))[:, 3:-3, 3:-3, 3:-3]


@pytest.mark.parametrize("nc", GRIDS)
def test_flux_differences_every_fitting_tile(nc):
    W, P, gas = make_case(nc, seed=nc)
    dt = po.cfl_dt(W, 0.5, gas)
    refs = [_flux_diff_ref(W, dt, gas, d) for d in range(3)]
    fits = 0
    for v in range(3):
        for d in range(3):
            with gl.Context(nc, P) as c:
                base = list(c.tiles())
                tiles = base.copy()
                tiles[d] = v
                try:
                    c.set_tiles(tuple(tiles))
                except gl.GkmError as e:
                    assert e.code == 1  # the variant does not fit this grid
                    continue
                c.upload(W)
                out = c.sweep_flux_diff(d, dt)
            fits += 1
            e = rel_err(out, refs[d])
            print(f"nc {nc} direction {d} tile {v}: rel err {e.max():.2e}", flush=True)
            assert np.all(e < 1e-11), (nc, d, v, e)
    assert fits >= 3


@pytest.mark.parametrize("nc", GRIDS)
def test_steps_match_oracle(nc):
    W, P, gas = make_case(nc, seed=nc + 1)
    dt = po.cfl_dt(W, 0.5, gas)
    ref = po.eulerian_step(W, dt, gas, nthreads=8)
    with gl.Context(nc, P) as c:
        c.upload(W)
        c.step(dt)
        out = c.download()
    e = rel_err(out, ref)
    inc = rel_err(out - W, ref - W)
    print(f"nc {nc}: one step {e.max():.2e}, increments {inc.max():.2e}", flush=True)
    assert np.all(e < 1e-12), e
    assert np.all(inc < 1e-9), inc
    with gl.Context(nc, P, precision=gl.GKM_MIXED) as c:
        c.upload(W)
        c.step(dt)
        em = rel_err(c.download(), ref)
    print(f"nc {nc}: mixed one step {em.max():.2e}", flush=True)
    assert np.all(em < 1e-7), em
    # a short run with the device CFL
    Wr, dts = W.copy(), []
    for _ in range(4):
        d = po.cfl_dt(Wr, 0.5, gas)
        dts.append(d)
        Wr = po.eulerian_step(Wr, d, gas, nthreads=8)
    with gl.Context(nc, P, cfl=0.5) as c:
        c.upload(W)
        hist = c.run(4, record_dt=True)
        out = c.download()
    assert np.allclose(hist, dts, rtol=1e-12, atol=0)
    assert np.all(rel_err(out, Wr) < 1e-11)


def test_coupled_particle_step_at_24():
    nc = 24
    W, P, gas = make_case(nc, seed=5)
    pos = np.ascontiguousarray(icm.uniform_particles(1000, 5)[:, [1, 0, 2]])
    dt = po.cfl_dt(W, 0.5, gas)
    U, _, _, _ = po.particle_sample(po.velocity_coefs(W), gas.dx, pos)
    Wr = po.eulerian_step(W, dt, gas, nthreads=8)
    xr = po.particle_heun(po.velocity_coefs(Wr), gas.dx, pos, U, dt)
    with gl.Context(nc, P) as c:
        c.upload(W)
        c.lpt_upload(pos)
        c.lpt_coupled_step(dt)
        x, _ = c.lpt_download()
        out = c.download()
    assert np.all(rel_err(out, Wr) < 1e-12)
    d = float(np.max(torus_dist(x, xr)))
    print("24^3 coupled step: tracer error", d, flush=True)
    assert d < 1e-11


def test_ensemble_of_uneven_grids_equals_single_runs():
    cases = [(24, 11), (40, 12)]
    states = [make_case(nc, seed=s) for nc, s in cases]
    e = gl.Ensemble([nc for nc, _ in cases], [st[1] for st in states], cfl=0.5, precision=gl.GKM_MIXED)
    try:
        for k, st in enumerate(states):
            e.members[k].upload(st[0])
        e.set_tiles((0, 0, 0))
        e.run(3, graph=False)
        outs = [m.download() for m in e.members]
    finally:
        e.close()
    for (nc, _), st, out in zip(cases, states, outs):
        with gl.Context(nc, st[1], cfl=0.5, precision=gl.GKM_MIXED) as c:
            c.set_tiles((0, 0, 0))
            c.upload(st[0])
            c.run(3)
            ref = c.download()
        print(f"ensemble member {nc}^3 vs single run: max |diff| {np.max(np.abs(out - ref)):.3e}", flush=True)
        assert np.array_equal(out, ref)


def test_grid_limits():
    _, P, _ = make_case(16)
    for bad in (15, 1025):
        with pytest.raises(gl.GkmError) as ei:
            gl.Context(bad, P)
        assert ei.value.code == 1 and "16, 1024" in str(ei.value)
    with gl.Context(16, P) as c:
        with pytest.raises(gl.GkmError) as ei:
            c.set_tiles((0, 0, 0))  # the z variant 16 x 4 x 4 needs nc >= 19
        assert ei.value.code == 1
