"""Edge cases of the GPU path: empty and single-tracer sets, tracers exactly on
the faces of the periodic box, positions outside it, a one-member ensemble, and
calls made in the wrong state.

Tolerances as in the parity tests: samples 1e-11 of the field's largest
magnitude, positions 1e-11 (minimum-image distance on the 2 pi torus); a
field advanced with an empty tracer set equals the plain step bit for bit.
"""
import numpy as np
import pytest

import pyoracle as po
from cases import make_case, torus_dist
from dns_gkm_lpt_amd import _lib as gl
from dns_gkm_lpt_amd import ic as icm

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def test_empty_tracer_set():
    W, P, gas = make_case(32, seed=40)
    with gl.Context(32, P, cfl=0.5) as c, gl.Context(32, P, cfl=0.5) as ref:
        c.upload(W)
        ref.upload(W)
        c.lpt_upload(np.zeros((0, 3)))
        for _ in range(2):
            dt = ref.cfl_dt()
            c.lpt_coupled_step(dt)
            ref.step(dt)
        x, v = c.lpt_download()
        assert x.shape == (0, 3) and v.shape == (0, 3)
        assert c.lpt_properties().shape == (0, gl.PROP_WIDTH)
        assert np.array_equal(c.download(), ref.download())


@pytest.mark.parametrize("n", [1, 257])
def test_ragged_tracer_counts_match_oracle(n):
    nc = 32
    W, P, gas = make_case(nc, seed=41)
    pos = np.ascontiguousarray(icm.uniform_particles(n, seed=41)[:, [1, 0, 2]])
    dt = po.cfl_dt(W, 0.5, gas)
    U, _, _, _ = po.particle_sample(po.velocity_coefs(W), gas.dx, pos)
    Wr = po.eulerian_step(W, dt, gas, nthreads=8)
    xr = po.particle_heun(po.velocity_coefs(Wr), gas.dx, pos, U, dt)
    with gl.Context(nc, P) as c:
        c.upload(W)
        c.lpt_upload(pos)
        c.lpt_coupled_step(dt)
        x, _ = c.lpt_download()
    assert x.shape == (n, 3)
    assert np.max(torus_dist(x, xr)) < 1e-11


def test_tracers_on_the_box_faces():
    """Coordinates 0 and 2 pi are both accepted (Part_Gen.m draws from [0, 2 pi]);
    samples there match the oracle's spline and a Heun step wraps into [0, 2 pi]."""
    nc = 32
    W, P, gas = make_case(nc, seed=42)
    edge = np.array([0.0, TWO_PI, np.nextafter(TWO_PI, 0.0), 1e-300])
    rng = np.random.default_rng(42)
    pos = rng.choice(edge, size=(64, 3))
    pos[:8] = rng.random((8, 3)) * TWO_PI
    pos[8] = (0.0, 0.0, 0.0)
    pos[9] = (TWO_PI, TWO_PI, TWO_PI)
    cs = po.velocity_coefs(W)
    U, gU, _, _ = po.particle_sample(cs, gas.dx, pos)
    dt = po.cfl_dt(W, 0.5, gas)
    xr = po.particle_heun(po.velocity_coefs(po.eulerian_step(W, dt, gas, nthreads=8)), gas.dx, pos, U, dt)
    with gl.Context(nc, P) as c:
        c.upload(W)
        c.lpt_upload(pos)
        U2, gU2, _, _ = c.lpt_sample()
        c.lpt_coupled_step(dt)
        x, _ = c.lpt_download()
    assert np.max(np.abs(U2 - U)) <= 1e-11 * np.max(np.abs(U))
    assert np.max(np.abs(gU2 - gU)) <= 1e-11 * np.max(np.abs(gU))
    assert np.all((x >= 0.0) & (x <= TWO_PI))
    assert np.max(torus_dist(x, xr)) < 1e-11


@pytest.mark.parametrize("bad", [-1e-12, TWO_PI * (1 + 1e-15), np.nan, np.inf])
def test_positions_outside_the_box_rejected(bad):
    W, P, gas = make_case(32, seed=43)
    pos = icm.uniform_particles(5, seed=43)
    pos[3, 1] = bad
    with gl.Context(32, P) as c:
        c.upload(W)
        with pytest.raises(gl.GkmError) as ei:
            c.lpt_upload(pos)
        assert ei.value.code == 1


def test_one_member_ensemble_equals_single_run():
    W, P, gas = make_case(48, seed=44)
    with gl.Ensemble([48], [P], cfl=0.5, precision=gl.GKM_FP64) as e:
        e.set_tiles((0, 0, 0))
        e.members[0].upload(W)
        e.run(3, graph=True)
        got = e.members[0].download()
    with gl.Context(48, P, cfl=0.5, precision=gl.GKM_FP64) as c:
        c.set_tiles((0, 0, 0))
        c.upload(W)
        c.run(3)
        ref = c.download()
    assert np.array_equal(got, ref)


def test_calls_in_the_wrong_state_fail_cleanly():
    W, P, gas = make_case(32, seed=45)
    with gl.Context(32, P) as c:
        with pytest.raises(gl.GkmError) as ei:
            c.cfl_dt()  # no state uploaded
        assert ei.value.code == 5
        with pytest.raises(gl.GkmError):
            c.lpt_stats()  # no sample taken
    with gl.Context(32, P, precision=gl.GKM_MIXED) as c:
        c.upload(W)
        with pytest.raises(gl.GkmError) as ei:
            c.tl_run(1e-4, 1, W)  # the tangent-linear step is FP64 only
        assert ei.value.code == 1
    with gl.Ensemble([32, 32], [P, P]) as e:
        e.members[0].upload(W)
        with pytest.raises(gl.GkmError) as ei:
            e.run(1)  # member 1 has no state
        assert ei.value.code == 5
