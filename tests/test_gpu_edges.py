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


@pytest.mark.parametrize("drop", [False, True])
def test_reference_drop_of_tracers_at_two_pi(drop):
    """lpt_set_reference_drop: in a uniform flow (u = -1e-15, v = 3, w = 0) a
    tracer at x = 0 moves to x = -u dt ~ -1e-18, which the wrap maps to 2 pi
    exactly (2 pi - 1e-18 rounds to 2 pi): block id int(2 pi / 2 pi) = 1, so
    the reference's single-rank run pops it (update.cpp:1056-1109); a tracer
    uploaded at x = 2 pi is not located on rank 0 (update.cpp:1304-1316).
    Without the mode both are kept on the face x = 2 pi."""
    nc = 16
    z = np.zeros((nc,) * 3)
    u, v = -1e-15, 3.0
    W = icm.prim_to_cons(z + u, z + v, z, z + 1.1, z + 86100.0)
    P = {"gam": 1.4, "rho0": 1.0, "Mt": 0.3, "Re": 30.0, "mu0": 1e-3, "p0": 86100.0, "t0": 1.0}
    pos = np.array([[1.0, 0.0, 1.0],      # x -> 2 pi after one step: dropped
                    [1.0, 1.0, 1.0],      # kept
                    [1.0, TWO_PI, 1.0],   # at 2 pi on upload: dropped at upload
                    [0.0, 3.0, 0.0]])     # y = 0 moves to 3 dt: kept
    dt = 1e-3
    with gl.Context(nc, P) as c:
        c.upload(W)
        if drop:
            c.lpt_set_reference_drop(True)
        c.lpt_upload(pos)
        assert c.lpt_count() == (3 if drop else 4)
        c.lpt_coupled_step(dt)
        x, _ = c.lpt_download()
        assert c.lpt_count() == (2 if drop else 4)
        U, gU, gV, gW = c.lpt_sample()
        s = c.lpt_stats()
        props = c.lpt_properties()
        out = c.download()
    assert np.array_equal(out, W)  # a uniform state does not change
    expect_y = pos[:, 0] + v * dt
    for k in (1, 3):
        assert x[k, 0] == pytest.approx(expect_y[k], abs=1e-14)
        assert x[k, 1] == pytest.approx(pos[k, 1], abs=1e-14)
        assert np.allclose(U[k], [u, v, 0.0], atol=1e-12)
    if drop:
        assert np.all(np.isnan(x[[0, 2]]))
        assert np.all(U[[0, 2]] == 0.0) and np.all(gU[[0, 2]] == 0.0)
        assert np.all(np.isnan(props[[0, 2]])) and np.all(np.isfinite(props[[1, 3]]))
        # GetParticleStats sums over the tracers a rank holds, divides by numparticles
        assert s.tke == pytest.approx(2 * (u * u + v * v) / 4, rel=1e-12)
    else:
        assert x[0, 1] == TWO_PI and x[2, 1] == TWO_PI
        assert np.all(np.isfinite(props))
        assert s.tke == pytest.approx(u * u + v * v, rel=1e-12)
