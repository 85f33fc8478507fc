"""Ensemble of independent runs (one launch per direction for all members,
hipGraph-captured steps, tracers of every member in one launch per phase).
Pins: every member equals its single run (the ensemble kernels run the same
bodies; measured bit for bit, asserted to 1e-13 since they are separate
instantiations), tracers equal the single context's coupled steps, graph
replays equal eager launches bit for bit, autotuning leaves the states
unchanged."""
import numpy as np
import pytest

from cases import make_case, rel_err, torus_dist
from dns_gkm_lpt_amd import _lib as gl
from dns_gkm_lpt_amd import ic as icm

pytestmark = pytest.mark.gpu

MEMBERS = [(32, 0, 0.3, 30.0), (48, 1, 0.5, 50.0), (64, 2, 0.2, 20.0), (32, 3, 0.6, 60.0)]


def cases():
    out = []
    for nc, seed, Mt, Re in MEMBERS:
        W, P, gas = make_case(nc, seed=seed, Mt=Mt, Re=Re)
        out.append((nc, W, P))
    return out


def make_ensemble(cs, precision):
    e = gl.Ensemble([c[0] for c in cs], [c[2] for c in cs], cfl=0.5, precision=precision)
    for m, (nc, W, P) in enumerate(cs):
        e.members[m].upload(W)
    return e


@pytest.mark.parametrize("precision,tol", [(gl.GKM_FP64, 1e-13), (gl.GKM_MIXED, 1e-12)])
def test_members_equal_single_runs(precision, tol):
    cs = cases()
    nsteps = 5
    with make_ensemble(cs, precision) as e:
        e.set_tiles((0, 0, 0))
        e.run(nsteps, graph=True)
        outs = [m.download() for m in e.members]
        dts = e.dt()
    for m, (nc, W, P) in enumerate(cs):
        with gl.Context(nc, P, cfl=0.5, precision=precision) as c:
            c.set_tiles((0, 0, 0))
            c.upload(W)
            hist = c.run(nsteps, record_dt=True)
            ref = c.download()
        err = rel_err(outs[m], ref)
        print("member", m, "nc", nc, "rel err vs single run", err, "dt", dts[m], hist[-1], flush=True)
        assert np.all(err < tol), (m, err)
        assert abs(dts[m] - hist[-1]) <= 1e-13 * hist[-1]


@pytest.mark.parametrize("precision", [gl.GKM_FP64, gl.GKM_MIXED])
def test_members_with_tracers_equal_single_coupled_runs(precision):
    """Member 1 holds no tracers; the others advance theirs with the field
    (sample at x^n, sweeps, spline of W^{n+1}, Heun), eager and replayed steps
    mixed, and a second upload of tracers between runs."""
    cs = cases()
    counts = [300, 0, 1000, 257]
    pos = [icm.uniform_particles(n, seed=10 + m) for m, n in enumerate(counts)]
    pos2 = icm.uniform_particles(123, seed=99)
    with make_ensemble(cs, precision) as e:
        e.set_tiles((0, 0, 0))
        for m, n in enumerate(counts):
            if n:
                e.members[m].lpt_upload(pos[m])
        e.run(2, graph=True)
        e.run(1, graph=False)
        e.members[3].lpt_upload(pos2)  # new buffers: the tables are rebuilt and the graphs recaptured
        e.run(2, graph=True)
        outs = [(e.members[m].download(), e.members[m].lpt_download() if counts[m] else None)
                for m in range(len(cs))]
        pst = e.members[0].lpt_stats()
    for m, (nc, W, P) in enumerate(cs):
        with gl.Context(nc, P, cfl=0.5, precision=precision) as c:
            c.set_tiles((0, 0, 0))
            c.upload(W)
            if counts[m]:
                c.lpt_upload(pos[m])
            for k in range(5):
                if k == 3 and m == 3:
                    c.lpt_upload(pos2)
                dt = c.cfl_dt()
                if counts[m]:
                    c.lpt_coupled_step(dt)
                else:
                    c.step(dt)
            ref = c.download()
            rpos = c.lpt_download() if counts[m] else None
            rst = c.lpt_stats() if m == 0 else None
        err = rel_err(outs[m][0], ref)
        assert np.all(err < 1e-13), (m, err)
        if counts[m]:
            dp = np.max(torus_dist(outs[m][1][0], rpos[0]))
            dv = np.max(np.abs(outs[m][1][1] - rpos[1])) / np.max(np.abs(rpos[1]))
            print("member", m, "tracers", len(rpos[0]), "max position diff", dp, "velocity", dv, flush=True)
            assert dp <= 1e-13 and dv <= 1e-13, (m, dp, dv)
        if m == 0:
            assert abs(pst.tke - rst.tke) <= 1e-13 * abs(rst.tke)


@pytest.mark.parametrize("precision", [gl.GKM_FP64, gl.GKM_MIXED])
def test_persistent_step_equals_per_phase_launches(precision):
    """One persistent-kernel launch per step (grid barriers between the CFL,
    sweep and tracer phases) gives the per-phase launches' results bit for bit:
    fields, tracer positions and velocities, dt; replayed graphs and eager
    launches, over several runs (the barrier counter carries across launches)."""
    cs = cases()
    counts = [300, 0, 1000, 257]
    pos = [icm.uniform_particles(n, seed=30 + m) for m, n in enumerate(counts)]

    def run(persistent):
        with make_ensemble(cs, precision) as e:
            e.set_tiles((2, 0, 1))
            for m, n in enumerate(counts):
                if n:
                    e.members[m].lpt_upload(pos[m])
            e.set_persistent(persistent)
            e.run(3, graph=True)
            e.run(2, graph=False)
            e.run(2, graph=True)
            return ([mb.download() for mb in e.members],
                    [e.members[m].lpt_download() if counts[m] else None for m in range(len(cs))], e.dt())

    ref, got = run(False), run(True)
    assert np.array_equal(ref[2], got[2])
    for m in range(len(cs)):
        assert np.array_equal(ref[0][m], got[0][m]), m
        if counts[m]:
            assert np.array_equal(ref[1][m][0], got[1][m][0]), m
            assert np.array_equal(ref[1][m][1], got[1][m][1]), m


def test_graph_replays_equal_eager_launches():
    cs = cases()
    with make_ensemble(cs, gl.GKM_MIXED) as eager, make_ensemble(cs, gl.GKM_MIXED) as graph:
        for rep in range(3):
            eager.run(2, graph=False)
            graph.run(2, graph=True)  # replays the two captured graphs (one per buffer parity)
            for m in range(len(cs)):
                a, b = eager.members[m].download(), graph.members[m].download()
                assert np.array_equal(a, b), (rep, m)
            assert np.array_equal(eager.dt(), graph.dt())


def test_ensemble_autotune_keeps_state():
    cs = cases()
    with make_ensemble(cs, gl.GKM_MIXED) as e:
        tiles, ms = e.autotune(reps=1)
        print("ensemble autotune", tiles, ms.tolist(), flush=True)
        assert all(0 <= t < 3 for t in tiles) and np.all(ms > 0)
        for m, (nc, W, P) in enumerate(cs):
            assert np.array_equal(e.members[m].download(), W)
        e.run(1)


def test_ensemble_rejects_mixed_precisions():
    cs = cases()[:2]
    cfgs = (gl.Config * 2)(gl.make_config(cs[0][0], cs[0][2], precision=gl.GKM_FP64),
                           gl.make_config(cs[1][0], cs[1][2], precision=gl.GKM_MIXED))
    import ctypes as C
    h = C.c_void_p()
    rc = gl.lib().gkm_ensemble_create(2, cfgs, C.byref(h))
    assert rc == 1
    assert b"precision" in gl.lib().gkm_ensemble_last_error(h)
    gl.lib().gkm_ensemble_destroy(h)
