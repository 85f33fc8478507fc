"""The batched-ensemble workload at its full size: 64 independent members
(24 x 32^3 + 24 x 48^3 + 16 x 64^3, each with its own seed, k0, Mt and
Re_lambda; dns_gkm_lpt_amd.ic.ensemble_members, the members bench.py times),
mixed precision, 8192 tracers each, 200 graph-replayed steps with each
member's own CFL dt.

Checked:
  * every member conserves mass, momentum and energy over the 200 steps
    (periodic telescoping fluxes): |sum W_N - sum W_0| <= 1e-12 sum |W_0|
    per conserved variable;
  * every member's kinetic energy <rho |u|^2 / 2> decays from each block of
    20 steps to the next (decaying turbulence, no forcing);
  * one member of each grid size equals its single-context run -- the same
    tiles, lpt_coupled_step with the context's own CFL dt each step -- bit for
    bit, fields and tracers (the single-context path is the one the oracle
    pins);
  * the persistent-kernel step (one launch per step, graph-replayed) equals
    the per-phase launches bit for bit over all 200 steps, every member,
    fields and tracers.
"""
import numpy as np
import pytest

from dns_gkm_lpt_amd import _lib as gl
from dns_gkm_lpt_amd import ic as icm

pytestmark = pytest.mark.gpu

NTR = 8192
NSTEPS = 200
BLOCK = 20
TILES = (2, 0, 2)


@pytest.fixture(scope="module")
def members():
    return icm.ensemble_members()


def _ke(W):
    inner = W[:, 3:-3, 3:-3, 3:-3]
    return 0.5 * float(np.sum((inner[1] ** 2 + inner[2] ** 2 + inner[3] ** 2) / inner[0]))


def _sums(W):
    inner = W[:, 3:-3, 3:-3, 3:-3]
    return inner.reshape(5, -1).sum(axis=1), np.abs(inner).reshape(5, -1).sum(axis=1)


def _make(members, persistent):
    e = gl.Ensemble([m[0] for m in members], [m[1] for m in members], cfl=0.5, precision=gl.GKM_MIXED)
    for k, (nc, P, W, tseed) in enumerate(members):
        e.members[k].upload(W)
        e.members[k].lpt_upload(icm.uniform_particles(NTR, seed=tseed))
    e.set_tiles(TILES)
    if persistent:
        e.set_persistent(True)
    return e


def test_full_ensemble_200_steps(members):
    n = len(members)
    assert n == 64 and sum(m[0] == 32 for m in members) == 24 and sum(m[0] == 64 for m in members) == 16
    s0 = [_sums(m[2]) for m in members]
    ke = [[_ke(m[2])] for m in members]
    with _make(members, persistent=False) as e:
        for b in range(NSTEPS // BLOCK):
            e.run(BLOCK, graph=True)
            for k in range(n):
                ke[k].append(_ke(e.members[k].download()))
            print(f"c4: per-phase graph, {BLOCK * (b + 1)} steps", flush=True)
        fields = [e.members[k].download() for k in range(n)]
        tracers = [e.members[k].lpt_download()[0] for k in range(n)]
    for k in range(n):
        s, a = _sums(fields[k])
        drift = np.abs(s - s0[k][0]) / s0[k][1]
        assert np.all(drift <= 1e-12), (k, drift)
        d = np.diff(ke[k])
        assert np.all(d < 0), (k, members[k][0], ke[k])
        assert np.all(np.isfinite(tracers[k]))

    # one member of each grid size against its single-context run
    for k in (0, 24, 48):
        nc, P, W, tseed = members[k]
        print(f"c4: single-context run of member {k} ({nc}^3)", flush=True)
        with gl.Context(nc, P, cfl=0.5, precision=gl.GKM_MIXED) as c:
            c.set_tiles(TILES)
            c.upload(W)
            c.lpt_upload(icm.uniform_particles(NTR, seed=tseed))
            for _ in range(NSTEPS):
                c.lpt_coupled_step(c.cfl_dt())
            ref = c.download()
            xref, _ = c.lpt_download()
        assert np.array_equal(fields[k], ref), (k, np.max(np.abs(fields[k] - ref)))
        assert np.array_equal(tracers[k], xref), (k, np.max(np.abs(tracers[k] - xref)))

    # the persistent-kernel step, graph-replayed, over the same 200 steps
    with _make(members, persistent=True) as e:
        e.run(NSTEPS, graph=True)
        print("c4: persistent graph, 200 steps", flush=True)
        for k in range(n):
            assert np.array_equal(e.members[k].download(), fields[k]), k
            assert np.array_equal(e.members[k].lpt_download()[0], tracers[k]), k
