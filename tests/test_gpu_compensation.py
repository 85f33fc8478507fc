"""Compensated accumulation of the mixed-precision update.

The mixed path keeps the state in FP64 and, per cell and variable, an FP32
carry of the rounding error its last update left out (on by default for
GKM_MIXED).  The update W^{n+1} = W^n - (dFx + dFy + dFz)/dx (update.cpp:505-516)
becomes t = W + y, e = (W + y) - t with y = increment + carry (TwoSum), so
that W + carry is the exact running sum of the rounded increments.

Checked:
  * one step from a zero carry: the increment y is rebuilt on the host from
    the three sweeps' flux differences (the same device kernels, one
    direction at a time), and for every cell the new state is y + W rounded
    once (t == W + y in IEEE double) and, in exact rational arithmetic,
    t + carry = W + y up to the FP32 rounding of the carry (|carry - e| <=
    2^-24 |e|); |carry| <= ulp(t)/2 everywhere;
  * conservation: over 100 steps of the 32^3 case the exact sums
    (math.fsum) of W + carry drift less than those of W with the plain
    update for every conserved variable, and the three momenta's drifts
    together at least 10x less (periodic fluxes telescope, so the drift is
    rounding only: the plain update's rounding scales with |W|, the
    compensated one's with the increments; both are random walks, so one
    component's ratio varies from run to run -- measured 4.7-190x -- and the
    bound is put on the sum; rho and rho E drift by 0 in both runs);
  * the compensated and the plain mixed runs agree to 1e-6 normwise (the
    carry moves the state by a rounding error per step, which the FP32 slope
    terms turn into FP32-rounding-sized changes of later fluxes, as any
    rounding change would; measured 0.8-3.1e-8 after 100 steps), and both
    stay within 1e-5 of the FP64 GPU run (the mixed path's tolerance).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from cases import make_case
from dns_gkm_lpt_amd import _lib as gl

pytestmark = pytest.mark.gpu

INNER = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))


def test_default_is_on_for_mixed_only():
    _, P, _ = make_case(32)
    with gl.Context(32, P, precision=gl.GKM_MIXED) as c:
        assert c.compensated
    with gl.Context(32, P, precision=gl.GKM_FP64) as c:
        assert not c.compensated


def test_carry_is_the_rounding_error_of_the_update():
    nc = 32
    W, P, gas = make_case(nc, seed=2)
    with gl.Context(nc, P, precision=gl.GKM_MIXED) as c:
        c.upload(W)
        dt = c.cfl_dt()
        dF = [c.sweep_flux_diff(d, dt) for d in range(3)]
        c.step(dt)
        W1 = c.download()[INNER]
        e32 = c.download_carry()
    W0 = W[INNER]
    inv_dx = 1.0 / gas.dx
    s = (dF[0] + dF[1]) + dF[2]
    y = -inv_dx * s
    assert np.array_equal(W1, W0 + y), "the compensated update rounds W + y once"
    ulp = np.spacing(np.abs(W1))
    assert np.all(np.abs(e32.astype(np.float64)) <= 0.5 * ulp)
    assert np.count_nonzero(e32) > 0.5 * e32.size
    rng = np.random.default_rng(0)
    flat = [a.reshape(-1) for a in (W0, y, W1, e32.astype(np.float64))]
    for i in rng.choice(flat[0].size, 3000, replace=False):
        w, yy, t, e = (float(a[i]) for a in flat)
        exact = Fraction(w) + Fraction(yy) - Fraction(t)
        assert abs(Fraction(e) - exact) <= abs(exact) * Fraction(1, 1 << 24), (i, w, yy, t, e)


def _sums(W, carry=None):
    out = []
    for q in range(5):
        vals = W[q].reshape(-1).tolist()
        if carry is not None:
            vals += carry[q].astype(np.float64).reshape(-1).tolist()
        out.append(math.fsum(vals))
    return np.array(out)


def _run(W, P, nsteps, compensate, precision=gl.GKM_MIXED):
    with gl.Context(32, P, precision=precision) as c:
        c.set_compensation(compensate)
        c.upload(W)
        c.run(nsteps)
        W1 = c.download()[INNER]
        carry = c.download_carry() if compensate else None
    return W1, carry


def test_compensation_conserves_to_the_increments_rounding():
    W, P, _ = make_case(32, seed=0)
    n = 100
    s0 = _sums(W[INNER])
    Wc, carry = _run(W, P, n, True)
    Wp, _ = _run(W, P, n, False)
    Wf, _ = _run(W, P, n, False, precision=gl.GKM_FP64)
    drift_c = np.abs(_sums(Wc, carry) - s0)
    drift_p = np.abs(_sums(Wp) - s0)
    scale = _sums(np.abs(W[INNER]))
    print("conservation drift / sum|W|: compensated", drift_c / scale, "plain", drift_p / scale, flush=True)
    assert np.all(drift_c <= drift_p), (drift_c, drift_p)
    assert np.sum(drift_c[1:4]) * 10 <= np.sum(drift_p[1:4]), (drift_c, drift_p)
    diff = np.array([np.max(np.abs(Wc[q] - Wp[q])) / np.max(np.abs(Wp[q])) for q in range(5)])
    print("compensated vs plain mixed, normwise", diff, flush=True)
    assert np.all(diff < 1e-6), diff
    for Wm in (Wc, Wp):
        e = np.array([np.max(np.abs(Wm[q] - Wf[q])) / np.max(np.abs(Wf[q])) for q in range(5)])
        assert np.all(e < 1e-5), e
