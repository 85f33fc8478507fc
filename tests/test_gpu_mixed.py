"""Mixed-precision path (GKM_MIXED) against the FP64 oracle and the FP64 GPU path.

Stated tolerances:
  * flux():   normwise relative error <= 1e-6 against the FP64 oracle
  * WENO:     normwise relative error <= 1e-6 against the FP64 oracle
  * one step: conserved fields <= 1e-7 normwise relative, step increments
              <= 1e-4 normwise relative (teacher-forced, same input). The
              FP32 slope terms carry ~3e-7 of the increment, and one CFL-0.5
              step changes the fields by a few % at 32^3, so the fields sit
              near 1e-8; 1e-7 leaves a factor ~5.
  * KE decay: <rho |u|^2> over a free run within 1e-4 relative of the FP64
              GPU run of the same seed
"""
import numpy as np
import pytest

import pyoracle as po
from cases import make_case, random_faces, random_stencils, rel_err
from dns_gkm_lpt_amd import _lib as gl

pytestmark = pytest.mark.gpu


def test_mixed_flux_batch():
    W, P, gas = make_case(32)
    X = np.concatenate([random_faces(256, seed=3, jump=0.05), random_faces(256, seed=4, jump=0.3)])
    tau = np.array([po.tau_of(X[i, :5], gas) for i in range(X.shape[0])])
    dt = 2e-4
    ref = po.flux(X, tau, dt, gas)
    with gl.Context(32, P) as c:
        F = c.flux_batch(X, tau, dt, precision=gl.GKM_MIXED)
    err = np.max(np.abs(F - ref), axis=0) / np.max(np.abs(ref), axis=0)
    print("mixed flux normwise rel err", err, flush=True)
    assert np.all(err < 1e-6), err


def test_mixed_weno_batch():
    W, P, gas = make_case(32)
    S = random_stencils(512, seed=5, jump=0.1)
    ref = po.weno(S, gas)
    with gl.Context(32, P) as c:
        out = c.weno_batch(S, precision=gl.GKM_MIXED)
    err = np.max(np.abs(out - ref), axis=0) / np.max(np.abs(ref), axis=0)
    print("mixed weno normwise rel err", err, flush=True)
    assert np.all(err < 1e-6), err


def test_mixed_one_step():
    nc = 32
    W, P, gas = make_case(nc, seed=0)
    dt = po.cfl_dt(W, 0.5, gas)
    ref = po.eulerian_step(W, dt, gas, nthreads=8)
    with gl.Context(nc, P, precision=gl.GKM_MIXED) as c:
        c.upload(W)
        c.step(dt)
        out = c.download()
    e = rel_err(out, ref)
    inc = rel_err(out - W, ref - W)
    print("mixed step rel err", e, "increment rel err", inc, flush=True)
    assert np.all(e < 1e-7), e
    assert np.all(inc < 1e-4), inc


def test_mixed_ke_decay_tracks_fp64():
    nc, nsteps, every = 64, 60, 10
    W, P, gas = make_case(nc, seed=1, Mt=0.5, Re=100.0, perturb=0.0)
    curves = {}
    for prec in (gl.GKM_FP64, gl.GKM_MIXED):
        ke = []
        with gl.Context(nc, P, cfl=0.5, precision=prec) as c:
            c.upload(W)
            ke.append(c.stats().tke)
            for _ in range(nsteps // every):
                c.run(every)
                ke.append(c.stats().tke)
        curves[prec] = np.array(ke)
    rel = np.abs(curves[gl.GKM_MIXED] - curves[gl.GKM_FP64]) / curves[gl.GKM_FP64]
    decay = 1 - curves[gl.GKM_FP64][-1] / curves[gl.GKM_FP64][0]
    print("KE decay over run", decay, "max rel KE diff mixed vs fp64", rel.max(), flush=True)
    assert rel.max() < 1e-4, rel
