"""The 32^3 CPU-runnable configuration run to the end on the GPU beside the oracle.

32^3 periodic cube, k0 4, Mt 0.3, Re_lambda 30, seed 0 (uniform density and
pressure, as the reference's initial condition), CFL 0.5 with dt from the
device CFL every step, 100 coupled steps (Eulerian step + Heun tracer step,
Evolve order, update.cpp:124-170), 4096 seeded uniform tracers.

Stated tolerances, FP64 path against the oracle over the whole run:
  * dt history: 1e-10 relative
  * KE, dissipation, S_ij S_ij, velocity-derivative skewness every 10 steps: 1e-9 relative
  * final conserved fields: 1e-10 normwise relative
  * final tracer positions: 1e-9 (minimum-image distance on the 2 pi torus)
Mixed path against the same oracle run (errors are printed):
  * KE and dissipation curves: 1e-4 relative
  * skewness curves: 1e-3 relative
  * final fields: 1e-5 normwise relative; tracer positions: 1e-5
"""
import numpy as np
import pytest

import pyoracle as po
from cases import rel_err, torus_dist
from dns_gkm_lpt_amd import _lib as gl
from dns_gkm_lpt_amd import ic as icm

pytestmark = pytest.mark.gpu

NC, NSTEPS, EVERY, NPART = 32, 100, 10, 4096


@pytest.fixture(scope="module")
def oracle_run():
    ic = icm.make_ic(NC, k0=4.0, Mt=0.3, Re=30.0, seed=0)
    P = ic["params"]
    W0 = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"], gam=P["gam"])
    gas = po.make_gas(NC, gam=P["gam"], mu0=P["mu0"], p0=P["p0"], den0=P["rho0"])
    x0 = np.ascontiguousarray(icm.uniform_particles(NPART, 0)[:, [1, 0, 2]])
    W, x = W0.copy(), x0.copy()
    dts, curves = [], []
    for n in range(1, NSTEPS + 1):
        dt = po.cfl_dt(W, 0.5, gas)
        dts.append(dt)
        U, _, _, _ = po.particle_sample(po.velocity_coefs(W), gas.dx, x)
        W = po.eulerian_step(W, dt, gas, nthreads=8)
        x = po.particle_heun(po.velocity_coefs(W), gas.dx, x, U, dt)
        if n % EVERY == 0:
            s = po.field_stats(W, po.velocity_coefs(W), gas, P["p0"], P["Mt"], P["rho0"])
            curves.append([s.tke, s.diss, s.sijsij] + list(s.skew))
    return {"P": P, "W0": W0, "x0": x0, "dts": np.array(dts), "curves": np.array(curves), "W": W, "x": x}


def gpu_run(ref, precision):
    dts, curves = [], []
    with gl.Context(NC, ref["P"], cfl=0.5, precision=precision) as c:
        c.upload(ref["W0"])
        c.lpt_upload(ref["x0"])
        for n in range(1, NSTEPS + 1):
            dt = c.cfl_dt()
            dts.append(dt)
            c.lpt_coupled_step(dt)
            if n % EVERY == 0:
                s = c.stats()
                curves.append([s.tke, s.diss, s.sijsij] + list(s.skew))
        W = c.download()
        x, _ = c.lpt_download()
    return np.array(dts), np.array(curves), W, x


def curve_errors(curves, ref_curves):
    return np.max(np.abs(curves - ref_curves) / np.abs(ref_curves), axis=0)


def test_fp64_run_tracks_oracle(oracle_run):
    dts, curves, W, x = gpu_run(oracle_run, gl.GKM_FP64)
    e_dt = np.max(np.abs(dts - oracle_run["dts"]) / oracle_run["dts"])
    e_curves = curve_errors(curves, oracle_run["curves"])
    e_W = rel_err(W, oracle_run["W"])
    e_x = float(np.max(torus_dist(x, oracle_run["x"])))
    ke = oracle_run["curves"][:, 0]
    print(f"fp64 100 steps: KE decay {1 - ke[-1] / ke[0]:.4f}; dt err {e_dt:.2e}; curve errs "
          f"(tke, diss, sijsij, skew x3) {e_curves}; field err {e_W}; particle err {e_x:.2e}", flush=True)
    assert e_dt < 1e-10
    assert np.all(e_curves < 1e-9), e_curves
    assert np.all(e_W < 1e-10), e_W
    assert e_x < 1e-9


def test_mixed_run_errors_against_oracle(oracle_run):
    dts, curves, W, x = gpu_run(oracle_run, gl.GKM_MIXED)
    e_curves = curve_errors(curves, oracle_run["curves"])
    e_W = rel_err(W, oracle_run["W"])
    e_x = float(np.max(torus_dist(x, oracle_run["x"])))
    e_dt = np.max(np.abs(dts - oracle_run["dts"]) / oracle_run["dts"])
    print(f"mixed 100 steps: dt err {e_dt:.2e}; curve errs (tke, diss, sijsij, skew x3) {e_curves}; "
          f"field err {e_W}; particle err {e_x:.2e}", flush=True)
    assert np.all(e_curves[:3] < 1e-4), e_curves
    assert np.all(e_curves[3:] < 1e-3), e_curves
    assert np.all(e_W < 1e-5), e_W
    assert e_x < 1e-5
