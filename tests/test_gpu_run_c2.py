"""The 128^3 GPU configuration run to the end, checked against the oracle step
by step from identical input at points along the run.

128^3 periodic cube, k0 4, Mt 0.3, Re_lambda 50, seed 0 (uniform density and
pressure, the reference's initial condition), CFL 0.5 with dt from the device
CFL every step, 500 coupled steps (Eulerian step + Heun tracer step, Evolve
order, update.cpp:124-170), 2^18 seeded uniform tracers, FP64.  The oracle
cannot run 500 steps of 128^3 inside a test, so the comparison is teacher
forced: at steps 0, 1, 250 and 499 the GPU's state and tracers are the input,
and the oracle's step from that input (CFL, Eulerian step, tracer sample at
x^n, spline of W^{n+1}, Heun) is compared with the GPU's own next step.

Stated tolerances: dt 1e-12 relative; conserved fields 1e-12 normwise
relative per variable; tracer positions 1e-11 (minimum-image distance on the
2 pi torus).  Over the whole run: the interior sums of mass, momentum and
energy stay constant to 1e-11 of the sum of magnitudes (periodic telescoping
of the flux differences), and the kinetic energy decays monotonically.
"""
import time

import numpy as np
import pytest

import pyoracle as po
from cases import rel_err, torus_dist
from dns_gkm_lpt_amd import _lib as gl
from dns_gkm_lpt_amd import ic as icm

pytestmark = pytest.mark.gpu

NC, NSTEPS, NPART = 128, 500, 1 << 18
CHECK = (0, 1, 250, 499)


def interior_sums(W):
    inner = (slice(None),) + (slice(3, NC + 3),) * 3
    return W[inner].sum(axis=(1, 2, 3)), np.abs(W[inner]).sum(axis=(1, 2, 3))


def test_128_run_matches_oracle_steps():
    ic = icm.make_ic(NC, k0=4.0, Mt=0.3, Re=50.0, seed=0)
    P = ic["params"]
    W0 = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"], gam=P["gam"])
    gas = po.make_gas(NC, gam=P["gam"], mu0=P["mu0"], p0=P["p0"], den0=P["rho0"])
    x0 = np.ascontiguousarray(icm.uniform_particles(NPART, 0)[:, [1, 0, 2]])
    snaps, tke = {}, []
    t0 = time.perf_counter()
    with gl.Context(NC, P, cfl=0.5, precision=gl.GKM_FP64) as c:
        c.upload(W0)
        c.lpt_upload(x0)
        for n in range(NSTEPS):
            if n in CHECK:
                snaps[n] = {"W": c.download(), "x": c.lpt_download()[0]}
            dt = c.cfl_dt()
            c.lpt_coupled_step(dt)
            if n in CHECK:
                snaps[n].update(dt=dt, W1=c.download(), x1=c.lpt_download()[0])
            if (n + 1) % 50 == 0:
                tke.append(c.stats().tke)
                print(f"step {n + 1}: tke {tke[-1]:.10e} dt {dt:.6e} ({time.perf_counter() - t0:.1f} s)", flush=True)
        Wend = c.download()
    s0, m0 = interior_sums(W0)
    s1, _ = interior_sums(Wend)
    drift = np.abs(s1 - s0) / m0
    print("conservation drift over the run", drift, flush=True)
    assert np.all(drift < 1e-11), drift
    assert np.all(np.diff(tke) < 0), tke
    for n, s in snaps.items():
        t = time.perf_counter()
        dto = po.cfl_dt(s["W"], 0.5, gas)
        U, _, _, _ = po.particle_sample(po.velocity_coefs(s["W"]), gas.dx, s["x"])
        Wo = po.eulerian_step(s["W"], s["dt"], gas, nthreads=16)
        xo = po.particle_heun(po.velocity_coefs(Wo), gas.dx, s["x"], U, s["dt"])
        e_dt = abs(s["dt"] - dto) / dto
        e_w = rel_err(s["W1"], Wo)
        e_x = np.max(torus_dist(s["x1"], xo))
        print(f"step {n}: dt {e_dt:.2e} fields {np.max(e_w):.2e} tracers {e_x:.2e} "
              f"(oracle {time.perf_counter() - t:.1f} s)", flush=True)
        assert e_dt <= 1e-12
        assert np.all(e_w < 1e-12), e_w
        assert e_x < 1e-11
