"""The drop-in driver run as the reference's Main.cpp would run it
(Main.cpp:10-49 -> Evolve, update.cpp:39-188), against the oracle.

A 32^3 case of the reference recipe (k0 4, Mt 0.3, Re_lambda 30, seed 0, 4096
seeded tracers) is written in the reference's input formats (params,
Init000, particleCount, unirand) and dns_gkm_lpt_amd/host/dnsgkm_run runs on
it as a fresh child process (DNSGKM_NSTEPS=21, DNSGKM_CFL=0.5): GetParams,
Initialize_DHIT, Evolve's loop with the statistics line and the particle
file every 10 iterations, the field and restart files at iteration 0, dt
from the device CFL after every step and t += dt (the reference's order).
This test process never touches the GPU.  The oracle replays the same loop
on the CPU from the same files.

Tolerances (the FP64 32^3 run's): statistics lines at their print precision
(the reference prints 9 significant digits, std::scientific with precision 8:
1e-8 relative per value, plus 1e-12 absolute for the normalised P_rms and
T_rms, which are zero up to rounding at iteration 0 where p and T are
uniform, and for t / t0, which is 0 there); P_ files (printed with 16
significant digits): positions 1e-9 (minimum-image distance), every property
column 1e-9 of its largest magnitude over iterations 0, 10 and 20 (at
iteration 0 rho, p and T are uniform, so their gradients and the pressure
Hessian are zero up to rounding there); T_ files 1e-12; the
iteration-0 field and restart files equal the initial condition to print
precision (1e-14 relative).
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from dns_gkm_lpt_amd import ic as icm
from dns_gkm_lpt_amd import output as outm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "dns_gkm_lpt_amd", "host", "dnsgkm_run")
NC, NPART, NSTEPS, CFL = 32, 4096, 21, 0.5


def oracle_loop(ic, xyz):
    """Evolve (update.cpp:39-188) on the oracle: the records of iterations 0,
    10, 20 (statistics line, particle file, time file)."""
    P = ic["params"]
    gas = po.make_gas(NC, gam=P["gam"], mu0=P["mu0"], p0=P["p0"], den0=P["rho0"])
    W = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"], gam=P["gam"])
    x = np.ascontiguousarray(xyz[:, [1, 0, 2]])  # initializeParticle: pos = (y, x, z)
    dt = po.cfl_dt(W, CFL, gas)
    t = 0.0
    rec = {}
    for it in range(NSTEPS):
        cs = po.velocity_coefs(W)
        U, gU, gV, gW = po.particle_sample(cs, gas.dx, x)
        if it % 10 == 0:
            ps = po.particle_stats(U, gU, gV, gW)
            fs = po.field_stats(W, cs, gas, P["p0"], P["Mt"], P["rho0"])
            line = [t / P["t0"], fs.tke, fs.diss, fs.sijsij, *fs.skew, *fs.kurt, fs.prms, fs.trms,
                    ps.tke, ps.sijsij, ps.skew_u, ps.skew_w]
            rec[it] = {"line": np.array(line), "pos": x.copy(), "props": po.particle_props(W, P["gam"], gas.dx, x),
                       "time": (t, dt, P["t0"])}
        W = po.eulerian_step(W, dt, gas, nthreads=8)
        x = po.particle_heun(po.velocity_coefs(W), gas.dx, x, U, dt)
        dt = po.cfl_dt(W, CFL, gas)
        t += dt
    return rec


def stats_lines(stdout):
    rows = []
    for line in stdout.splitlines():
        tok = line.split()
        if len(tok) != 16:
            continue
        try:
            rows.append([float(v) for v in tok])
        except ValueError:
            continue
    return np.array(rows)


def read_p_file(path):
    a = np.loadtxt(path, ndmin=2)
    tags = a[:, 0].astype(int)
    vals = a[:, 1:]
    pos = vals[:, [1, 0, 2]]  # written x, y, z
    props = np.full((a.shape[0], 41), np.nan)  # columns the file does not print stay NaN
    for col, src in enumerate(outm.P_COLUMNS):
        if not isinstance(src, tuple):
            props[:, src] = vals[:, col]
    return tags, pos, props


@pytest.mark.skipif(not os.access(DRIVER, os.X_OK), reason="dnsgkm_run not built (make -C dns_gkm_lpt_amd/host)")
def test_driver_run_matches_oracle(tmp_path):
    ic = icm.make_ic(NC, k0=4.0, Mt=0.3, Re=30.0, seed=0)
    xyz = icm.uniform_particles(NPART, 0)
    icm.write_case(str(tmp_path / "InitialConditions"), ic, xyz)
    env = dict(os.environ, DNSGKM_NSTEPS=str(NSTEPS), DNSGKM_CFL=str(CFL))
    print("dropin: running dnsgkm_run 32^3, 21 iterations", flush=True)
    run = subprocess.run([DRIVER], cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=300)
    print(run.stdout[-3000:], run.stderr[-2000:], flush=True)
    assert run.returncode == 0, run.stderr
    rec = oracle_loop(ic, xyz)
    P = ic["params"]

    rows = stats_lines(run.stdout)
    assert rows.shape == (3, 16), rows.shape
    for r, it in zip(rows, (0, 10, 20)):
        ref = rec[it]["line"]
        assert abs(r[0] - ref[0]) <= 1e-8 * abs(ref[0]) + 1e-12
        d = np.abs(r[1:] - ref[1:])
        print(f"dropin: stats line {it}: worst relative error "
              f"{np.max(d / np.maximum(np.abs(ref[1:]), 1e-300)):.2e}", flush=True)
        assert np.all(d <= 1e-8 * np.abs(ref[1:]) + 1e-12), (it, r, ref)

    cols = [c for c in outm.P_COLUMNS if not isinstance(c, tuple)]
    scale = np.max([np.max(np.abs(rec[it]["props"][:, cols]), axis=0) for it in (0, 10, 20)], axis=0)
    for it in (0, 10, 20):
        tags, pos, props = read_p_file(str(tmp_path / "Results" / "Particles" / f"P_0_{it}"))
        assert np.array_equal(tags, np.arange(NPART))
        d = np.abs(pos - rec[it]["pos"]) % (2 * np.pi)
        assert np.max(np.minimum(d, 2 * np.pi - d)) <= 1e-9
        ref = rec[it]["props"][:, cols]
        err = np.max(np.abs(props[:, cols] - ref), axis=0) / scale
        print(f"dropin: P_0_{it}: worst column error {err.max():.2e}", flush=True)
        assert np.all(err <= 1e-9), (it, err)
        tt = np.array([float(v) for v in open(tmp_path / "Results" / "Time" / f"T_{it}").read().split()])
        assert np.allclose(tt, rec[it]["time"], rtol=1e-12, atol=0)

    # iteration 0: the field file and the restart file hold the initial condition
    f0 = np.loadtxt(str(tmp_path / "Results" / "Field" / "Field_0_0"))
    assert f0.shape == (NC ** 3, 8)
    j, i, k = f0[:, 1].astype(int), f0[:, 0].astype(int), f0[:, 2].astype(int)
    for col, key in zip(range(3, 8), ("rho", "u", "v", "w", "p")):
        ref = ic[key][j, i, k]
        assert np.max(np.abs(f0[:, col] - ref)) <= 1e-14 * np.max(np.abs(ref)), key
    back = icm.read_init(str(tmp_path / "InitialCondition" / "Init000"), NC)
    for key in ("u", "v", "w", "rho", "p"):
        assert np.max(np.abs(back[key] - ic[key])) <= 1e-14 * np.max(np.abs(ic[key])), key
