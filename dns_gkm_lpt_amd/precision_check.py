"""Mixed precision against FP64 on the same seeded run.

Runs the reference's decaying-turbulence case (Initialize_DHIT,
FlowField.cpp:60-114, with the seeded initial-condition recipe of ic.py) twice
on the GPU -- FP64 and mixed precision -- with tracers, dt from each run's own
device CFL every step, the statistics of PrintStats (update.cpp:688-842) every
`every` steps, and reports the largest relative error of each statistics
curve, the final fields' normwise error and the final tracer positions'
largest minimum-image distance, mixed against FP64.  Curves are compared at
equal step numbers (the two runs' dt histories differ by rounding).

    python -m dns_gkm_lpt_amd.precision_check --nc 512 --steps 1000 --particles 4194304
"""
import argparse
import json
import time

import numpy as np

from . import _lib as gl
from . import ic as icm


def run(nc, P, W0, x0, steps, every, precision, tiles, log):
    curves, dts = [], []
    t0 = time.perf_counter()
    with gl.Context(nc, P, cfl=0.5, precision=precision) as c:
        c.set_tiles(tiles)
        c.upload(W0)
        if len(x0):
            c.lpt_upload(x0)
        for n in range(1, steps + 1):
            dt = c.cfl_dt()
            dts.append(dt)
            if len(x0):
                c.lpt_coupled_step(dt)
            else:
                c.step(dt)
            if n % every == 0:
                s = c.stats()
                curves.append([s.tke, s.diss, s.sijsij] + list(s.skew))
            if n % 100 == 0:
                log(f"{'mixed' if precision == gl.GKM_MIXED else 'fp64'}: step {n}, "
                    f"{time.perf_counter() - t0:.1f} s")
        W = c.download()
        x = c.lpt_download()[0] if len(x0) else np.zeros((0, 3))
    return np.array(curves), np.array(dts), W, x, (time.perf_counter() - t0) / steps


def compare(nc=512, steps=1000, every=10, particles=1 << 22, Mt=0.5, Re=100.0, k0=4.0, seed=0,
            tiles_fp64=(0, 2, 0), tiles_mixed=(2, 2, 0), log=print):
    ic = icm.make_ic(nc, k0=k0, Mt=Mt, Re=Re, seed=seed)
    P = ic["params"]
    W0 = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"], gam=P["gam"])
    del ic
    x0 = np.ascontiguousarray(icm.uniform_particles(particles, seed)[:, [1, 0, 2]]) if particles else np.zeros((0, 3))
    log(f"initial condition ready: {nc}^3, {particles} tracers")
    c64, dt64, W64, x64, s64 = run(nc, P, W0, x0, steps, every, gl.GKM_FP64, tiles_fp64, log)
    cmx, dtmx, Wmx, xmx, smx = run(nc, P, W0, x0, steps, every, gl.GKM_MIXED, tiles_mixed, log)
    rel = np.abs(cmx - c64) / np.abs(c64)
    inner = (slice(None),) + (slice(3, nc + 3),) * 3
    fields = [float(np.max(np.abs(Wmx[inner][q] - W64[inner][q])) / np.max(np.abs(W64[inner][q]))) for q in range(5)]
    d = np.abs(xmx - x64) % (2 * np.pi)
    dist = float(np.max(np.minimum(d, 2 * np.pi - d))) if particles else 0.0
    names = ["tke", "diss", "sijsij", "skew_x", "skew_y", "skew_z"]
    return {
        "nc": nc, "steps": steps, "every": every, "particles": particles, "Mt": Mt, "Re_lambda": Re, "k0": k0,
        "seed": seed,
        "max_rel_err": {k: float(v) for k, v in zip(names, rel.max(axis=0))},
        "final_rel_err": {k: float(v) for k, v in zip(names, rel[-1])},
        "ke_decay": {"fp64": float(c64[-1, 0] / c64[0, 0]), "mixed": float(cmx[-1, 0] / cmx[0, 0])},
        "dt_max_rel_diff": float(np.max(np.abs(dtmx - dt64) / dt64)),
        "final_fields_rel_err": fields,
        "final_tracer_max_distance": dist,
        "s_per_step": {"fp64": s64, "mixed": smx},
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--nc", type=int, default=512)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--particles", type=int, default=1 << 22)
    ap.add_argument("--Mt", type=float, default=0.5)
    ap.add_argument("--Re", type=float, default=100.0)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    out = compare(a.nc, a.steps, a.every, a.particles, a.Mt, a.Re, seed=a.seed,
                  log=lambda m: print(m, flush=True))
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
