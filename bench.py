"""Benchmark of the GKM-WENO step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nc 512] [--precision fp64|mixed]

A step is one Eulerian GKM-WENO step (CFL dt on the device, x/y/z fused
sweeps with the update) of the 512^3 periodic cube, state resident in HBM.
Prints one JSON line: cell-updates/s with the HBM-roofline fraction of the
dominant kernel (HIP events on the solver's stream) and the CPU baseline
(the reference's own evolveEulerian when oracle/_ref is built, else the
oracle restatement) timed on a bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
METRIC = "GKM-WENO cell-updates/s at 512^3, FP64 vs mixed, % of HBM roofline"
# algorithmic HBM bytes per cell of one launch: the x sweep reads the 5 FP64
# conserved values and writes the 5 FP64 flux-difference sums (40 + 40 B); the
# y and z sweeps also read the running sum (+40 B).  The whole step's minimum
# is 80 B per cell-update (read W^n, write W^{n+1}).
SWEEP_BYTES = {0: 80.0, 1: 120.0, 2: 120.0}
STEP_BYTES = 80.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(seconds_cap=30.0):
    """Time the CPU Eulerian step on a bounded sample of the same workload
    (one step of a 128^3 cube of the same initial-condition recipe, ~20 s on
    one core)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes

    import pyoracle as po
    from dns_gkm_lpt_amd import ic as icm

    nc = 128
    ic = icm.make_ic(nc, k0=4.0, Mt=0.5, Re=100.0, seed=0)
    P = ic["params"]
    W = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"])
    gas = po.make_gas(nc, gam=P["gam"], mu0=P["mu0"], p0=P["p0"], den0=P["rho0"])
    dt = po.cfl_dt(W, 0.5, gas)
    R = po.ref()
    if R is not None:
        R.ref_set_params(nc, P["gam"], P["mu0"], P["p0"], P["rho0"], dt)
        Wc = np.ascontiguousarray(W.copy())
        t = time.perf_counter()
        R.ref_eulerian(Wc)
        el = time.perf_counter() - t
        kind = "reference"
        what = "reference evolveEulerian (GKMflux.cpp + update.cpp CPU sweeps, 1 thread)"
    else:
        t = time.perf_counter()
        po.eulerian_step(W, dt, gas, nthreads=1)
        el = time.perf_counter() - t
        kind = "port"
        what = "oracle restatement of evolveEulerian (1 thread)"
    return {"value": nc ** 3 / el, "unit": "cell-updates/s", "cores": 1, "kind": kind,
            "sample": f"1 step of {nc}^3 (Mt 0.5, Re_lambda 100, k0 4, seed 0), {what}, {el:.2f} s"}


def copy_calibration(mib):
    """One plain device copy of `mib` MiB (a fill, then an elementwise negation
    out of place) so that FETCH_SIZE / WRITE_SIZE of a counter run can be
    scaled by what they read for a kernel of known traffic."""
    import torch

    n = mib * (1 << 20) // 8
    a = torch.full((n,), 2.0, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.neg(a, out=b)
    torch.cuda.synchronize()
    log(f"bench: copy calibration {mib} MiB read + {mib} MiB written (torch.neg), fill {mib} MiB written")
    del a, b


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nc", type=int, default=512)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=2, help="steps timed per kernel with HIP events")
    ap.add_argument("--tiles", default="auto",
                    help="sweep tile variants x,y,z (e.g. 0,0,0) or 'auto' to time them before the run")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain copy of this size first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the periodic cube is resident on a single MI355X and this build has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from dns_gkm_lpt_amd import _lib as gl
    from dns_gkm_lpt_amd import ic as icm

    if args.profile_steps < 1:
        ap.error("--profile-steps must be >= 1")
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
    nc = args.nc
    t0 = time.perf_counter()
    ic = icm.make_ic(nc, k0=4.0, Mt=0.5, Re=100.0, seed=0)
    P = ic["params"]
    W = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"])
    del ic
    log(f"bench: initial condition {nc}^3 built in {time.perf_counter() - t0:.1f} s")
    prec = gl.GKM_FP64 if args.precision == "fp64" else gl.GKM_MIXED
    ctx = gl.Context(nc, P, cfl=0.5, precision=prec)
    ctx.upload(W)
    del W
    log("bench: state uploaded")
    if args.tiles == "auto":
        tiles, tms = ctx.autotune(reps=2)
        log(f"bench: autotuned sweep tiles {tiles}; ms per variant (rows x, y, z) {np.round(tms, 2).tolist()}")
    else:
        ctx.set_tiles([int(v) for v in args.tiles.split(",")])
        tiles = ctx.tiles()
    ctx.run(args.warmup)
    ctx.sync()
    log(f"bench: {args.warmup} warm-up steps done")

    ctx.sync()
    t = time.perf_counter()
    ctx.run(args.steps)
    ctx.sync()
    el = time.perf_counter() - t
    ms_per_step = 1e3 * el / args.steps
    value = nc ** 3 * args.steps / el
    log(f"bench: {args.steps} steps in {el:.3f} s ({ms_per_step:.2f} ms/step)")

    # per-kernel durations, HIP events on the solver's stream
    ctx.timing(True)
    ctx.run(args.profile_steps)
    ms, cnt = ctx.timing_read()
    ctx.timing(False)
    names = {0: "sweep_kernel<0> (x)", 1: "sweep_kernel<1> (y)", 2: "sweep_kernel<2> (z, with update)",
             3: "cfl reduction"}
    kern = {names[i]: {"avg_ms": ms[i] / cnt[i], "launches": int(cnt[i])} for i in range(4) if cnt[i]}
    dom = max(range(3), key=lambda i: ms[i] / max(cnt[i], 1))
    dom_ms = ms[dom] / cnt[dom]
    achieved = SWEEP_BYTES[dom] * nc ** 3 / (dom_ms * 1e-3) / 1e9
    step_gbs = STEP_BYTES * value / 1e9
    log(f"bench: kernel times {kern}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 flux/WENO + f64 state",
        "data": "synthetic",
        "config": {"workload": f"GKM-WENO Eulerian step (CFL 0.5 dt on device + fused x/y/z sweeps + update), "
                               f"{nc}^3 periodic cube, Mt 0.5, Re_lambda 100, k0 4, seed 0",
                   "nc": nc, "precision": args.precision, "sweep_tiles": list(tiles)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": names[dom], "avg_ms": dom_ms,
                     "algorithmic_bytes_per_cell": SWEEP_BYTES[dom]},
        "step_hbm_fraction": step_gbs / HBM_PEAK_GBS,
        "kernels": kern,
    }
    ctx.close()
    if not args.no_cpu_baseline:
        log("bench: timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
