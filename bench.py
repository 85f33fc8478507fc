"""Benchmark of the GKM-WENO step and the particle tracker on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nc 512] [--precision both|mixed|fp64]
                    [--particles N] [--tiles auto|x,y,z] [--ensemble-steps N] [--c2-steps N] [--adjoint-steps N]
                    [--no-cpu-baseline]

A step is one Eulerian GKM-WENO step (CFL dt on the device, x/y/z fused sweeps
with the update) of the periodic cube, state resident in HBM.  The workload is
the 512^3 mixed-precision case (Mt 0.5, Re_lambda 100, k0 4, seed 0) with 2^22
tracers; the FP64 step of the same state is timed beside it ("FP64 vs mixed").
Prints one JSON line: cell-updates/s of the mixed step (`value`), the FP64
figures, the roofline of the dominant kernel (HIP events on the solver's
stream): its VALU rate against the FP64 peak with the reference's operation
count, and beside it its HBM fraction,
particle-updates/s, the 64-member ensemble (per-phase graph, eager and
persistent-kernel steps), the 128^3 FP64 run with 2^18 tracers (500 coupled
steps), the 64^3 discrete-adjoint gradient, and the CPU
baseline (the reference's own evolveEulerian when oracle/_ref is built, else
the oracle restatement) timed on a bounded sample on this host.
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
METRIC = "GKM-WENO cell-updates/s at 512^3, FP64 vs mixed, % of HBM roofline"
# algorithmic HBM bytes per cell of one launch: the x sweep reads the 5 FP64
# conserved values and writes the 5 FP64 flux-difference sums (40 + 40 B); the
# y and z sweeps also read the running sum (+40 B).  The whole step's minimum
# is 80 B per cell-update (read W^n, write W^{n+1}).
SWEEP_BYTES = {0: 80.0, 1: 120.0, 2: 120.0}
STEP_BYTES = 80.0
# the mixed path's compensated update also reads and writes an FP32 carry per
# cell and variable in the z sweep (20 + 20 B)
CARRY_BYTES = 40.0
# Arithmetic operations of the reference's step with one WENO reconstruction
# per interface (each +, -, *, / and transcendental counted once), counted
# through the oracle restatement by oracle/count_ops.cpp: per interface (one
# sweep launch does one interface per cell) and per cell-update (3 interfaces
# + the update).
ALG_OPS_PER_FACE = 8972
ALG_OPS_PER_CELL_UPDATE = 26951
# HBM bytes per sweep launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (each scaled by what the counter reads for a plain 1 GiB copy:
# FETCH x 2.0, WRITE x 1.0), GiB read + GiB written, keyed by
# (precision, direction, tile variant, nc); the commands and the table are in
# profiles/sweep_traffic_512.md.  Used for roofline.traffic when the dominant
# kernel of the run is one of these configurations.
# Mixed rows: the compensated update (the default); x and y sweeps do not
# touch the carry.
MEASURED_TRAFFIC_GIB = {
    ("mixed", 0, 0, 512): (7.96, 5.00),
    ("mixed", 0, 2, 512): (21.88, 5.00),
    ("mixed", 1, 0, 512): (14.59, 5.00),
    ("mixed", 1, 2, 512): (17.22, 5.00),
    ("mixed", 2, 0, 512): (23.18, 7.50),
    ("mixed", 2, 2, 512): (53.09, 7.85),
    ("fp64", 0, 0, 512): (52.88, 21.27),
    ("fp64", 0, 2, 512): (65.76, 17.07),
    ("fp64", 1, 0, 512): (56.53, 16.70),
    ("fp64", 1, 2, 512): (65.35, 18.61),
    ("fp64", 2, 0, 512): (59.67, 19.91),
    ("fp64", 2, 2, 512): (85.34, 19.45),
}
# rocprofv3 VALUBusy (% of time the VALU issues) of each sweep launch, keyed
# like MEASURED_TRAFFIC_GIB; commands and table in profiles/sweep_valu_512.md.
MEASURED_VALU_BUSY = {
    ("mixed", 0, 0, 512): 89.0,
    ("mixed", 0, 2, 512): 75.3,
    ("mixed", 1, 0, 512): 75.2,
    ("mixed", 1, 2, 512): 68.1,
    ("mixed", 2, 0, 512): 71.7,
    ("mixed", 2, 2, 512): 65.7,
    ("fp64", 0, 0, 512): 57.6,
    ("fp64", 0, 2, 512): 52.3,
    ("fp64", 1, 0, 512): 52.4,
    ("fp64", 1, 2, 512): 47.7,
    ("fp64", 2, 0, 512): 54.7,
    ("fp64", 2, 2, 512): 51.4,
}
KERNEL_NAMES = {0: "sweep_kernel<0> (x)", 1: "sweep_kernel<1> (y)", 2: "sweep_kernel<2> (z, with update)",
                3: "cfl reduction", 4: "spline prefilter (u, v, w)", 5: "particle sample + Heun",
                6: "particle cell sort"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_info():
    """Host CPU model and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def _cpu_step(W, nc, P, dt, use_ref):
    """One CPU Eulerian step of the padded state W; returns the wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    if use_ref:
        R = po.ref()
        R.ref_set_params(nc, P["gam"], P["mu0"], P["p0"], P["rho0"], dt)
        Wc = np.ascontiguousarray(W.copy())
        t = time.perf_counter()
        R.ref_eulerian(Wc)
        return time.perf_counter() - t
    gas = po.make_gas(nc, gam=P["gam"], mu0=P["mu0"], p0=P["p0"], den0=P["rho0"])
    t = time.perf_counter()
    po.eulerian_step(W, dt, gas, nthreads=1)
    return time.perf_counter() - t


def _cpu_worker(args):
    W, nc, P, dt, use_ref, q = args
    q.put(_cpu_step(W, nc, P, dt, use_ref))


def cpu_baseline(max_workers=16):
    """Time the CPU Eulerian step on a bounded sample of the same workload: one
    step of a 128^3 cube of the same initial-condition recipe (Mt 0.5, Re_lambda
    100, k0 4, seed 0), with the reference's own evolveEulerian when
    oracle/_ref is built (else the oracle restatement), (1) on one core and (2)
    as one independent single-threaded step per usable core (at most
    max_workers) run at once -- what the reference's MPI path gets from the
    same cores when each rank owns a block (its halo exchange not counted).
    Runs in forked worker processes before this process touches the GPU."""
    import multiprocessing as mp

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from dns_gkm_lpt_amd import ic as icm

    nc = 128
    ic = icm.make_ic(nc, k0=4.0, Mt=0.5, Re=100.0, seed=0)
    P = ic["params"]
    W = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"])
    gas = po.make_gas(nc, gam=P["gam"], mu0=P["mu0"], p0=P["p0"], den0=P["rho0"])
    dt = po.cfl_dt(W, 0.5, gas)
    use_ref = po.ref_available()
    info = cpu_info()
    ctx = mp.get_context("fork")
    q = ctx.Queue()
    # (1) one core
    w = ctx.Process(target=_cpu_worker, args=((W, nc, P, dt, use_ref, q),))
    w.start()
    el1 = q.get()
    w.join()
    # (2) every usable core at once
    nw = max(1, min(max_workers, info["usable"]))
    t = time.perf_counter()
    ws = [ctx.Process(target=_cpu_worker, args=((W, nc, P, dt, use_ref, q),)) for _ in range(nw)]
    for w in ws:
        w.start()
    els = [q.get() for _ in ws]
    for w in ws:
        w.join()
    wall = time.perf_counter() - t
    kind = "reference" if use_ref else "port"
    what = ("the reference's evolveEulerian (GKMflux.cpp + update.cpp CPU sweeps, oracle/_ref)" if use_ref
            else "the oracle restatement of evolveEulerian")
    out = {"value": nw * nc ** 3 / wall, "unit": "cell-updates/s", "cores": nw, "kind": kind,
           "sample": f"{nw} concurrent single-threaded steps of {nc}^3 (Mt 0.5, Re_lambda 100, k0 4, seed 0), "
                     f"one per core, {what}; {wall:.2f} s wall (per-step {min(els):.2f}-{max(els):.2f} s)",
           "cpu_model": info["model"], "nproc": info["nproc"], "usable_cores": info["usable"],
           "one_core": {"value": nc ** 3 / el1, "s_step": el1}}
    log(f"bench: CPU baseline {out}")
    return out


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


def kernel_table(ms, cnt):
    return {KERNEL_NAMES[i]: {"avg_ms": ms[i] / cnt[i], "launches": int(cnt[i])} for i in range(7) if cnt[i]}


def run_eulerian(gl, nc, P, W, prec_name, args):
    """Warm-up, K timed steps, per-kernel event timing; returns (ctx, result)."""
    prec = gl.GKM_FP64 if prec_name == "fp64" else gl.GKM_MIXED
    ctx = gl.Context(nc, P, cfl=0.5, precision=prec)
    ctx.upload(W)
    if args.tiles == "auto":
        tiles, tms = ctx.autotune(reps=2)
        log(f"bench: {prec_name}: autotuned sweep tiles {tiles}; ms per variant (rows x, y, z) "
            f"{np.round(tms, 2).tolist()}")
    else:
        ctx.set_tiles([int(v) for v in args.tiles.split(",")])
        tiles = ctx.tiles()
    ctx.run(args.warmup)
    ctx.sync()
    t = time.perf_counter()
    ctx.run(args.steps)
    ctx.sync()
    el = time.perf_counter() - t
    ms_per_step = 1e3 * el / args.steps
    value = nc ** 3 * args.steps / el
    log(f"bench: {prec_name}: {args.steps} steps in {el:.3f} s ({ms_per_step:.2f} ms/step)")
    extra = {}
    if prec == gl.GKM_MIXED and ctx.compensated:
        # the same steps with the plain FP64 update (no carried rounding error)
        ctx.set_compensation(False)
        ctx.run(1)
        ctx.sync()
        t = time.perf_counter()
        ctx.run(args.steps)
        ctx.sync()
        el_plain = time.perf_counter() - t
        ctx.set_compensation(True)
        extra["no_compensation"] = {"ms_per_step": 1e3 * el_plain / args.steps,
                                    "value": nc ** 3 * args.steps / el_plain}
        log(f"bench: {prec_name}: without compensation {1e3 * el_plain / args.steps:.2f} ms/step")
    ctx.timing(True)
    ctx.run(args.profile_steps)
    ms, cnt = ctx.timing_read()
    ctx.timing(False)
    kern = kernel_table(ms, cnt)
    log(f"bench: {prec_name}: kernel times {kern}")
    return ctx, {"value": value, "ms_per_step": ms_per_step, "sweep_tiles": list(tiles), "kernels": kern,
                 "ms": ms, "cnt": cnt, **extra}


def run_particles(ctx, npart, steps):
    """Coupled steps (Eulerian step + spline of the new field + Heun) with npart
    tracers; the particle share is timed with HIP events per kernel."""
    from dns_gkm_lpt_amd import ic as icm

    pos = np.ascontiguousarray(icm.uniform_particles(npart, 0)[:, [1, 0, 2]])
    ctx.lpt_upload(pos)
    del pos
    dt = ctx.cfl_dt()
    ctx.lpt_coupled_step(dt)  # warm-up (builds the first spline)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        ctx.lpt_coupled_step(dt)
    ctx.sync()
    coupled_ms = 1e3 * (time.perf_counter() - t) / steps
    ctx.timing(True)
    for _ in range(steps):
        ctx.lpt_coupled_step(dt)
    ms, cnt = ctx.timing_read()
    ctx.timing(False)
    kern = kernel_table(ms, cnt)
    # per step: one spline build, one sample + one Heun launch (slot 5), a sort every 10th step
    part_ms = (ms[4] + ms[5] + ms[6]) / steps
    log(f"bench: particles: {npart} tracers, coupled step {coupled_ms:.2f} ms, particle share {part_ms:.2f} ms; "
        f"{kern}")
    return {"n": npart, "value": npart / (part_ms * 1e-3), "unit": "particle-updates/s",
            "ms_per_step": part_ms, "coupled_ms_per_step": coupled_ms,
            "kernels": {k: v for k, v in kern.items() if not k.startswith("sweep") and k != "cfl reduction"},
            "note": "particle share = spline prefilter of u, v, w + sample + Heun + amortised cell sort"}


def run_ensemble(gl, steps, warmup):
    """The batched-ensemble configuration: 64 independent realizations (24 at
    32^3, 24 at 48^3, 16 at 64^3), each with its own seed, k0 in [3, 6], Mt in
    [0.1, 0.6], Re_lambda in [20, 60], mixed precision, 8192 tracers each,
    tiles autotuned; steps replayed from captured HIP graphs (one launch per
    phase), timed beside the same steps launched eagerly and as one
    persistent-kernel launch per step (also graph-replayed)."""
    from dns_gkm_lpt_amd import ic as icm

    cases = icm.ensemble_members()
    ncs = [c[0] for c in cases]
    cells = sum(nc ** 3 for nc in ncs)
    ntr = 8192
    out = {"members": len(ncs), "grids": "24 x 32^3 + 24 x 48^3 + 16 x 64^3", "precision": "mixed",
           "tracers_per_member": ntr}
    with gl.Ensemble(ncs, [c[1] for c in cases], cfl=0.5, precision=gl.GKM_MIXED) as e:
        for m, c in enumerate(cases):
            e.members[m].upload(c[2])
            e.members[m].lpt_upload(icm.uniform_particles(ntr, seed=c[3]))
        tiles, _ = e.autotune(reps=2)
        e.run(warmup, graph=True)
        for mode, graph, persistent in (("graph", True, False), ("eager", False, False),
                                        ("persistent", True, True)):
            e.set_persistent(persistent)
            if persistent:
                e.run(warmup, graph=True)
            t = time.perf_counter()
            e.run(steps, graph=graph)
            el = time.perf_counter() - t
            out[f"ms_per_step_{mode}"] = 1e3 * el / steps
            out[f"member_cell_updates_per_s_{mode}"] = cells * steps / el
            out[f"particle_updates_per_s_{mode}"] = ntr * len(ncs) * steps / el
        # phase times of one persistent step as workgroup 0 saw the grid barriers
        names = ["cfl maxima + tracer velocities", "dt", "x sweep", "y sweep", "z sweep",
                 "spline z pass", "spline x pass", "spline y pass", "Heun"]
        out["persistent_phase_ms"] = dict(zip(names, (round(float(v), 4) for v in e.persistent_profile())))
        out["sweep_tiles"] = list(tiles)
        out["steps"] = steps
    out["value"] = out["member_cell_updates_per_s_graph"]
    out["unit"] = "member cell-updates/s"
    log(f"bench: ensemble {out}")
    return out


def run_c2(steps):
    """The 128^3 FP64 configuration run end to end: Mt 0.3, Re_lambda 50, k0 4,
    seed 0, 2^18 seeded tracers, `steps` coupled steps (device CFL dt, Eulerian
    step, tracer sample and Heun step each step), tiles autotuned; timed from
    the first step to the last, state and tracers resident."""
    from dns_gkm_lpt_amd import _lib as gl
    from dns_gkm_lpt_amd import ic as icm

    nc, ntr = 128, 1 << 18
    ic = icm.make_ic(nc, k0=4.0, Mt=0.3, Re=50.0, seed=0)
    P = ic["params"]
    W = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"], gam=P["gam"])
    x0 = np.ascontiguousarray(icm.uniform_particles(ntr, 0)[:, [1, 0, 2]])
    with gl.Context(nc, P, cfl=0.5, precision=gl.GKM_FP64) as c:
        c.upload(W)
        tiles, _ = c.autotune(reps=2)
        c.upload(W)
        c.lpt_upload(x0)
        c.sync()
        t = time.perf_counter()
        for _ in range(steps):
            c.lpt_coupled_step(c.cfl_dt())
        c.sync()
        el = time.perf_counter() - t
    out = {"grid": "128^3", "precision": "fp64", "tracers": ntr, "steps": steps, "s_run": el,
           "ms_per_coupled_step": 1e3 * el / steps, "cell_updates_per_s": nc ** 3 * steps / el,
           "particle_updates_per_s": ntr * steps / el, "sweep_tiles": list(tiles)}
    log(f"bench: 128^3 run {out}")
    return out


def run_adjoint(steps):
    """The discrete-adjoint configuration: dJ/du0 of the final kinetic energy
    J = <rho |u|^2 / 2> with respect to the initial velocity, 64^3 periodic
    cube (Mt 0.3, Re_lambda 40, k0 4, seed 0), FP64, dt fixed from the initial
    CFL 0.5, checkpoints every 20 steps.  Timed separately: one gradient
    (forward run keeping checkpoints + adjoint run recomputing each segment),
    one tangent-linear run along a seeded band-limited unit direction, and one
    forward run; the adjoint and tangent-linear directional derivatives are
    compared."""
    from dns_gkm_lpt_amd import adjoint as adj

    pb, u = adj.make_problem(64, seed=0, nsteps=steps, every=20)
    d = adj.band_limited_direction(u.shape, 100)
    t = time.perf_counter()
    J, g = pb.gradient(u)
    s_grad = time.perf_counter() - t
    t = time.perf_counter()
    tl = pb.directional(u, d)
    s_tl = time.perf_counter() - t
    t = time.perf_counter()
    pb.objective(u)
    s_fwd = time.perf_counter() - t
    ad = float(np.sum(g * d))
    # central differences at the configuration's step h = 1e-6 |u0| (and 1e-8)
    # along two band-limited and two white-noise unit directions (untimed)
    dirs = {"band_limited": [adj.band_limited_direction(u.shape, 100 + k) for k in range(2)],
            "white_noise": [adj.white_noise_direction(u.shape, 200 + k) for k in range(2)]}
    fd = {}
    for kind, ds in dirs.items():
        _, _, rep = adj.check_gradient(pb, u, ds, (1e-6, 1e-8), g=g)
        fd[kind] = {"err_fd": [r["err_fd"] for r in rep["rows"]], "err_tl": [r["err_tl"] for r in rep["rows"]]}
    out = {"grid": "64^3", "steps": steps, "checkpoint_every": 20, "precision": "fp64", "dt": pb.dt,
           "s_gradient": s_grad, "s_tangent_linear": s_tl, "s_forward": s_fwd,
           "gradient_over_forward": s_grad / s_fwd, "adjoint_vs_tangent_linear": abs(ad - tl) / abs(tl),
           "finite_differences": fd,
           "fd_note": "relative error of central differences of J against the adjoint gradient, h relative to |u0|; "
                      "white-noise directions cross the step's branch switches (WENO maximum, |p_l - p_r|) "
                      "in O(h) of the cells, so their error falls with h"}
    log(f"bench: adjoint {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nc", type=int, default=512)
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="'both': value is the mixed step, the FP64 step of the same state is timed beside it")
    ap.add_argument("--particles", type=int, default=1 << 22, help="tracers for the particle-update rate (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=200,
                    help="steps of the 64-member ensemble configuration (0: skip)")
    ap.add_argument("--c2-steps", type=int, default=500,
                    help="coupled steps of the 128^3 FP64 configuration with 2^18 tracers (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the 64^3 discrete-adjoint gradient configuration (0: skip)")
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
    if args.profile_steps < 1 or args.steps < 1:
        ap.error("--steps and --profile-steps must be >= 1")

    cpu = None
    if not args.no_cpu_baseline:  # first: its worker processes fork before this process touches the GPU
        log("bench: timing the CPU baseline")
        cpu = cpu_baseline()

    from dns_gkm_lpt_amd import _lib as gl
    from dns_gkm_lpt_amd import ic as icm

    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
    nc = args.nc
    t0 = time.perf_counter()
    ic = icm.make_ic(nc, k0=4.0, Mt=0.5, Re=100.0, seed=0)
    P = ic["params"]
    W = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"])
    del ic
    log(f"bench: initial condition {nc}^3 built in {time.perf_counter() - t0:.1f} s")

    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res = {}
    particles = None
    for prec in precs:
        ctx, r = run_eulerian(gl, nc, P, W, prec, args)
        if prec == precs[-1] and args.particles > 0:
            particles = run_particles(ctx, args.particles, args.steps)
        ctx.close()
        res[prec] = r
    del W
    head = res[precs[-1]]
    ms, cnt = head["ms"], head["cnt"]
    dom = max(range(3), key=lambda i: ms[i] / max(cnt[i], 1))
    dom_ms = ms[dom] / cnt[dom]
    launch_bytes = SWEEP_BYTES[dom] + (CARRY_BYTES if dom == 2 and "no_compensation" in head else 0.0)
    achieved = launch_bytes * nc ** 3 / (dom_ms * 1e-3) / 1e9
    hbm80 = STEP_BYTES * nc ** 3 / (dom_ms * 1e-3) / 1e9
    alg_tops = ALG_OPS_PER_FACE * nc ** 3 / (dom_ms * 1e-3) / 1e12
    measured = MEASURED_TRAFFIC_GIB.get((precs[-1], dom, head["sweep_tiles"][dom], nc))
    traffic = None if measured is None else (measured[0] + measured[1]) * 2.0 ** 30

    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precs[-1] == "fp64" else "f32 WENO weights + f32 flux slope terms, f64 state/Euler part",
        "data": "synthetic",
        "config": {"workload": f"GKM-WENO Eulerian step (CFL 0.5 dt on device + fused x/y/z sweeps + update), "
                               f"{nc}^3 periodic cube, Mt 0.5, Re_lambda 100, k0 4, seed 0",
                   "nc": nc, "precision": precs[-1], "sweep_tiles": head["sweep_tiles"]},
        # The sweep is VALU-bound (about 340 reference operations per algorithmic
        # byte against a ridge near 10): the roofline is the FP64 VALU peak with
        # the reference's operation count; the HBM figures stand beside it.
        "roofline": {"bound": "valu", "achieved": alg_tops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": alg_tops / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": KERNEL_NAMES[dom], "avg_ms": dom_ms, "alg_ops_per_face": ALG_OPS_PER_FACE,
                     "note": "achieved = reference operations per face (oracle/count_ops.cpp, one WENO per "
                             "interface) x nc^3 faces / average launch time (HIP events on the solver's "
                             "stream); peak = MI355X FP64 vector peak (non-packed FP32 issues at the same "
                             "rate)",
                     "traffic_note": "HBM bytes per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, "
                                     "calibrated on a 1 GiB copy; profiles/sweep_traffic_512.md)"
                                     if traffic else "no counter measurement for this configuration",
                     "hbm": {"achieved_gbs": hbm80, "frac": hbm80 / HBM_PEAK_GBS, "bytes_per_cell": STEP_BYTES,
                             "launch_alg_bytes_per_cell": launch_bytes,
                             "launch_alg_gbs": achieved, "launch_alg_frac": achieved / HBM_PEAK_GBS,
                             "traffic_over_launch_alg": None if traffic is None else
                             traffic / (launch_bytes * nc ** 3),
                             "peak_gbs": HBM_PEAK_GBS},
                     "valu_busy_pct_rocprof": MEASURED_VALU_BUSY.get((precs[-1], dom, head["sweep_tiles"][dom], nc)),
                     "valu_busy_note": "rocprofv3 VALUBusy of this kernel configuration "
                                       "(profiles/sweep_valu_512.md)"},
        "step_hbm_fraction": STEP_BYTES * head["value"] / 1e9 / HBM_PEAK_GBS,
        "kernels": head["kernels"],
    }
    if "fp64" in res and "mixed" in res:
        out["fp64"] = {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["ms_per_step"],
                       "sweep_tiles": res["fp64"]["sweep_tiles"], "kernels": res["fp64"]["kernels"]}
        out["mixed_over_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    if "no_compensation" in head:
        out["mixed_no_compensation"] = dict(head["no_compensation"], note=(
            "the same mixed steps with the plain FP64 update instead of the compensated one "
            "(gkm_set_compensation(ctx, 0))"))
    if particles is not None:
        out["particles"] = particles
    if args.ensemble_steps > 0:
        out["ensemble"] = run_ensemble(gl, args.ensemble_steps, args.warmup)
    if args.c2_steps > 0:
        out["run_128"] = run_c2(args.c2_steps)
    if args.adjoint_steps > 0:
        out["adjoint"] = run_adjoint(args.adjoint_steps)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
