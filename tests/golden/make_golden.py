"""Write the golden vectors under tests/golden/ from the reference's own code.

Runs oracle/_ref/libgkm_ref.so, which oracle/Makefile (target `ref`) compiles
from the reference's GKMflux.cpp, param.cpp and the CPU Eulerian-step
functions of update.cpp.  Inputs are seeded synthetic states around the
reference's dimensional base state; outputs are what the reference computes.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import pyoracle as po  # noqa: E402
from cases import make_case, random_faces, random_stencils  # noqa: E402


def main():
    R = po.ref()
    if R is None:
        po.build(ref=True)
        R = po.ref()
    # --- flux(): 96 faces, mild and strong jumps, plus a uniform-state face
    nc = 16
    W, P, gas = make_case(nc, seed=0)
    X = np.concatenate([random_faces(48, seed=11, jump=0.05), random_faces(48, seed=12, jump=0.4)])
    X[0, 5:10] = X[0, 0:5]
    X[0, 10:20] = np.tile(X[0, 0:5], 2)
    X[0, 20:] = 0.0
    tau = np.array([po.tau_of(X[i, :5], gas) for i in range(X.shape[0])])
    dt = 1.7e-4
    R.ref_set_params(nc, P["gam"], P["mu0"], P["p0"], P["rho0"], dt)
    F = np.empty((X.shape[0], 5))
    R.ref_flux_batch(X.shape[0], X, tau, dt, R.ref_get_dx(), F)
    np.savez_compressed(os.path.join(HERE, "flux.npz"), X=X, tau=tau, dt=dt, nc=nc, gam=P["gam"], mu0=P["mu0"],
                        p0=P["p0"], rho0=P["rho0"], F=F)
    # --- ApplyWENO: smooth, rough and discontinuous stencils
    S = np.concatenate([random_stencils(48, seed=21, jump=0.02), random_stencils(48, seed=22, jump=0.5)])
    S[:8, 15:] = S[:8, 15:] * 1.5  # a jump between cells i and i+1
    out = np.empty((S.shape[0], 10))
    R.ref_weno_batch(S.shape[0], S, out)
    np.savez_compressed(os.path.join(HERE, "weno.npz"), S=S, out=out, gam=P["gam"])
    # --- evolveEulerian on a 12^3 cube, two seeds
    for seed in (0, 1):
        nc = 12
        W, P, gas = make_case(nc, seed=seed, perturb=0.08)
        dt = po.cfl_dt(W, 0.5, gas)
        R.ref_set_params(nc, P["gam"], P["mu0"], P["p0"], P["rho0"], dt)
        Wn = np.ascontiguousarray(W.copy())
        R.ref_eulerian(Wn)
        np.savez_compressed(os.path.join(HERE, f"step12_seed{seed}.npz"), W=W, Wn=Wn, dt=dt, nc=nc, gam=P["gam"],
                            mu0=P["mu0"], p0=P["p0"], rho0=P["rho0"])
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
