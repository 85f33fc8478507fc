"""Gradient of the final kinetic energy with respect to the initial velocity.

The objective is the volume average of (1/2) rho |u|^2 over the interior cells
of the state after nsteps fixed-dt FP64 steps (the statistics of the reference
print <rho |u|^2> without the 1/2, PrintStats, update.cpp:688-842).  The
initial state is the reference's initial-condition map (Initialize_DHIT,
FlowField.cpp:89-105) of the velocity field with uniform density and pressure.

    J(u0) = (1/nc^3) sum (1/2) |m_N|^2 / rho_N,   W_N = step^N(W_0(u0))

The gradient is one adjoint run (gkm_adjoint_forward: the trajectory with
checkpoints every `every` steps, ending at W_N, from which J and lam_N are
formed; gkm_adjoint_backward: each segment but the last recomputed on the way
back) followed by the transpose of the initial-condition map.  KineticEnergy
wraps it as a torch.autograd.Function.
"""
import numpy as np

from . import _lib as gl
from . import ic as icm


def cell_kinetic_energy(W):
    """(1/2) rho |u|^2 of each interior cell of a padded state."""
    nc = W.shape[-1] - 6
    inner = (slice(3, nc + 3),) * 3
    rho = W[0][inner]
    m2 = W[1][inner] ** 2 + W[2][inner] ** 2 + W[3][inner] ** 2
    return 0.5 * m2 / rho


def kinetic_energy(W):
    """Volume-averaged (1/2) rho |u|^2 of a padded state."""
    return float(np.mean(cell_kinetic_energy(W)))


def kinetic_energy_adjoint(W):
    """dJ/dW of kinetic_energy (padded layout, zero ghost layers)."""
    nc = W.shape[-1] - 6
    inner = (slice(3, nc + 3),) * 3
    lam = np.zeros_like(W)
    rho = W[0][inner]
    m = [W[q][inner] for q in (1, 2, 3)]
    n3 = float(nc ** 3)
    lam[0][inner] = -0.5 * (m[0] ** 2 + m[1] ** 2 + m[2] ** 2) / rho ** 2 / n3
    for q in range(3):
        lam[1 + q][inner] = m[q] / rho / n3
    return lam


def initial_state(u, rho, p, gam):
    """W_0 of a velocity field u (3, nc, nc, nc) with uniform rho and p."""
    return icm.prim_to_cons(u[0], u[1], u[2], np.full_like(u[0], rho), np.full_like(u[0], p), gam=gam)


def initial_state_transpose(lam0, u, rho):
    """Transpose of the initial-condition map: dJ/du from lam0 = dJ/dW_0.
    m = rho u and E = p/(gam-1) + rho |u|^2 / 2, so dJ/du_i = rho lam_{1+i} + rho u_i lam_4
    (the ghost layers of W_0 are copies of interior cells and are not independent)."""
    nc = u.shape[-1]
    inner = (slice(3, nc + 3),) * 3
    return np.stack([rho * lam0[1 + i][inner] + rho * u[i] * lam0[4][inner] for i in range(3)])


class Problem:
    """The fixed-dt problem: grid, gas parameters, uniform rho and p, dt, steps."""

    def __init__(self, nc, params, dt, nsteps, every, rho=None, p=None):
        self.nc, self.params, self.dt, self.nsteps, self.every = nc, params, dt, nsteps, every
        self.rho = params["rho0"] if rho is None else rho
        self.p = params["p0"] if p is None else p

    def context(self):
        return gl.Context(self.nc, self.params, precision=gl.GKM_FP64)

    def final_state(self, u):
        """W_N of a forward run from u."""
        with self.context() as c:
            c.upload(initial_state(u, self.rho, self.p, self.params["gam"]))
            for _ in range(self.nsteps):
                c.step(self.dt)
            return c.download()

    def objective(self, u):
        """J(u) by a forward run."""
        return kinetic_energy(self.final_state(u))

    def central_difference(self, u, d, h):
        """(J(u + h d) - J(u - h d)) / 2h, summed as per-cell differences so
        that the rounding of J itself (an ulp of J, ~1e-16 J) does not enter."""
        ep = cell_kinetic_energy(self.final_state(u + h * d))
        em = cell_kinetic_energy(self.final_state(u - h * d))
        return float(np.mean(ep - em)) / (2 * h)

    def gradient(self, u):
        """(J(u), dJ/du) by a forward run and an adjoint run."""
        W0 = initial_state(u, self.rho, self.p, self.params["gam"])
        with self.context() as c:
            c.upload(W0)
            c.adjoint_forward(self.dt, self.nsteps, self.every)
            WN = c.download()
            J = kinetic_energy(WN)
            lam0 = c.adjoint_backward(kinetic_energy_adjoint(WN))
        return J, initial_state_transpose(lam0, u, self.rho)

    def directional(self, u, d):
        """dJ/du . d by a tangent-linear run (an independent check of the adjoint)."""
        W0 = initial_state(u, self.rho, self.p, self.params["gam"])
        nc = self.nc
        inner = (slice(3, nc + 3),) * 3
        dW0 = np.zeros_like(W0)
        for i in range(3):
            dW0[1 + i][inner] = self.rho * d[i]
        dW0[4][inner] = self.rho * (u[0] * d[0] + u[1] * d[1] + u[2] * d[2])
        dW0 = icm.fill_halo(dW0)
        with self.context() as c:
            c.upload(W0)
            dWN = c.tl_run(self.dt, self.nsteps, dW0)
            WN = c.download()
        lamN = kinetic_energy_adjoint(WN)
        return float(np.sum(lamN * dWN))


def KineticEnergy(problem):
    """A torch.autograd.Function computing J(u) for a (3, nc, nc, nc) float64
    tensor u; backward is the adjoint gradient."""
    import torch

    class _KE(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u):
            un = u.detach().cpu().numpy().astype(np.float64)
            J, g = problem.gradient(un)
            ctx.save_for_backward(torch.from_numpy(g).to(u.dtype))
            return torch.tensor(J, dtype=u.dtype)

        @staticmethod
        def backward(ctx, gJ):
            (g,) = ctx.saved_tensors
            return gJ * g

    return _KE


def band_limited_direction(shape, seed, kmax=4):
    """A seeded random unit vector field (shape (3, nc, nc, nc)) holding only
    the Fourier modes |k| <= kmax (the initial condition's band)."""
    rng = np.random.default_rng(seed)
    nc = shape[-1]
    k = np.fft.fftfreq(nc, 1.0 / nc)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    d = np.stack([np.real(np.fft.ifftn(np.fft.fftn(rng.standard_normal((nc,) * 3)) * (kk <= kmax)))
                  for _ in range(shape[0])])
    return d / np.linalg.norm(d)


def white_noise_direction(shape, seed):
    """A seeded random unit vector field with independent Gaussian cells."""
    d = np.random.default_rng(seed).standard_normal(shape)
    return d / np.linalg.norm(d)


def make_problem(nc, k0=4.0, Mt=0.3, Re=40.0, seed=0, nsteps=200, every=20, cfl=0.5):
    """The reference's initial condition (Initialize_DHIT, FlowField.cpp:89-105)
    with dt fixed from its CFL number; returns (Problem, u0)."""
    ic = icm.make_ic(nc, k0=k0, Mt=Mt, Re=Re, seed=seed)
    P = ic["params"]
    u = np.stack([ic["u"], ic["v"], ic["w"]])
    with gl.Context(nc, P, precision=gl.GKM_FP64, cfl=cfl) as c:
        c.upload(initial_state(u, P["rho0"], P["p0"], P["gam"]))
        dt = c.cfl_dt()
    return Problem(nc, P, dt, nsteps, every), u


def check_gradient(problem, u, directions, h_rels=(1e-6,), g=None):
    """Compare the adjoint gradient with central differences of J along each
    unit direction d, h = h_rel |u| for every h_rel given, and with the
    tangent-linear derivative along d.  A given gradient g is reused (J is then
    None).  Returns (J, g, report): the gradient's wall time and, per
    direction, the adjoint, tangent-linear and finite-difference derivatives
    with their relative errors (keyed by h_rel) and wall times."""
    import time
    t0 = time.perf_counter()
    J, g = problem.gradient(u) if g is None else (None, g)
    t_grad = time.perf_counter() - t0
    rows = []
    for d in directions:
        ad = float(np.sum(g * d))
        t0 = time.perf_counter()
        tl = problem.directional(u, d)
        row = {"adjoint": ad, "tl": tl, "err_tl": abs(tl - ad) / abs(ad), "s_tl": time.perf_counter() - t0,
               "fd": {}, "err_fd": {}}
        for hr in h_rels:
            fd = problem.central_difference(u, d, hr * np.linalg.norm(u))
            row["fd"][f"{hr:g}"] = fd
            row["err_fd"][f"{hr:g}"] = abs(fd - ad) / abs(ad)
        rows.append(row)
    return J, g, {"s_gradient": t_grad, "rows": rows}


def main(argv=None):
    """Gradient check of the 64^3 / 200-step / checkpoint-every-20 problem
    (Mt 0.3, Re_lambda 40, dt fixed from the initial CFL 0.5): seeded
    band-limited and white-noise unit directions, one JSON line."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--nc", type=int, default=64)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--dirs", type=int, default=8)
    ap.add_argument("--white", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--h", default="1e-6,1e-8", help="finite-difference steps relative to |u0|")
    a = ap.parse_args(argv)
    hs = [float(x) for x in a.h.split(",")]
    pb, u = make_problem(a.nc, seed=a.seed, nsteps=a.steps, every=a.every)
    dirs = [band_limited_direction(u.shape, 100 + k) for k in range(a.dirs)]
    J, g, rep = check_gradient(pb, u, dirs, hs)
    wdirs = [white_noise_direction(u.shape, 200 + k) for k in range(a.white)]
    _, _, wrep = check_gradient(pb, u, wdirs, hs, g=g)
    out = {"nc": a.nc, "steps": a.steps, "every": a.every, "dt": pb.dt, "J": J, "grad_norm": float(np.linalg.norm(g)),
           "s_gradient": rep["s_gradient"], "band_limited": rep["rows"], "white_noise": wrep["rows"],
           "max_err_fd": {f"{h:g}": max(r["err_fd"][f"{h:g}"] for r in rep["rows"]) for h in hs},
           "max_err_tl": max(r["err_tl"] for r in rep["rows"])}
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
