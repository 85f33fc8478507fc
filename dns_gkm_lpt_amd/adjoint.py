"""Gradient of the final kinetic energy with respect to the initial velocity.

The objective is the volume average of (1/2) rho |u|^2 over the interior cells
of the state after nsteps fixed-dt FP64 steps (the statistics of the reference
print <rho |u|^2> without the 1/2, PrintStats, update.cpp:688-842).  The
initial state is the reference's initial-condition map (Initialize_DHIT,
FlowField.cpp:89-105) of the velocity field with uniform density and pressure.

    J(u0) = (1/nc^3) sum (1/2) |m_N|^2 / rho_N,   W_N = step^N(W_0(u0))

The gradient is one adjoint run (gkm_adjoint_run: checkpoints every `every`
steps, each segment recomputed) followed by the transpose of the
initial-condition map.  KineticEnergy wraps it as a torch.autograd.Function.
"""
import numpy as np

from . import _lib as gl
from . import ic as icm


def kinetic_energy(W):
    """Volume-averaged (1/2) rho |u|^2 of a padded state."""
    nc = W.shape[-1] - 6
    inner = (slice(3, nc + 3),) * 3
    rho = W[0][inner]
    m2 = W[1][inner] ** 2 + W[2][inner] ** 2 + W[3][inner] ** 2
    return float(np.sum(0.5 * m2 / rho) / nc ** 3)


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

    def objective(self, u):
        """J(u) by a forward run."""
        with self.context() as c:
            c.upload(initial_state(u, self.rho, self.p, self.params["gam"]))
            for _ in range(self.nsteps):
                c.step(self.dt)
            return kinetic_energy(c.download())

    def gradient(self, u):
        """(J(u), dJ/du) by a forward run and an adjoint run."""
        W0 = initial_state(u, self.rho, self.p, self.params["gam"])
        with self.context() as c:
            c.upload(W0)
            for _ in range(self.nsteps):
                c.step(self.dt)
            WN = c.download()
            J = kinetic_energy(WN)
            c.upload(W0)
            lam0 = c.adjoint_run(self.dt, self.nsteps, self.every, kinetic_energy_adjoint(WN))
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
