"""Tangent-linear and discrete adjoint of the fixed-dt FP64 step.

Pins (none retypes the adjoint): the dot-product identity
<lam_N, dW_N> = <lam_0, dW_0> between the tangent-linear and the adjoint runs;
the tangent-linear run against central finite differences of the production
step; the gradient of the final kinetic energy against central finite
differences and against the tangent-linear directional derivative along seeded
random unit directions.  Finite differences use h = 1e-6 relative to the norm
of the perturbed field (an absolute 1e-6 on a unit direction would change J by
~1e-12 of itself, below what FP64 differences resolve).

The step is piecewise smooth: the WENO smoothness-indicator maximum and the
|p_l - p_r| of the collision time switch branch on surfaces in state space.
A perturbation of size h crosses such a surface in O(h) of the cells, so a
central difference carries an O(h) error there that the derivative does not
have.  Random directions are therefore drawn band-limited (|k| <= 4, the
initial condition's band), where the measured FD error is ~1e-8; for
white-noise directions the FD error falls linearly with h toward the adjoint
value (checked below), and field-wise comparisons use the median cell.
Stated tolerances: dot-product identity 1e-10; tangent-linear fields vs
finite differences 1e-8 (median cell, relative to the field's largest
magnitude) and 1e-4 (relative L2); kinetic-energy gradient vs finite
differences 1e-6; tangent-linear vs adjoint directional derivative 1e-10.
"""
import numpy as np
import pytest

from cases import make_case
from dns_gkm_lpt_amd import _lib as gl
from dns_gkm_lpt_amd import adjoint as adj
from dns_gkm_lpt_amd import ic as icm

pytestmark = pytest.mark.gpu

NC, NSTEPS, EVERY = 32, 6, 4


def interior(a):
    nc = a.shape[-1] - 6
    return a[(slice(None),) + (slice(3, nc + 3),) * 3]


def random_state_vector(shape, seed, scale):
    rng = np.random.default_rng(seed)
    v = np.zeros(shape)
    nc = shape[-1] - 6
    inner = (slice(3, nc + 3),) * 3
    for q in range(5):
        v[q][inner] = rng.standard_normal((nc,) * 3) * scale[q]
    return icm.fill_halo(v)


def test_dot_product_identity():
    W, P, gas = make_case(NC, seed=21)
    with gl.Context(NC, P, precision=gl.GKM_FP64) as c:
        c.upload(W)
        dt = c.cfl_dt()
        scale = np.abs(interior(W)).max(axis=(1, 2, 3))
        dW0 = random_state_vector(W.shape, 1, 1e-3 * scale)
        lamN = random_state_vector(W.shape, 2, 1.0 / scale)
        c.upload(W)
        dWN = c.tl_run(dt, NSTEPS, dW0)
        c.upload(W)
        lam0 = c.adjoint_run(dt, NSTEPS, EVERY, lamN)
    lhs = float(np.sum(interior(lamN) * interior(dWN)))
    rhs = float(np.sum(interior(lam0) * interior(dW0)))
    print("dot-product identity", lhs, rhs, abs(lhs - rhs) / abs(lhs), flush=True)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def smooth_state(steps=20):
    """The reference's initial condition (uniform rho and p) advanced a few
    steps, so that density and pressure vary smoothly: at uniform pressure the
    collision time's |p_l - p_r| term sits on its kink at every face."""
    W, P, gas = make_case(NC, seed=22, perturb=0.0)
    with gl.Context(NC, P, precision=gl.GKM_FP64) as c:
        c.upload(W)
        c.run(steps)
        return c.download(), P


def test_tangent_linear_matches_finite_differences():
    W, P = smooth_state()
    scale = np.abs(interior(W)).max(axis=(1, 2, 3))
    x = np.arange(NC) * 2 * np.pi / NC
    dW0 = np.zeros_like(W)
    inner = (slice(3, NC + 3),) * 3
    for q in range(5):
        dW0[q][inner] = 1e-2 * scale[q] * (np.sin(x[:, None, None] + q) * np.cos(2 * x[None, :, None]) *
                                            np.sin(x[None, None, :] + 2 * q))
    dW0 = icm.fill_halo(dW0)
    with gl.Context(NC, P, precision=gl.GKM_FP64) as c:
        c.upload(W)
        dt = c.cfl_dt()
        c.upload(W)
        dWN = c.tl_run(dt, NSTEPS, dW0)
        h = 1e-6 / np.sqrt(np.mean(interior(dW0 / scale[:, None, None, None]) ** 2))
        outs = []
        for sgn in (1.0, -1.0):
            c.upload(W + sgn * h * dW0)
            for _ in range(NSTEPS):
                c.step(dt)
            outs.append(c.download())
    fd = (outs[0] - outs[1]) / (2 * h)
    d = interior(fd - dWN)
    l2 = np.sqrt(np.sum(d ** 2, axis=(1, 2, 3)) / np.sum(interior(dWN) ** 2, axis=(1, 2, 3)))
    mx = np.max(np.abs(d), axis=(1, 2, 3)) / np.max(np.abs(interior(dWN)), axis=(1, 2, 3))
    med = np.median(np.abs(d), axis=(1, 2, 3)) / np.max(np.abs(interior(dWN)), axis=(1, 2, 3))
    print("tangent-linear vs FD: median", med, "relative L2", l2, "max", mx, flush=True)
    assert np.all(med < 1e-8), med
    assert np.all(l2 < 1e-4), l2


def problem(nc=NC, seed=23, Mt=0.3, Re=40.0, nsteps=NSTEPS, every=EVERY):
    ic = icm.make_ic(nc, k0=4.0, Mt=Mt, Re=Re, seed=seed)
    P = ic["params"]
    u = np.stack([ic["u"], ic["v"], ic["w"]])
    W0 = adj.initial_state(u, P["rho0"], P["p0"], P["gam"])
    with gl.Context(nc, P, precision=gl.GKM_FP64, cfl=0.5) as c:
        c.upload(W0)
        dt = c.cfl_dt()  # fixed dt from the initial CFL 0.5
    return adj.Problem(nc, P, dt, nsteps, every), u


def band_limited_direction(shape, seed, kmax=4):
    """A seeded random unit vector field with Fourier modes |k| <= kmax only."""
    rng = np.random.default_rng(seed)
    nc = shape[-1]
    k = np.fft.fftfreq(nc, 1.0 / nc)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    d = np.stack([np.real(np.fft.ifftn(np.fft.fftn(rng.standard_normal((nc,) * 3)) * (kk <= kmax)))
                  for _ in range(shape[0])])
    return d / np.linalg.norm(d)


def check_gradient(pb, u, ndir, seed):
    J, g = pb.gradient(u)
    h = 1e-6 * np.linalg.norm(u)
    worst_fd = worst_tl = 0.0
    for k in range(ndir):
        d = band_limited_direction(u.shape, seed + k)
        ad = float(np.sum(g * d))
        fd = (pb.objective(u + h * d) - pb.objective(u - h * d)) / (2 * h)
        tl = pb.directional(u, d)
        e_fd, e_tl = abs(fd - ad) / abs(ad), abs(tl - ad) / abs(ad)
        print(f"direction {k}: adjoint {ad:.12e} FD {fd:.12e} TL {tl:.12e} rel {e_fd:.2e} {e_tl:.2e}", flush=True)
        worst_fd, worst_tl = max(worst_fd, e_fd), max(worst_tl, e_tl)
    return J, g, worst_fd, worst_tl


def test_kinetic_energy_gradient_small():
    pb, u = problem()
    J, g, e_fd, e_tl = check_gradient(pb, u, 3, 5)
    assert e_fd < 1e-6 and e_tl < 1e-10, (e_fd, e_tl)
    # white-noise direction: the central difference converges to the adjoint as h -> 0
    d = np.random.default_rng(9).standard_normal(u.shape)
    d /= np.linalg.norm(d)
    ad = float(np.sum(g * d))
    errs = []
    for hr in (1e-5, 1e-7):
        h = hr * np.linalg.norm(u)
        fd = (pb.objective(u + h * d) - pb.objective(u - h * d)) / (2 * h)
        errs.append(abs(fd - ad) / abs(ad))
    print("white-noise direction: FD error at h = 1e-5, 1e-7:", errs, flush=True)
    assert errs[1] < 1e-6 and errs[1] < errs[0]


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    pb, u = problem(nsteps=3, every=2)
    ut = torch.tensor(u, dtype=torch.float64, requires_grad=True)
    J = adj.KineticEnergy(pb).apply(ut)
    J.backward()
    Jref, g = pb.gradient(u)
    assert abs(float(J) - Jref) <= 1e-14 * abs(Jref)
    assert np.array_equal(ut.grad.numpy(), g)
