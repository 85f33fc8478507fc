"""Seeded test states shared by the CPU and GPU tests."""
import numpy as np

import pyoracle as po
from dns_gkm_lpt_amd import ic as icm


def make_case(nc, seed=0, Mt=0.3, Re=30.0, k0=4.0, perturb=0.05):
    """IC of the reference recipe plus seeded density/pressure perturbations
    (the IC has uniform rho and p; perturbing them exercises every WENO and
    flux branch from the first step)."""
    ic = icm.make_ic(nc, k0=k0, Mt=Mt, Re=Re, seed=seed)
    rho, p = ic["rho"].copy(), ic["p"].copy()
    if perturb:
        rng = np.random.default_rng(seed + 17)
        rho *= 1 + perturb * rng.uniform(-1, 1, rho.shape)
        p *= 1 + perturb * rng.uniform(-1, 1, p.shape)
    W = icm.prim_to_cons(ic["u"], ic["v"], ic["w"], rho, p, gam=ic["params"]["gam"])
    P = ic["params"]
    gas = po.make_gas(nc, gam=P["gam"], mu0=P["mu0"], p0=P["p0"], den0=P["rho0"])
    return W, P, gas


def random_faces(n, seed=0, jump=0.1):
    """n random face inputs (WL, WR, Wl, Wr, 6 slopes) of the flux() signature
    around the reference's dimensional state (rho ~ 1, p ~ 86100 Pa)."""
    rng = np.random.default_rng(seed)

    def state(k):
        rho = 1 + jump * rng.uniform(-1, 1, k)
        u, v, w = (60 * rng.standard_normal(k) for _ in range(3))
        p = 86100 * (1 + jump * rng.uniform(-1, 1, k))
        return np.stack([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)], 1)

    WL, WR = state(n), state(n)
    Wl = WL + 0.3 * (WR - WL) + 0.02 * rng.standard_normal((n, 5)) * np.abs(WL)
    Wr = WR - 0.3 * (WR - WL) + 0.02 * rng.standard_normal((n, 5)) * np.abs(WR)
    D = rng.standard_normal((n, 6, 5)) * np.abs(WL)[:, None, :] * 0.5
    X = np.concatenate([WL, WR, Wl, Wr, D.reshape(n, 30)], 1)
    return np.ascontiguousarray(X)


def random_stencils(n, seed=0, jump=0.1):
    rng = np.random.default_rng(seed)
    rho = 1 + jump * rng.uniform(-1, 1, (n, 6))
    u, v, w = (60 * rng.standard_normal((n, 6)) for _ in range(3))
    p = 86100 * (1 + jump * rng.uniform(-1, 1, (n, 6)))
    S = np.stack([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)], 2)
    return np.ascontiguousarray(S.reshape(n, 30))


def rel_err(a, b):
    """max |a - b| / max |b| per leading component."""
    a = np.asarray(a)
    b = np.asarray(b)
    return np.array([np.max(np.abs(a[q] - b[q])) / max(np.max(np.abs(b[q])), 1e-300) for q in range(a.shape[0])])


def torus_dist(a, b):
    d = np.abs(a - b) % (2 * np.pi)
    return np.minimum(d, 2 * np.pi - d)
