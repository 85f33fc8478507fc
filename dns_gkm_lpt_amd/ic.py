"""Seeded initial conditions for decaying isotropic turbulence, and the
reference's binary input formats.

The recipe follows InitialConditions/init_DNS.m of the reference:
  * random velocity components in physical space -> FFT (init_DNS.m:28-32),
  * solenoidal projection u_k -= (k.u_k) k / |k|^2 (init_DNS.m:35-43),
  * shell spectrum E(k) = A k^4 exp(-2 (k/k0)^2) imposed on 1 <= |k| <= 8 with
    |u_k|^2 = E(|k|) / (4 pi k^2) (init_DNS.m:17-18, 47-63),
  * scaled to the turbulent Mach number Mt (init_DNS.m:78-92),
  * viscosity from Re_lambda, mu = 2 ek0 sqrt(5 / (3 <A_ij A_ij>)) / Re with
    second-order periodic central differences (init_DNS.m:96-120),
  * eddy time lambda / u_rms (init_DNS.m:124-126),
  * uniform density rho0 = 1 and pressure p0 = rho0 R T0, T0 = 300 K, gamma = 7/5.
Two deliberate departures: the random field is seeded Gaussian (numpy
default_rng) instead of MATLAB's unseeded uniform rand, so runs are
reproducible, and the spectrum uses |u_k|^2 where init_DNS.m:54 squares the
complex values.

Arrays are indexed [y, x, z] (j, i, k), z fastest, as the solver stores them.
"""
import os
import struct

import numpy as np

try:  # multi-threaded FFTs when scipy is present; numpy's otherwise
    import scipy.fft as _fft

    def _rfftn(a):
        return _fft.rfftn(a, workers=-1)

    def _irfftn(a, s):
        return _fft.irfftn(a, s=s, workers=-1)
except ImportError:  # pragma: no cover
    def _rfftn(a):
        return np.fft.rfftn(a)

    def _irfftn(a, s):
        return np.fft.irfftn(a, s=s)

R_GAS = 287.0
GAMMA = 7.0 / 5.0
T0_REF = 300.0
RHO0 = 1.0


def _kgrid(n):
    k = np.fft.fftfreq(n, d=1.0 / n)
    kz = np.fft.rfftfreq(n, d=1.0 / n)
    return k[:, None, None], k[None, :, None], kz[None, None, :]


def make_ic(nc, k0=4.0, Mt=0.3, Re=30.0, seed=0, kcmin=1.0, kcmax=8.0):
    """Solenoidal random velocity with E(k) ~ k^4 exp(-2 (k/k0)^2).

    Returns a dict with u, v, w, rho, p (each nc^3, [y, x, z]) and the params
    record (gam, rho0, Mt, Re, mu0, p0, t0) the reference's GetParams reads.
    """
    rng = np.random.default_rng(seed)
    ky, kx, kz = _kgrid(nc)
    # u <-> kx, v <-> ky, w <-> kz (init_DNS.m:36-39 with MATLAB dims (y, x, z))
    uk = _rfftn(rng.standard_normal((nc, nc, nc)))
    vk = _rfftn(rng.standard_normal((nc, nc, nc)))
    wk = _rfftn(rng.standard_normal((nc, nc, nc)))
    ksq = np.maximum(kx * kx + ky * ky + kz * kz, 1e-14)
    kdot = kx * uk + ky * vk + kz * wk
    uk -= kdot * kx / ksq
    vk -= kdot * ky / ksq
    wk -= kdot * kz / ksq
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    A = np.sqrt(2 / np.pi) * (16.0 / 3.0) * Mt ** 2 / k0 ** 5
    shell = (kk >= kcmin) & (kk <= kcmax)
    ee = np.where(shell, A * kk ** 4 * np.exp(-2 * (kk / k0) ** 2) / (4 * np.pi * np.maximum(kk, 1) ** 2), 0.0)
    umag2 = np.abs(uk) ** 2 + np.abs(vk) ** 2 + np.abs(wk) ** 2
    fac = np.where(shell & (umag2 > 0), np.sqrt(ee / np.where(umag2 > 0, umag2, 1.0)), 0.0)
    u = _irfftn(uk * fac, s=(nc, nc, nc))
    v = _irfftn(vk * fac, s=(nc, nc, nc))
    w = _irfftn(wk * fac, s=(nc, nc, nc))
    del uk, vk, wk, kdot, fac

    p0 = RHO0 * R_GAS * T0_REF
    ek0 = 0.5 * np.mean(u * u + v * v + w * w)
    M = np.sqrt(2 * ek0) / np.sqrt(GAMMA * p0 / RHO0)
    eta = Mt / M
    u *= eta
    v *= eta
    w *= eta
    ek0 = 0.5 * np.mean(u * u + v * v + w * w)
    Mt_out = np.sqrt(2 * ek0) / np.sqrt(GAMMA * p0 / RHO0)

    dx = 2 * np.pi / nc

    def d(f, ax):
        return 0.5 * (np.roll(f, -1, axis=ax) - np.roll(f, 1, axis=ax)) / dx

    # axis 0 = y, 1 = x, 2 = z
    sq = 0.0
    for f in (u, v, w):
        for ax in (0, 1, 2):
            g = d(f, ax)
            sq += np.mean(g * g)
    mu = 2 * ek0 * np.sqrt(5 / (3 * sq)) / Re
    lam = np.sqrt(10 * ek0 / sq)
    urms = np.sqrt(2 * ek0 / 3)
    t0 = lam / urms
    rho = np.full((nc, nc, nc), RHO0)
    p = np.full((nc, nc, nc), p0)
    return dict(u=u, v=v, w=w, rho=rho, p=p, nc=nc,
                params=dict(gam=GAMMA, rho0=RHO0, Mt=float(Mt_out), Re=float(Re), mu0=float(mu),
                            p0=float(p0), t0=float(t0)))


def ensemble_members(seed=4):
    """The batched-ensemble workload: 64 independent realizations, 24 at 32^3,
    24 at 48^3 and 16 at 64^3, member m with its own seed 100 + m and k0, Mt,
    Re_lambda drawn uniformly from [3, 6], [0.1, 0.6], [20, 60] (seeded by
    `seed`).  Returns [(nc, params, W padded, tracer seed)]."""
    rng = np.random.default_rng(seed)
    ncs = [32] * 24 + [48] * 24 + [64] * 16
    out = []
    for m, nc in enumerate(ncs):
        k0, Mt, Re = rng.uniform(3, 6), rng.uniform(0.1, 0.6), rng.uniform(20, 60)
        ic = make_ic(nc, k0=k0, Mt=Mt, Re=Re, seed=100 + m)
        out.append((nc, ic["params"], prim_to_cons(ic["u"], ic["v"], ic["w"], ic["rho"], ic["p"]), 100 + m))
    return out


def uniform_particles(n, seed=0):
    """Tracer positions uniform in [0, 2 pi)^3 (Part_Gen.m); columns x, y, z."""
    rng = np.random.default_rng(seed + 1_000_003)
    return rng.random((n, 3)) * (2 * np.pi)


def prim_to_cons(u, v, w, rho, p, gam=GAMMA):
    """Primitive -> conservative as Initialize_DHIT does (FlowField.cpp:89-105),
    into the padded layout (5, nt, nt, nt) with periodic ghost layers."""
    nc = u.shape[0]
    nt = nc + 6
    W = np.empty((5, nt, nt, nt))
    Vsqr = u * u + v * v + w * w
    inner = (slice(3, nt - 3),) * 3
    W[0][inner] = rho
    W[1][inner] = u * rho
    W[2][inner] = v * rho
    W[3][inner] = w * rho
    W[4][inner] = rho * (p / (rho * (gam - 1.0)) + 0.5 * Vsqr)
    return fill_halo(W)


def fill_halo(W):
    """Periodic ghost layers: ghost(i) = cell(((i - 3) mod nc) + 3) on each axis."""
    out = np.empty_like(W)
    for q in range(W.shape[0]):
        out[q] = np.pad(W[q][3:-3, 3:-3, 3:-3], 3, mode="wrap")
    return out


# ----------------------------------------------------------------------------
# Reference binary formats
# ----------------------------------------------------------------------------

def write_params(path, nc, params):
    """InitialConditions/params: int32 N then doubles gam, rho0, Mt, Re, mu, p0,
    eddy time (init_DNS.m:165-175, read by FlowField.cpp:16-40)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<i", int(nc)))
        f.write(struct.pack("<7d", params["gam"], params["rho0"], params["Mt"], params["Re"],
                            params["mu0"], params["p0"], params["t0"]))


def read_params(path):
    with open(path, "rb") as f:
        (n,) = struct.unpack("<i", f.read(4))
        vals = struct.unpack("<7d", f.read(56))
    keys = ("gam", "rho0", "Mt", "Re", "mu0", "p0", "t0")
    return n, dict(zip(keys, vals))


def write_init(path, ic):
    """Init<a><b><c>: u, v, w, rho, p, each in MATLAB column-major order over
    (y, x, z), i.e. element (j, i, k) at k*n*n + i*n + j (FlowField.cpp:97)."""
    with open(path, "wb") as f:
        for key in ("u", "v", "w", "rho", "p"):
            f.write(np.ascontiguousarray(ic[key].transpose(2, 1, 0), dtype="<f8").tobytes())


def read_init(path, nc):
    raw = np.fromfile(path, dtype="<f8")
    if raw.size != 5 * nc ** 3:
        raise ValueError(f"{path}: expected {5 * nc ** 3} doubles, found {raw.size}")
    fields = raw.reshape(5, nc, nc, nc)  # (field, k, i, j)
    return {key: np.ascontiguousarray(fields[q].transpose(2, 1, 0))
            for q, key in enumerate(("u", "v", "w", "rho", "p"))}


def write_particles(dirpath, xyz):
    """unirand (x[Np], y[Np], z[Np] doubles) and particleCount (int32 Np);
    Part_Gen.m, read by update.cpp:1296-1302 and FlowField.cpp:54-57."""
    xyz = np.asarray(xyz, dtype="<f8")
    with open(os.path.join(dirpath, "unirand"), "wb") as f:
        for c in range(3):
            f.write(np.ascontiguousarray(xyz[:, c]).tobytes())
    with open(os.path.join(dirpath, "particleCount"), "wb") as f:
        f.write(struct.pack("<i", xyz.shape[0]))


def read_particles(dirpath):
    with open(os.path.join(dirpath, "particleCount"), "rb") as f:
        (n,) = struct.unpack("<i", f.read(4))
    raw = np.fromfile(os.path.join(dirpath, "unirand"), dtype="<f8", count=3 * n)
    return np.ascontiguousarray(raw.reshape(3, n).T)


def write_case(dirpath, ic, xyz):
    """All reference inputs for a 1x1x1 decomposition: params, Init000,
    particleCount, unirand."""
    os.makedirs(dirpath, exist_ok=True)
    write_params(os.path.join(dirpath, "params"), ic["nc"], ic["params"])
    write_init(os.path.join(dirpath, "Init000"), ic)
    write_particles(dirpath, xyz)
