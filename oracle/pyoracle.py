"""ctypes binding of the CPU oracle (oracle/libgkm_oracle.so) and of the
reference build (oracle/_ref/libgkm_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package never
imports this module.
"""
import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libgkm_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libgkm_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Gas(C.Structure):
    _fields_ = [("nc", C.c_int), ("K", C.c_int), ("gam", C.c_double), ("R", C.c_double),
                ("prn", C.c_double), ("mu0", C.c_double), ("T0", C.c_double),
                ("dx", C.c_double), ("eps", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("tke", C.c_double), ("diss", C.c_double), ("sijsij", C.c_double),
                ("skew", C.c_double * 3), ("kurt", C.c_double * 3), ("prms", C.c_double),
                ("trms", C.c_double), ("vrms", C.c_double), ("pmean", C.c_double),
                ("tmean", C.c_double), ("vmean", C.c_double), ("ens", C.c_double)]


class PStats(C.Structure):
    _fields_ = [("tke", C.c_double), ("sijsij", C.c_double), ("skew_u", C.c_double),
                ("skew_w", C.c_double)]


def make_gas(nc, gam=1.4, mu0=1e-3, p0=86100.0, den0=1.0, R=287.0, prn=0.70, eps=1e-10):
    """Gas constants exactly as GetParams derives them (FlowField.cpp:41-50)."""
    pi = np.arctan(1.0) * 4.0
    return Gas(nc=nc, K=int((5.0 - 3.0 * gam) / (gam - 1.0)), gam=gam, R=R, prn=prn, mu0=mu0,
               T0=p0 / (R * den0), dx=2.0 * pi / nc, eps=eps)


def build(ref=False):
    subprocess.check_call(["make", "-s", "-C", HERE] + (["ref"] if ref else []))


_lib = None
_ref = None


def _configure(L):
    G = C.POINTER(Gas)
    L.orc_weno_batch.argtypes = [C.c_long, _dp, G, _dp]
    L.orc_flux_batch.argtypes = [C.c_long, _dp, _dp, C.c_double, G, _dp]
    L.orc_tau.argtypes = [_dp, G]
    L.orc_tau.restype = C.c_double
    L.orc_halo_fill.argtypes = [C.c_int, _dp]
    L.orc_fluxes.argtypes = [C.c_int, _dp, C.c_double, G, C.c_int, _dp, _dp, _dp]
    L.orc_update.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double]
    L.orc_eulerian_step.argtypes = [C.c_int, _dp, C.c_double, G, C.c_int]
    L.orc_cfl_dt.argtypes = [C.c_int, _dp, C.c_double, G]
    L.orc_cfl_dt.restype = C.c_double
    L.orc_spline_coefs.argtypes = [C.c_int, _dp, _dp]
    L.orc_spline_eval.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_velocity_coefs.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    L.orc_particle_sample.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_long, _dp, _dp,
                                      _dp, _dp, _dp]
    L.orc_particle_heun.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_long, _dp, _dp,
                                    C.c_double]
    L.orc_lap_field.argtypes = [C.c_int, _dp, C.c_double, _dp]
    L.orc_particle_props.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_long, _dp, _dp]
    L.orc_field_stats.argtypes = [C.c_int, _dp, _dp, _dp, _dp, G, C.c_double, C.c_double,
                                  C.c_double, C.POINTER(Stats)]
    L.orc_particle_stats.argtypes = [C.c_long, C.c_long, _dp, _dp, _dp, _dp, C.POINTER(PStats)]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _configure(C.CDLL(LIB))
    return _lib


@contextlib.contextmanager
def using(path):
    """Route every helper below through the oracle library at `path` (a
    variant built from a modified copy of gkm_oracle.c, for the tests that
    check a pin fails on a planted mistake)."""
    global _lib
    saved = _lib
    _lib = _configure(C.CDLL(path))
    try:
        yield
    finally:
        _lib = saved


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    """The reference's own code (GKMflux.cpp, update.cpp CPU step); None if not built."""
    global _ref
    if _ref is None:
        if not ref_available():
            return None
        L = C.CDLL(REF_LIB)
        L.ref_set_params.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double]
        L.ref_flux_batch.argtypes = [C.c_long, _dp, _dp, C.c_double, C.c_double, _dp]
        L.ref_weno_batch.argtypes = [C.c_long, _dp, _dp]
        L.ref_eulerian.argtypes = [_dp]
        L.ref_get_dx.restype = C.c_double
        L.ref_get_T0.restype = C.c_double
        _ref = L
    return _ref


# ----------------------------------------------------------------------------
# numpy-level helpers
# ----------------------------------------------------------------------------

def weno(S, gas):
    S = np.ascontiguousarray(S, dtype=np.float64).reshape(-1, 30)
    out = np.empty((S.shape[0], 10))
    lib().orc_weno_batch(S.shape[0], S, C.byref(gas), out)
    return out


def flux(X, tau, dt, gas):
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 50)
    tau = np.ascontiguousarray(tau, dtype=np.float64).reshape(-1)
    F = np.empty((X.shape[0], 5))
    lib().orc_flux_batch(X.shape[0], X, tau, dt, C.byref(gas), F)
    return F


def tau_of(WL, gas):
    return lib().orc_tau(np.ascontiguousarray(WL, dtype=np.float64), C.byref(gas))


def halo_fill(W):
    W = np.ascontiguousarray(W, dtype=np.float64)
    nc = W.shape[-1] - 6
    lib().orc_halo_fill(nc, W)
    return W


def fluxes(W, dt, gas, nthreads=1):
    W = np.ascontiguousarray(W, dtype=np.float64)
    nc = W.shape[-1] - 6
    F = [np.zeros_like(W) for _ in range(3)]
    lib().orc_fluxes(nc, W, dt, C.byref(gas), nthreads, F[0], F[1], F[2])
    return F


def eulerian_step(W, dt, gas, nthreads=1):
    """evolveEulerian + periodic halo on a copy of W (5, nt, nt, nt)."""
    W = np.array(W, dtype=np.float64, order="C", copy=True)
    nc = W.shape[-1] - 6
    lib().orc_eulerian_step(nc, W, dt, C.byref(gas), nthreads)
    return W


def cfl_dt(W, cfl, gas):
    W = np.ascontiguousarray(W, dtype=np.float64)
    return lib().orc_cfl_dt(W.shape[-1] - 6, W, cfl, C.byref(gas))


def spline_coefs(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    nc = f.shape[-1] - 6
    c = np.empty((nc + 8,) * 3)
    lib().orc_spline_coefs(nc, f, c)
    return c


def spline_eval(coefs, dx, y, x, z):
    nc = coefs.shape[-1] - 8
    v = C.c_double()
    g = (C.c_double * 3)()
    h = (C.c_double * 9)()
    lib().orc_spline_eval(nc, np.ascontiguousarray(coefs), dx, y, x, z, C.byref(v), g, h)
    return v.value, np.array(g[:]), np.array(h[:]).reshape(3, 3)


def velocity_coefs(W):
    W = np.ascontiguousarray(W, dtype=np.float64)
    nc = W.shape[-1] - 6
    cs = [np.empty((nc + 8,) * 3) for _ in range(3)]
    lib().orc_velocity_coefs(nc, W, cs[0], cs[1], cs[2])
    return cs


def particle_sample(cs, dx, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    nc = cs[0].shape[-1] - 8
    U, gU, gV, gW = (np.empty((n, 3)) for _ in range(4))
    lib().orc_particle_sample(nc, cs[0], cs[1], cs[2], dx, n, pos, U, gU, gV, gW)
    return U, gU, gV, gW


def lap_field(coefs, dx):
    nc = coefs.shape[-1] - 8
    out = np.empty((nc + 6,) * 3)
    lib().orc_lap_field(nc, np.ascontiguousarray(coefs), dx, out)
    return out


def particle_props(W, gam, dx, pos):
    """np x 41 Particle members (getParticleProp) from padded W."""
    W = np.ascontiguousarray(W, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    nc = W.shape[-1] - 6
    out = np.empty((pos.shape[0], 41))
    lib().orc_particle_props(nc, W, gam, dx, pos.shape[0], pos, out)
    return out


def particle_heun(cs, dx, pos, U, dt):
    pos = np.array(pos, dtype=np.float64, order="C", copy=True)
    U = np.ascontiguousarray(U, dtype=np.float64)
    nc = cs[0].shape[-1] - 8
    lib().orc_particle_heun(nc, cs[0], cs[1], cs[2], dx, pos.shape[0], pos, U, dt)
    return pos


def field_stats(W, cs, gas, p0, Mt, den0=1.0):
    W = np.ascontiguousarray(W, dtype=np.float64)
    s = Stats()
    lib().orc_field_stats(W.shape[-1] - 6, W, cs[0], cs[1], cs[2], C.byref(gas), p0, Mt, den0,
                          C.byref(s))
    return s


def particle_stats(U, gU, gV, gW, n_total=None):
    n = U.shape[0]
    s = PStats()
    lib().orc_particle_stats(n, n if n_total is None else n_total, np.ascontiguousarray(U),
                             np.ascontiguousarray(gU), np.ascontiguousarray(gV),
                             np.ascontiguousarray(gW), C.byref(s))
    return s
