"""ctypes binding of libdnsgkm.so (the C ABI of include/dnsgkm.h).

The HIP library is the only compute path: if libdnsgkm.so is missing this
module raises instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdnsgkm.so")

GKM_FP64 = 0
GKM_MIXED = 1
ERRORS = {1: "GKM_EINVAL", 2: "GKM_EHIP", 3: "GKM_ENOMEM", 4: "GKM_ESTATE", 5: "GKM_ENOTREADY"}

# every entry point include/dnsgkm.h declares
EXPORTS = [
    "gkm_create", "gkm_destroy", "gkm_last_error", "gkm_get_K", "gkm_set_compensation", "gkm_get_compensation",
    "gkm_download_carry", "gkm_upload_state", "gkm_download_state",
    "gkm_upload_state_packed", "gkm_download_state_packed", "gkm_step", "gkm_cfl_dt", "gkm_run", "gkm_sync",
    "gkm_sweep_flux_diff", "gkm_set_sweep_tiles", "gkm_get_sweep_tiles", "gkm_autotune", "gkm_step_noalloc",
    "gkm_tl_run", "gkm_adjoint_run", "gkm_adjoint_forward", "gkm_adjoint_backward",
    "gkm_ensemble_create", "gkm_ensemble_destroy", "gkm_ensemble_last_error", "gkm_ensemble_member",
    "gkm_ensemble_set_tiles", "gkm_ensemble_autotune", "gkm_ensemble_run", "gkm_ensemble_dt",
    "gkm_ensemble_set_persistent", "gkm_ensemble_persistent_profile",
    "gkm_stats_compute", "gkm_timing_enable", "gkm_timing_read", "gkm_flux_batch", "gkm_weno_batch",
    "lpt_upload", "lpt_set_reference_drop", "lpt_count", "lpt_download", "lpt_sample", "lpt_stats", "lpt_properties", "lpt_coupled_step", "lpt_heun",
]


# columns of lpt_properties: the reference's Particle members (ParticleTrack.h:21-22),
# vectors over (y, x, z), hessP row-major over (y, x, z)
PROP_NAMES = (["DEN"] + [f"U{a}" for a in range(3)] + ["P"] +
              [f"{v}{a}" for v in ("gradD", "gradU", "gradV", "gradW", "gradP") for a in range(3)] +
              ["lapU", "lapV", "lapW"] + [f"hessP{a}" for a in range(9)] +
              [f"{v}{a}" for v in ("gradLapU", "gradLapV", "gradLapW") for a in range(3)])
PROP_WIDTH = len(PROP_NAMES)


class Config(C.Structure):
    _fields_ = [("nc", C.c_int), ("precision", C.c_int), ("gam", C.c_double), ("rho0", C.c_double),
                ("Mt", C.c_double), ("Re", C.c_double), ("mu0", C.c_double), ("p0", C.c_double),
                ("t0", C.c_double), ("cfl", C.c_double), ("prandtl", C.c_double), ("device", C.c_int),
                ("check_state", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("tke", C.c_double), ("diss", C.c_double), ("sijsij", C.c_double), ("skew", C.c_double * 3),
                ("kurt", C.c_double * 3), ("prms", C.c_double), ("trms", C.c_double), ("vrms", C.c_double),
                ("pmean", C.c_double), ("tmean", C.c_double), ("vmean", C.c_double),
                ("enstrophy", C.c_double)]


class PStats(C.Structure):
    _fields_ = [("tke", C.c_double), ("sijsij", C.c_double), ("skew_u", C.c_double), ("skew_w", C.c_double)]


class GkmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run dns_gkm_lpt_amd.build.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.gkm_create.argtypes = [C.POINTER(Config), C.POINTER(P)]
        L.gkm_destroy.argtypes = [P]
        L.gkm_destroy.restype = None
        L.gkm_last_error.argtypes = [P]
        L.gkm_last_error.restype = C.c_char_p
        L.gkm_get_K.argtypes = [P]
        L.gkm_set_compensation.argtypes = [P, C.c_int]
        L.gkm_step_noalloc.argtypes = [C.POINTER(Config), C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                       C.c_char_p, C.c_int]
        L.gkm_get_compensation.argtypes = [P]
        L.gkm_download_carry.argtypes = [P, np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")]
        L.gkm_upload_state_packed.argtypes = [P, _dp]
        L.gkm_download_state_packed.argtypes = [P, _dp]
        L.gkm_upload_state.argtypes = [P, C.POINTER(C.c_void_p)]
        L.gkm_download_state.argtypes = [P, C.POINTER(C.c_void_p)]
        L.gkm_step.argtypes = [P, C.c_double]
        L.gkm_cfl_dt.argtypes = [P, C.POINTER(C.c_double)]
        L.gkm_run.argtypes = [P, C.c_int, C.c_void_p]
        L.gkm_sync.argtypes = [P]
        L.gkm_sweep_flux_diff.argtypes = [P, C.c_int, C.c_double, _dp]
        L.gkm_set_sweep_tiles.argtypes = [P, C.POINTER(C.c_int * 3)]
        L.gkm_get_sweep_tiles.argtypes = [P, C.POINTER(C.c_int * 3)]
        L.gkm_autotune.argtypes = [P, C.c_int, C.POINTER(C.c_int * 3), C.POINTER(C.c_double * 9)]
        L.gkm_tl_run.argtypes = [P, C.c_double, C.c_int, _dp, _dp]
        L.gkm_adjoint_run.argtypes = [P, C.c_double, C.c_int, C.c_int, _dp, _dp]
        L.gkm_adjoint_forward.argtypes = [P, C.c_double, C.c_int, C.c_int]
        L.gkm_adjoint_backward.argtypes = [P, _dp, _dp]
        L.gkm_ensemble_create.argtypes = [C.c_int, C.POINTER(Config), C.POINTER(P)]
        L.gkm_ensemble_destroy.argtypes = [P]
        L.gkm_ensemble_destroy.restype = None
        L.gkm_ensemble_last_error.argtypes = [P]
        L.gkm_ensemble_last_error.restype = C.c_char_p
        L.gkm_ensemble_member.argtypes = [P, C.c_int]
        L.gkm_ensemble_member.restype = P
        L.gkm_ensemble_set_tiles.argtypes = [P, C.POINTER(C.c_int * 3)]
        L.gkm_ensemble_autotune.argtypes = [P, C.c_int, C.POINTER(C.c_int * 3), C.POINTER(C.c_double * 9)]
        L.gkm_ensemble_run.argtypes = [P, C.c_int, C.c_int]
        L.gkm_ensemble_dt.argtypes = [P, _dp]
        L.gkm_ensemble_set_persistent.argtypes = [P, C.c_int]
        L.gkm_ensemble_persistent_profile.argtypes = [P, _dp, C.POINTER(C.c_int)]
        L.gkm_stats_compute.argtypes = [P, C.POINTER(Stats)]
        L.gkm_timing_enable.argtypes = [P, C.c_int]
        L.gkm_timing_read.argtypes = [P, C.POINTER(C.c_double * 8), C.POINTER(C.c_long * 8)]
        L.gkm_flux_batch.argtypes = [P, C.c_long, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp]
        L.gkm_weno_batch.argtypes = [P, C.c_long, _dp, C.c_int, _dp]
        L.lpt_upload.argtypes = [P, C.c_long, _dp]
        L.lpt_set_reference_drop.argtypes = [P, C.c_int]
        L.lpt_count.argtypes = [P, C.POINTER(C.c_long)]
        L.lpt_download.argtypes = [P, C.c_void_p, C.c_void_p]
        L.lpt_sample.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.lpt_stats.argtypes = [P, C.POINTER(PStats)]
        L.lpt_properties.argtypes = [P, _dp]
        L.lpt_coupled_step.argtypes = [P, C.c_double]
        L.lpt_heun.argtypes = [P, C.c_double]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_config(nc, params, cfl=0.5, precision=GKM_FP64, prandtl=0.70, device=0, check_state=1):
    return Config(nc=nc, precision=precision, gam=params["gam"], rho0=params.get("rho0", 1.0),
                  Mt=params.get("Mt", 0.3), Re=params.get("Re", 30.0), mu0=params["mu0"], p0=params["p0"],
                  t0=params.get("t0", 1.0), cfl=cfl, prandtl=prandtl, device=device, check_state=check_state)


class Context:
    """One resident GKM-WENO run on one GPU (wraps gkm_ctx)."""

    def __init__(self, nc, params, cfl=0.5, precision=GKM_FP64, prandtl=0.70, device=0, check_state=1):
        self.nc = nc
        self.nt = nc + 6
        self._owned = True
        cfg = make_config(nc, params, cfl, precision, prandtl, device, check_state)
        h = C.c_void_p()
        rc = lib().gkm_create(C.byref(cfg), C.byref(h))
        self._h = h
        if rc:
            msg = lib().gkm_last_error(h).decode() if h.value else "gkm_create failed"
            if h.value:
                lib().gkm_destroy(h)
            self._h = C.c_void_p()
            raise GkmError(rc, msg)

    def _chk(self, rc):
        if rc:
            raise GkmError(rc, lib().gkm_last_error(self._h).decode())

    @classmethod
    def _view(cls, handle, nc):
        """A context owned by something else (an ensemble member)."""
        c = cls.__new__(cls)
        c.nc, c.nt, c._owned, c._h = nc, nc + 6, False, C.c_void_p(handle)
        return c

    def close(self):
        if self._h and self._h.value and self._owned:
            lib().gkm_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def K(self):
        return lib().gkm_get_K(self._h)

    @property
    def compensated(self):
        """Whether the update carries its rounding error (compensated accumulation)."""
        return bool(lib().gkm_get_compensation(self._h))

    def set_compensation(self, on):
        self._chk(lib().gkm_set_compensation(self._h, 1 if on else 0))

    def download_carry(self):
        """The carried rounding error of the resident state (5, nc, nc, nc) float32,
        unpadded: the exact running sum is W + carry."""
        out = np.empty((5, self.nc, self.nc, self.nc), dtype=np.float32)
        self._chk(lib().gkm_download_carry(self._h, out))
        return out

    def upload(self, W):
        W = np.ascontiguousarray(W, dtype=np.float64)
        if W.shape != (5, self.nt, self.nt, self.nt):
            raise ValueError(f"state must be (5, {self.nt}, {self.nt}, {self.nt})")
        self._chk(lib().gkm_upload_state_packed(self._h, W))

    def download(self):
        W = np.empty((5, self.nt, self.nt, self.nt))
        self._chk(lib().gkm_download_state_packed(self._h, W))
        return W

    def step(self, dt):
        self._chk(lib().gkm_step(self._h, dt))

    def cfl_dt(self):
        d = C.c_double()
        self._chk(lib().gkm_cfl_dt(self._h, C.byref(d)))
        return d.value

    def run(self, nsteps, record_dt=False):
        hist = np.empty(max(nsteps, 1)) if record_dt else None
        self._chk(lib().gkm_run(self._h, nsteps, _ptr(hist)))
        return hist[:nsteps] if record_dt else None

    def sweep_flux_diff(self, direction, dt):
        out = np.empty((5, self.nc, self.nc, self.nc))
        self._chk(lib().gkm_sweep_flux_diff(self._h, direction, dt, out))
        return out

    def sync(self):
        self._chk(lib().gkm_sync(self._h))

    def tl_run(self, dt, nsteps, dW0):
        """Tangent-linear run from the resident state (advanced by nsteps):
        returns dW_N for the perturbation dW0 (padded layout)."""
        dW0 = np.ascontiguousarray(dW0, dtype=np.float64)
        out = np.empty((5, self.nt, self.nt, self.nt))
        self._chk(lib().gkm_tl_run(self._h, dt, nsteps, dW0, out))
        return out

    def adjoint_run(self, dt, nsteps, every, lamN):
        """Adjoint run from the resident state (advanced by nsteps): returns
        lam_0 = (dW_N/dW_0)^T lam_N (padded layout)."""
        lamN = np.ascontiguousarray(lamN, dtype=np.float64)
        out = np.empty((5, self.nt, self.nt, self.nt))
        self._chk(lib().gkm_adjoint_run(self._h, dt, nsteps, every, lamN, out))
        return out

    def adjoint_forward(self, dt, nsteps, every):
        """The forward part of adjoint_run: the checkpointed trajectory from the
        resident state, which ends at W_N (resident)."""
        self._chk(lib().gkm_adjoint_forward(self._h, dt, nsteps, every))

    def adjoint_backward(self, lamN):
        """The backward part of adjoint_run from the last adjoint_forward's
        trajectory (used up): returns lam_0 (padded layout)."""
        lamN = np.ascontiguousarray(lamN, dtype=np.float64)
        out = np.empty((5, self.nt, self.nt, self.nt))
        self._chk(lib().gkm_adjoint_backward(self._h, lamN, out))
        return out

    def set_tiles(self, tiles):
        t = (C.c_int * 3)(*tiles)
        self._chk(lib().gkm_set_sweep_tiles(self._h, C.byref(t)))

    def tiles(self):
        t = (C.c_int * 3)()
        self._chk(lib().gkm_get_sweep_tiles(self._h, C.byref(t)))
        return tuple(t)

    def autotune(self, reps=2):
        """Time every sweep tile variant on the resident state, keep the fastest;
        returns (tiles, ms) with ms[d][v] the average launch time."""
        t = (C.c_int * 3)()
        ms = (C.c_double * 9)()
        self._chk(lib().gkm_autotune(self._h, reps, C.byref(t), C.byref(ms)))
        return tuple(t), np.array(ms[:]).reshape(3, 3)

    def stats(self):
        s = Stats()
        self._chk(lib().gkm_stats_compute(self._h, C.byref(s)))
        return s

    def timing(self, on=True):
        self._chk(lib().gkm_timing_enable(self._h, 1 if on else 0))

    def timing_read(self):
        ms = (C.c_double * 8)()
        cnt = (C.c_long * 8)()
        self._chk(lib().gkm_timing_read(self._h, C.byref(ms), C.byref(cnt)))
        return np.array(ms[:]), np.array(cnt[:])

    def flux_batch(self, X, tau, dt, precision=GKM_FP64, dx=None):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 50)
        tau = np.ascontiguousarray(tau, dtype=np.float64).reshape(-1)
        F = np.empty((X.shape[0], 5))
        dx = 2.0 * (np.arctan(1.0) * 4.0) / self.nc if dx is None else dx
        self._chk(lib().gkm_flux_batch(self._h, X.shape[0], X, tau, dt, dx, precision, F))
        return F

    def weno_batch(self, S, precision=GKM_FP64):
        S = np.ascontiguousarray(S, dtype=np.float64).reshape(-1, 30)
        out = np.empty((S.shape[0], 10))
        self._chk(lib().gkm_weno_batch(self._h, S.shape[0], S, precision, out))
        return out

    # ---- particles
    def lpt_upload(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.np = pos.shape[0]
        self._chk(lib().lpt_upload(self._h, self.np, pos))

    def lpt_set_reference_drop(self, on=True):
        """Drop tracers at block id != 0 (a coordinate of exactly 2 pi) as the
        reference's single-rank run does; call before lpt_upload."""
        self._chk(lib().lpt_set_reference_drop(self._h, int(on)))

    def lpt_count(self):
        n = C.c_long()
        self._chk(lib().lpt_count(self._h, C.byref(n)))
        return n.value

    def lpt_download(self):
        pos = np.empty((self.np, 3))
        vel = np.empty((self.np, 3))
        self._chk(lib().lpt_download(self._h, _ptr(pos), _ptr(vel)))
        return pos, vel

    def lpt_sample(self):
        U, gU, gV, gW = (np.empty((self.np, 3)) for _ in range(4))
        self._chk(lib().lpt_sample(self._h, _ptr(U), _ptr(gU), _ptr(gV), _ptr(gW)))
        return U, gU, gV, gW

    def lpt_stats(self):
        s = PStats()
        self._chk(lib().lpt_stats(self._h, C.byref(s)))
        return s

    def lpt_properties(self):
        """Particle::getParticleProp of every particle: np x 41 (PROP_NAMES columns)."""
        out = np.empty((self.np, PROP_WIDTH))
        if self.np:
            self._chk(lib().lpt_properties(self._h, out))
        return out

    def lpt_coupled_step(self, dt):
        self._chk(lib().lpt_coupled_step(self._h, dt))

    def lpt_heun(self, dt):
        self._chk(lib().lpt_heun(self._h, dt))


class Ensemble:
    """Independent runs stepped together (wraps gkm_ensemble).  members[m] is a
    Context view of member m (upload / download / stats / cfl_dt, lpt_upload /
    lpt_download / lpt_stats); tracers uploaded to a member advance with it,
    as lpt_coupled_step advances them in a single context."""

    def __init__(self, ncs, params_list, cfl=0.5, precision=GKM_MIXED, prandtl=0.70, device=0, check_state=1):
        n = len(ncs)
        cfgs = (Config * n)(*[make_config(nc, P, cfl, precision, prandtl, device, check_state)
                              for nc, P in zip(ncs, params_list)])
        h = C.c_void_p()
        rc = lib().gkm_ensemble_create(n, cfgs, C.byref(h))
        self._h = h
        if rc:
            msg = lib().gkm_ensemble_last_error(h).decode() if h.value else "gkm_ensemble_create failed"
            if h.value:
                lib().gkm_ensemble_destroy(h)
            self._h = C.c_void_p()
            raise GkmError(rc, msg)
        self.members = [Context._view(lib().gkm_ensemble_member(h, m), nc) for m, nc in enumerate(ncs)]

    def _chk(self, rc):
        if rc:
            raise GkmError(rc, lib().gkm_ensemble_last_error(self._h).decode())

    def close(self):
        if self._h and self._h.value:
            for m in self.members:
                m._h = C.c_void_p()
            lib().gkm_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tiles(self, tiles):
        t = (C.c_int * 3)(*tiles)
        self._chk(lib().gkm_ensemble_set_tiles(self._h, C.byref(t)))

    def autotune(self, reps=2):
        t = (C.c_int * 3)()
        ms = (C.c_double * 9)()
        self._chk(lib().gkm_ensemble_autotune(self._h, reps, C.byref(t), C.byref(ms)))
        return tuple(t), np.array(ms[:]).reshape(3, 3)

    def run(self, nsteps, graph=True):
        self._chk(lib().gkm_ensemble_run(self._h, nsteps, 1 if graph else 0))

    def set_persistent(self, on=True):
        """One persistent-kernel launch per step instead of one launch per phase."""
        self._chk(lib().gkm_ensemble_set_persistent(self._h, 1 if on else 0))

    def persistent_profile(self):
        """Run one persistent step; ms of each phase (5 or 9 entries)."""
        ms = np.zeros(9)
        n = C.c_int(0)
        self._chk(lib().gkm_ensemble_persistent_profile(self._h, ms, C.byref(n)))
        return ms[:n.value]

    def dt(self):
        out = np.empty(len(self.members))
        self._chk(lib().gkm_ensemble_dt(self._h, out))
        return out

