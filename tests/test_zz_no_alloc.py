"""gkm_step_noalloc: one step on caller-owned device buffers with no device
allocation, against gkm_step of a context (the same kernels, tiles and
segment lengths, so the results are equal bit for bit).

Two calls in a row on one GPU, the second on a smaller grid (48^3, then
32^3), in FP64 and in mixed precision (the context with compensation off:
the no-allocation step does the plain update); a grid above the maximum
(nc = 80 > 64) is refused with GKM_EINVAL and a message naming the limit.
Device buffers come from the HIP runtime the library itself uses (hipMalloc
and hipMemcpy through ctypes, plumbing only; torch bundles a second HIP
runtime, which cannot be initialised once this one has been), and the step
runs on the null stream."""
import ctypes as C

import numpy as np
import pytest

from cases import make_case
from dns_gkm_lpt_amd import _lib as gl

pytestmark = pytest.mark.gpu

INNER = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))


def _hip():
    gl.lib()  # libdnsgkm.so loads /opt/rocm's libamdhip64 first; this is that library
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc_step(W_inner, nc, P, precision, dt):
    hip = _hip()
    host = np.ascontiguousarray(W_inner, dtype=np.float64)
    nbytes = host.nbytes
    d_in, d_out = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_in), nbytes) == 0 and hip.hipMalloc(C.byref(d_out), nbytes) == 0
    try:
        assert hip.hipMemcpy(d_in, host.ctypes.data_as(C.c_void_p), nbytes, 1) == 0  # host to device
        cfg = gl.make_config(nc, P, cfl=0.5, precision=precision)
        msg = C.create_string_buffer(256)
        rc = gl.lib().gkm_step_noalloc(C.byref(cfg), d_in, d_out, dt, None, msg, len(msg))
        assert rc == 0, msg.value.decode()
        assert hip.hipDeviceSynchronize() == 0
        out = np.empty_like(host)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d_out, nbytes, 2) == 0  # device to host
        return out
    finally:
        hip.hipFree(d_in)
        hip.hipFree(d_out)


def _context_step(W, nc, P, precision, dt):
    with gl.Context(nc, P, precision=precision) as c:
        if precision == gl.GKM_MIXED:
            c.set_compensation(False)
        c.upload(W)
        c.step(dt)
        return c.download()[INNER]


@pytest.mark.parametrize("precision", [gl.GKM_FP64, gl.GKM_MIXED])
def test_noalloc_step_equals_context_step(precision):
    for nc, seed in ((48, 1), (32, 2)):  # the second call on a smaller grid
        W, P, _ = make_case(nc, seed=seed)
        with gl.Context(nc, P) as c:
            c.upload(W)
            dt = c.cfl_dt()
        ref = _context_step(W, nc, P, precision, dt)
        out = _noalloc_step(W[INNER], nc, P, precision, dt)
        print(f"nc {nc} precision {precision}: max |diff| {np.max(np.abs(out - ref)):.3e}", flush=True)
        assert np.array_equal(out, ref)


def test_noalloc_refuses_grids_above_its_maximum():
    _, P, _ = make_case(32)
    cfg = gl.make_config(80, P)
    msg = C.create_string_buffer(256)
    rc = gl.lib().gkm_step_noalloc(C.byref(cfg), C.c_void_p(1), C.c_void_p(2), 1e-4, None, msg, len(msg))
    assert rc == 1
    assert b"exceeds the no-allocation maximum of 64" in msg.value
