"""The drop-in host layer: the reference's own Main.cpp compiles and links
unchanged against dns_gkm_lpt_amd/host (FlowField / GKMflux / update /
ParticleTrack / param), and the driver binary is built."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "dns_gkm_lpt_amd", "host")
REF_MAIN = "/root/reference/src/Main.cpp"


def test_driver_built():
    subprocess.check_call(["make", "-s", "-C", HOST], stdout=subprocess.DEVNULL)
    assert os.access(os.path.join(HOST, "dnsgkm_run"), os.X_OK)


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="reference sources not present")
def test_reference_main_links_against_dropin():
    subprocess.check_call(["make", "-s", "-C", HOST, "refmain"], stdout=subprocess.DEVNULL)
    exe = os.path.join(HOST, "_refmain", "Main.out")
    assert os.access(exe, os.X_OK)
    syms = subprocess.run(["nm", "-C", exe], capture_output=True, text=True).stdout
    for name in ("GetParams()", "PrintParams()", "segment()", "Initialize_DHIT(double**)",
                 "Evolve(double**, int)"):
        assert name in syms, name


def test_dropin_flux_matches_oracle(tmp_path):
    """flux() of the drop-in layer (the FP64 rest-frame flux template evaluated
    on the host) against the oracle's restatement of GKMflux.cpp:11-229 on
    seeded faces (weak and strong jumps): within 1e-12 of each component's
    largest magnitude."""
    import ctypes as C

    import numpy as np

    import pyoracle as po
    from cases import random_faces

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    so = tmp_path / "libhostflux.so"
    subprocess.check_call([hipcc, "-xhip", "--cuda-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
                           "-I", HOST, os.path.join(ROOT, "tests", "host", "flux_shim.cpp"),
                           os.path.join(HOST, "GKMflux.cpp"), os.path.join(HOST, "param.cpp"), "-o", str(so)],
                          stdout=subprocess.DEVNULL)
    L = C.CDLL(str(so))
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.host_flux_batch.argtypes = [C.c_long, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp]
    nc = 32
    gas = po.make_gas(nc, mu0=1.2e-3)
    for jump in (0.05, 0.3):
        X = random_faces(2000, seed=31, jump=jump)
        tau = np.full(X.shape[0], 3e-6)
        dt = 1.5e-4
        ref = po.flux(X, tau, dt, gas)
        out = np.empty_like(ref)
        L.host_flux_batch(X.shape[0], X, tau, dt, 1.4, 1.2e-3, gas.T0, nc, out)
        err = np.max(np.abs(out - ref), axis=0) / np.max(np.abs(ref), axis=0)
        assert np.all(err < 1e-12), (jump, err)
