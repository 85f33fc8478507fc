"""Reverse-mode pieces of the adjoint (csrc/adj_math.hpp, csrc/adj_flux.hpp)
against forward-mode dual numbers (csrc/gkm_dual.hpp) on the host.

WENO gradient: 80000 seeded stencils (smooth, kinked, discontinuous, nearly
constant).  Tolerance 1e-10 of the largest partial derivative of each stencil:
the two modes associate the same products differently, and the reverse mode
forms q_k - out (candidate minus result), which cancels on nearly constant data
(measured 5.6e-12 there, <= 2.9e-13 on the other kinds).

Per-face flux adjoint (face_flux_vjp, one reverse pass through the whole
flux): 4000 seeded faces (air at 300 K, up to Mach 0.6, weak and strong
jumps), all 50 partials of the face flux (cell states, WENO states, six slopes;
collision time from the left cell) contracted with a random adjoint flux,
against one dual pass per input.  Tolerance 1e-12 of the largest partial times
its input (measured 2.8e-15).  Planted mistakes read far above it: the
half-space moment's d/d lambda scaled by 0.8 gives 6.6e-2, the energy slot of
c2p's adjoint dropped 1.0, the p3 coefficient's adjoint dropped 0.94, the
pressure-jump term's d/d lambda dropped 0.61, one term of d f3 / d theta
dropped 0.27."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "dns_gkm_lpt_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_reverse_mode_matches_dual_numbers(tmp_path):
    exe = tmp_path / "adj_math_check"
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "host", "adj_math_check.cpp"),
                           "-o", str(exe)], stdout=subprocess.DEVNULL)
    out = subprocess.run([str(exe), "1e-10", "1e-12"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_half_erfc_matches_the_c_library(tmp_path):
    """0.5 erfc(x) as the flux forms it (gkm_math.hpp half_erfc: exp(-x^2)
    times a fitted erfcx) against the C library's erfc on 4.4e6 points of
    [-12, 12] and of the flux's own range: absolute difference <= 1e-15
    (measured 5.6e-16, near x = 0; the relative difference reaches 1.5e-14
    only where erfc is below 1e-60, from the rounding of x^2 inside the
    exponential)."""
    exe = tmp_path / "erfc_check"
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "host", "erfc_check.cpp"),
                           "-o", str(exe)], stdout=subprocess.DEVNULL)
    out = subprocess.run([str(exe), "1e-15"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr

