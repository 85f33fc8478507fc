"""Reverse-mode WENO gradient of the adjoint (csrc/adj_math.hpp) against
forward-mode dual numbers (csrc/gkm_dual.hpp) on the host: 80000 seeded
stencils (smooth, kinked, discontinuous, nearly constant).  Tolerance 1e-10 of
the largest partial derivative of each stencil: the two modes associate the
same products differently, and the reverse mode forms q_k - out (candidate
minus result), which cancels on nearly constant data (measured 5.6e-12 there,
<= 2.9e-13 on the other kinds)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "dns_gkm_lpt_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_weno_side_grad_matches_dual_numbers(tmp_path):
    exe = tmp_path / "adj_math_check"
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "host", "adj_math_check.cpp"),
                           "-o", str(exe)], stdout=subprocess.DEVNULL)
    out = subprocess.run([str(exe), "1e-10"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
