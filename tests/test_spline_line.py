"""The blocked natural-spline line solve of the x/y prefilter passes
(csrc/lpt_device.hpp line_solve_blocked) on the host: every position read and
written once, three 16-row blocks in registers, the backward sweep of each
block entering from a warm-up over the next two (32 rows).

Checked against the definition of the natural cubic spline (knot values
(c_i + 4 c_{i+1} + c_{i+2})/6 = f_i and zero second differences at both ends;
residual <= 1e-14 of max |f|, measured 3.5e-16) and against the plain
two-sweep Thomas solve (<= 1e-15, measured 1.1e-16) on 60 lines: 4 .. 1030
knots (shorter than one block, exact multiples of it, ragged ends), strides 1
and 7, random and smooth data.  The same check rejects blocks of 8 rows, a
warm-up of 16 (error 0.268^16 of the line's values; measured 4.6e-10)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "dns_gkm_lpt_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SRC = os.path.join(HERE, "host", "spline_line_check.cpp")


def _build(exe, *defs):
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-I", CSRC, *defs, SRC, "-o", str(exe)],
                          stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_blocked_line_solve_is_the_natural_spline(tmp_path):
    exe = tmp_path / "spline_line_check"
    _build(exe)
    out = subprocess.run([str(exe), "1e-14", "1e-15"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_short_warm_up_is_rejected(tmp_path):
    exe = tmp_path / "spline_line_check_b8"
    _build(exe, "-DLINE_BLOCK=8")
    out = subprocess.run([str(exe), "1e-14", "1e-15"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 1, out.stdout + out.stderr
