"""Layout of the compensated update's FP32 carry (csrc/gkm_kernels.hpp:
carry_plane, carry_index), which only the z sweep's update reads and writes:
16-cell blocks along z, consecutive y lines adjacent.  On the host: every cell
of nc^3 has its own index below carry_plane(nc) and a z block of a line is 16
consecutive floats, for the grid sizes the GPU tests run (16, 24, 32, 40, 48,
64, 100, 128) and 512^3.  Leaving the block number out of the index must fail
the same check.  The GPU side (test_gpu_compensation.py) reads the carry back
through gkm_download_carry, which undoes the blocking."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "dns_gkm_lpt_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SRC = os.path.join(HERE, "host", "carry_layout_check.cpp")
SIZES = ["16", "24", "32", "40", "48", "64", "100", "128", "512"]

needs_hipcc = pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None,
                                 reason="hipcc not available")


def _build(exe, *defs):
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-I", CSRC, *defs, SRC, "-o", str(exe)],
                          stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


@needs_hipcc
def test_carry_layout_is_one_to_one(tmp_path):
    exe = tmp_path / "carry_layout_check"
    _build(exe)
    out = subprocess.run([str(exe), *SIZES], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr


@needs_hipcc
def test_dropped_block_number_is_rejected(tmp_path):
    exe = tmp_path / "carry_layout_check_plant"
    _build(exe, "-DPLANT_DROP_BLOCK")
    out = subprocess.run([str(exe), "24", "40"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 1, out.stdout + out.stderr
