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
