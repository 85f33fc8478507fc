"""The C-ABI library loads and exports every function include/dnsgkm.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

from dns_gkm_lpt_amd import _lib as gl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "dnsgkm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:gkm|lpt)_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared() == sorted(gl.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(gl.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", gl.LIB_PATH], capture_output=True, text=True).stdout
    syms = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    for name in declared():
        assert name in syms, name


def test_binding_loads_without_gpu():
    L = gl.lib()
    assert L.gkm_last_error(None) == b"null context"
