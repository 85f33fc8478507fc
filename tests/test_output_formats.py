"""Output files of the reference (writeField, writeParticleData + T_,
WriteReInit; update.cpp:1324-1422): the drop-in C++ writers and the Python
writers produce byte-identical files from the same data, the line layout is
the reference's, and the restart file reads back as the state's primitives."""
import os
import subprocess

import numpy as np
import pytest

from dns_gkm_lpt_amd import _lib as gl
from dns_gkm_lpt_amd import ic as icm
from dns_gkm_lpt_amd import output as out
from cases import make_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "dns_gkm_lpt_amd", "host")
MPI = "/opt/conda"


def build_driver(tmp):
    subprocess.check_call(["make", "-s", "-C", HOST, "libdnsgkm_host.a"], stdout=subprocess.DEVNULL)
    exe = os.path.join(tmp, "writers_main")
    libdir = os.path.join(ROOT, "dns_gkm_lpt_amd")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", HOST, "-I", os.path.join(MPI, "include"),
                           "-o", exe, os.path.join(ROOT, "tests", "host", "writers_main.cpp"),
                           os.path.join(HOST, "libdnsgkm_host.a"), "-L", libdir, "-ldnsgkm",
                           os.path.join(MPI, "lib", "libmpi.so"),
                           f"-Wl,-rpath,{libdir}:/usr/lib/x86_64-linux-gnu:{MPI}/lib"])
    return exe


def case(nc=4, npart=5):
    W, P, gas = make_case(nc, seed=11)
    rng = np.random.default_rng(12)
    pos = rng.uniform(0, 2 * np.pi, (npart, 3))
    props = rng.standard_normal((npart, gl.PROP_WIDTH)) * 10.0 ** rng.integers(-3, 4, (npart, gl.PROP_WIDTH))
    return W, P, pos, props


@pytest.mark.skipif(not os.path.exists(os.path.join(MPI, "lib", "libmpi.so")), reason="MPI library not present")
def test_cpp_and_python_writers_agree_bytewise(tmp_path):
    exe = build_driver(str(tmp_path))
    W, P, pos, props = case()
    nc = W.shape[-1] - 6
    wfile, pfile = tmp_path / "W.bin", tmp_path / "parts.bin"
    np.ascontiguousarray(W).tofile(wfile)
    np.concatenate([pos, props], 1).tofile(pfile)
    cdir, pdir = tmp_path / "cpp", tmp_path / "py"
    os.makedirs(cdir)
    subprocess.check_call([exe, str(nc), str(wfile), str(pfile), str(len(pos))], cwd=cdir)
    res = pdir / "Results"
    out.write_field(str(res), W, 7, 1.4)
    out.write_particles(str(res), 7, pos, props)
    out.write_time(str(res), 7, 0.25, 1e-3, 2.0)
    out.write_reinit(str(pdir / "InitialCondition" / "Init000"), W, 1.4)
    for rel in ("Results/Field/Field_0_7", "Results/Particles/P_0_7", "Results/Time/T_7",
                "InitialCondition/Init000"):
        a, b = (cdir / rel).read_bytes(), (pdir / rel).read_bytes()
        assert a == b, rel


def test_particle_line_layout(tmp_path):
    W, P, pos, props = case(npart=2)
    path = out.write_particles(str(tmp_path), 3, pos, props, tags=np.array([4, 9]))
    lines = open(path).read().splitlines()
    assert len(lines) == 2
    f = lines[1].split(" ")
    assert f[0] == "9" and len(f) == 1 + 3 + 1 + 3 + 1 + 8 * 3 + 9
    v = [float(x) for x in f[1:]]
    # position and vectors print x, y, z = members [1], [0], [2]
    assert v[0:3] == [float("%.15e" % pos[1, i]) for i in (1, 0, 2)]
    assert v[3] == float("%.15e" % props[1, 0])                      # DEN
    assert v[4:7] == [float("%.15e" % props[1, 1 + i]) for i in range(3)]  # U in member order
    assert v[8:11] == [float("%.15e" % props[1, 5 + i]) for i in (1, 0, 2)]  # gradD
    hess = [float("%.15e" % props[1, 23 + i]) for i in (4, 3, 5, 1, 0, 2, 7, 6, 8)]
    assert v[-9:] == hess


def test_reinit_reads_back_as_primitives(tmp_path):
    W, P, pos, props = case(nc=6)
    nc = 6
    path = out.write_reinit(str(tmp_path / "Init000"), W, 1.4)
    back = icm.read_init(path, nc)
    den, u, v, w, p = out.primitives(W, 1.4)
    inner = (slice(3, nc + 3),) * 3
    for key, ref in (("u", u), ("v", v), ("w", w), ("rho", den), ("p", p)):
        assert np.array_equal(back[key], ref[inner]), key
    # and the restart round-trips through the initial-condition path to the same state
    W2 = icm.prim_to_cons(back["u"], back["v"], back["w"], back["rho"], back["p"], gam=1.4)
    assert np.max(np.abs(W2 - W) / np.abs(W).max(axis=(1, 2, 3))[:, None, None, None]) < 1e-14


def test_field_file_layout(tmp_path):
    W, P, pos, props = case(nc=4)
    path = out.write_field(str(tmp_path), W, 0, 1.4)
    lines = open(path).read().splitlines()
    assert len(lines) == 64
    # loops y, x, z; columns x y z
    assert lines[0].split()[:3] == ["0", "0", "0"]
    assert lines[1].split()[:3] == ["0", "0", "1"]
    assert lines[4].split()[:3] == ["1", "0", "0"]
    assert lines[16].split()[:3] == ["0", "1", "0"]
