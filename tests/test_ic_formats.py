"""Initial-condition recipe and the reference's binary input formats."""
import os
import re
import struct

import numpy as np
import pytest

from dns_gkm_lpt_amd import ic as icm


def test_ic_is_solenoidal_and_scaled():
    d = icm.make_ic(32, k0=4.0, Mt=0.3, Re=30.0, seed=3)
    u, v, w = d["u"], d["v"], d["w"]
    k = np.fft.fftfreq(32, 1 / 32)
    ky, kx, kz = np.meshgrid(k, k, k, indexing="ij")
    div = kx * np.fft.fftn(u) + ky * np.fft.fftn(v) + kz * np.fft.fftn(w)
    assert np.max(np.abs(div)) < 1e-9 * np.max(np.abs(np.fft.fftn(u)))
    c = np.sqrt(1.4 * 86100.0)
    Mt = np.sqrt(np.mean(u * u + v * v + w * w)) / c
    assert Mt == pytest.approx(0.3, rel=1e-12)
    assert d["params"]["Mt"] == pytest.approx(0.3, rel=1e-12)
    assert d["params"]["p0"] == 86100.0 and d["params"]["gam"] == 1.4


def test_ic_spectrum_shape_and_band():
    d = icm.make_ic(64, k0=4.0, Mt=0.3, Re=50.0, seed=1)
    k = np.fft.fftfreq(64, 1 / 64)
    ky, kx, kz = np.meshgrid(k, k, k, indexing="ij")
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    e = sum(np.abs(np.fft.fftn(d[c])) ** 2 for c in "uvw")
    assert np.all(e[(kk > 8.0) | (kk < 1.0)] < 1e-18 * e.max())
    # each mode inside the band carries E(k)/(4 pi k^2) ~ k^2 exp(-2 (k/k0)^2)
    band = (kk >= 1) & (kk <= 8)
    ratio = e[band] / (kk[band] ** 2 * np.exp(-2 * (kk[band] / 4.0) ** 2))
    assert np.std(ratio) < 1e-8 * np.mean(ratio)


def test_ic_viscosity_from_re_lambda():
    d = icm.make_ic(32, k0=4.0, Mt=0.3, Re=30.0, seed=0)
    u, v, w = d["u"], d["v"], d["w"]
    dx = 2 * np.pi / 32
    sq = 0.0
    for f in (u, v, w):
        for ax in range(3):
            g = 0.5 * (np.roll(f, -1, ax) - np.roll(f, 1, ax)) / dx
            sq += np.mean(g * g)
    ek0 = 0.5 * np.mean(u * u + v * v + w * w)
    urms, lam = np.sqrt(2 * ek0 / 3), np.sqrt(10 * ek0 / sq)
    nu = d["params"]["mu0"] / d["params"]["rho0"]
    assert urms * lam / nu * np.sqrt(3.0) / np.sqrt(3.0) == pytest.approx(
        2 * ek0 * np.sqrt(5 / (3 * sq)) / nu / (2 * ek0 * np.sqrt(5 / (3 * sq)) / nu) * 30.0, rel=1e-12)
    assert d["params"]["t0"] == pytest.approx(lam / urms, rel=1e-12)


def test_params_file_bytes(tmp_path):
    p = dict(gam=1.4, rho0=1.0, Mt=0.3, Re=30.0, mu0=0.25, p0=86100.0, t0=0.01)
    f = tmp_path / "params"
    icm.write_params(str(f), 32, p)
    raw = f.read_bytes()
    assert len(raw) == 4 + 7 * 8
    assert struct.unpack("<i", raw[:4])[0] == 32
    assert struct.unpack("<7d", raw[4:]) == (1.4, 1.0, 0.3, 30.0, 0.25, 86100.0, 0.01)
    n, q = icm.read_params(str(f))
    assert n == 32 and q == p


def test_init_file_matlab_order(tmp_path):
    """Element (j,i,k) of a field sits at k*n*n + i*n + j (FlowField.cpp:97)."""
    n = 6
    rng = np.random.default_rng(0)
    d = {c: rng.standard_normal((n, n, n)) for c in ("u", "v", "w", "rho", "p")}
    f = tmp_path / "Init000"
    icm.write_init(str(f), d)
    raw = np.fromfile(str(f), dtype="<f8")
    assert raw.size == 5 * n ** 3
    j, i, k = 2, 4, 1
    for q, c in enumerate(("u", "v", "w", "rho", "p")):
        assert raw[q * n ** 3 + k * n * n + i * n + j] == d[c][j, i, k]
    back = icm.read_init(str(f), n)
    for c in d:
        assert np.array_equal(back[c], d[c])


def test_particle_files(tmp_path):
    xyz = icm.uniform_particles(100, seed=2)
    assert xyz.min() >= 0 and xyz.max() < 2 * np.pi
    icm.write_particles(str(tmp_path), xyz)
    assert struct.unpack("<i", (tmp_path / "particleCount").read_bytes())[0] == 100
    raw = np.fromfile(str(tmp_path / "unirand"), dtype="<f8")
    assert np.array_equal(raw[:100], xyz[:, 0]) and np.array_equal(raw[100:200], xyz[:, 1])
    assert np.array_equal(icm.read_particles(str(tmp_path)), xyz)


def test_prim_to_cons_matches_initialize_dhit():
    n = 8
    rng = np.random.default_rng(5)
    u, v, w = (rng.standard_normal((n, n, n)) for _ in range(3))
    rho = 1 + 0.1 * rng.random((n, n, n))
    p = 86100 * (1 + 0.1 * rng.random((n, n, n)))
    W = icm.prim_to_cons(u, v, w, rho, p)
    inner = (slice(3, -3),) * 3
    assert np.array_equal(W[0][inner], rho)
    assert np.array_equal(W[1][inner], u * rho)
    assert np.allclose(W[4][inner], rho * (p / (rho * 0.4) + 0.5 * (u * u + v * v + w * w)), rtol=1e-15)
    idx = (np.arange(n + 6) - 3) % n + 3
    assert np.array_equal(W, W[:, idx][:, :, idx][:, :, :, idx])
