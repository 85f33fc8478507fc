"""Closed-form pins of the oracle's statistics that the Taylor-Green test in
test_oracle_pins.py leaves open: the dissipation with the Sutherland viscosity
(update.cpp:751-752), the P_rms / T_rms / V_rms normalisations
(update.cpp:808-833) and the particle S_ij S_ij (update.cpp:1131).

Each pin is a function of its own so that test_pins_fail_on_planted_mistakes
can run it against oracle variants built from a copy of gkm_oracle.c with one
plausible slip planted (the T^0.76 power law in place of Sutherland, mu at
the wrong temperature, a dropped term or normaliser, swapped gradient axes):
every variant must fail the pin that covers it.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle as po
from dns_gkm_lpt_amd import ic as icm

R = 287.0


def _cell_centres(nc):
    dx = 2 * np.pi / nc
    c = (np.arange(nc) + 0.5) * dx
    return np.meshgrid(c, c, c, indexing="ij")  # y, x, z


def _taylor_green(nc):
    y, x, z = _cell_centres(nc)
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    return y, x, z, u, v, np.zeros_like(u)


def _sutherland(mu0, T, T0):
    """mu0 (T/T0)^{3/2} (T0 + S)/(T + S), S = 110.4 K."""
    return mu0 * (T / T0) ** 1.5 * (T0 + 110.4) / (T + 110.4)


# ------------------------------------------------------------------ pins
def pin_diss_uniform_temperature():
    """Taylor-Green at uniform T = 450 K != T0 = 300 K, mu0 != 0:
    <du_i/dx_j du_i/dx_j> = 6 x 1/8, so Diss = mu(450 K) 3/4 (spline
    gradients at 64^3: 2e-3; the T^0.76 law differs by 3.6 %, mu(T0) by 35 %)."""
    nc = 64
    _, _, _, u, v, w = _taylor_green(nc)
    rho = 1.2
    T = 450.0
    p = rho * R * T
    W = icm.prim_to_cons(u, v, w, np.full_like(u, rho), np.full_like(u, p))
    mu0 = 1.5e-3
    gas = po.make_gas(nc, mu0=mu0)  # T0 = 86100 / 287 = 300 K
    assert gas.T0 == pytest.approx(300.0, rel=1e-15)
    s = po.field_stats(W, po.velocity_coefs(W), gas, 86100.0, 0.3, 1.0)
    expect = _sutherland(mu0, T, 300.0) * 0.75
    assert s.diss == pytest.approx(expect, rel=2e-3), (s.diss, expect)


def pin_diss_temperature_varying_in_z():
    """Taylor-Green with T = T1 (1 + 0.3 cos 2z) (pressure varies, density
    uniform), so mu varies from cell to cell.  Averaged over x and y the
    squared gradients of the Taylor-Green field are cos^2 z + sin^2 z / 2, so
    Diss = (1/nc) sum_z mu(T(z)) (cos^2 z + sin^2 z / 2) (spline gradients)."""
    nc = 64
    _, _, z, u, v, w = _taylor_green(nc)
    rho = 0.9
    T = 380.0 * (1 + 0.3 * np.cos(2 * z))
    W = icm.prim_to_cons(u, v, w, np.full_like(u, rho), rho * R * T)
    mu0 = 2.0e-3
    gas = po.make_gas(nc, mu0=mu0)
    s = po.field_stats(W, po.velocity_coefs(W), gas, 86100.0, 0.3, 1.0)
    zc = (np.arange(nc) + 0.5) * 2 * np.pi / nc
    Tz = 380.0 * (1 + 0.3 * np.cos(2 * zc))
    expect = np.mean(_sutherland(mu0, Tz, 300.0) * (np.cos(zc) ** 2 + 0.5 * np.sin(zc) ** 2))
    assert s.diss == pytest.approx(expect, rel=2e-3), (s.diss, expect)


def pin_rms_normalisers():
    """p = p0 (1 + e sin x), rho = r0 (1 + d sin y), at rest.  Cell-centre
    means of periodic analytic functions are exact to rounding here, so
    mean(sin) = 0, mean(sin^2) = 1/2, mean(1/(1 + d sin)) = (1 - d^2)^(-1/2),
    mean(1/(1 + d sin)^2) = (1 - d^2)^(-3/2), and
      P_rms = p0 e / sqrt(2) / (gam p0 Mt^2),
      T_rms = T0' sqrt((1 + e^2/2) a2 - a1^2) / ((gam - 1) T0 Mt^2),
              T0' = p0 / (R r0), a1 = (1-d^2)^(-1/2), a2 = (1-d^2)^(-3/2),
      V_rms = sqrt(a2 - a1^2) / r0 / (Mt^2 / den0)."""
    nc = 32
    y, x, _ = _cell_centres(nc)
    eps, dlt = 0.07, 0.12
    p0, r0 = 86100.0, 1.3
    den0, Mt, gam = 1.3, 0.45, 1.4
    p = p0 * (1 + eps * np.sin(x))
    rho = r0 * (1 + dlt * np.sin(y))
    z = np.zeros_like(p)
    W = icm.prim_to_cons(z, z, z, rho, p, gam=gam)
    gas = po.make_gas(nc, gam=gam, mu0=1e-3, p0=p0, den0=den0)
    T0 = p0 / (R * den0)
    assert gas.T0 == pytest.approx(T0, rel=1e-15)
    s = po.field_stats(W, po.velocity_coefs(W), gas, p0, Mt, den0)
    a1 = (1 - dlt ** 2) ** -0.5
    a2 = (1 - dlt ** 2) ** -1.5
    Tp = p0 / (R * r0)
    assert s.pmean == pytest.approx(p0, rel=1e-13)
    assert s.tmean == pytest.approx(Tp * a1, rel=1e-13)
    assert s.vmean == pytest.approx(a1 / r0, rel=1e-13)
    assert s.prms == pytest.approx(p0 * eps / np.sqrt(2) / (gam * p0 * Mt * Mt), rel=1e-11)
    assert s.trms == pytest.approx(Tp * np.sqrt((1 + eps ** 2 / 2) * a2 - a1 ** 2) / ((gam - 1) * T0 * Mt * Mt),
                                   rel=1e-10)
    assert s.vrms == pytest.approx(np.sqrt(a2 - a1 ** 2) / r0 / (Mt * Mt / den0), rel=1e-10)


def pin_particle_sijsij_linear_field():
    """u = A x (a linear field, which the natural spline reproduces exactly,
    gradients included): every particle sees grad u = A, so
    part_S_ijS_ij = |sym(A)|_F^2 np / np_total and part_TKE = <|A x|^2>."""
    nc = 16
    nt = nc + 6
    dx = 2 * np.pi / nc
    g = -3 * dx + 0.5 * dx + dx * np.arange(nt)
    Y, X, Z = np.meshgrid(g, g, g, indexing="ij")
    # rows u, v, w; columns d/dx, d/dy, d/dz -- no symmetry, distinct entries
    A = np.array([[0.31, -0.72, 0.15], [0.44, -0.09, 0.63], [-0.27, 0.58, -0.22]])
    cs = [po.spline_coefs(np.ascontiguousarray(A[r, 0] * X + A[r, 1] * Y + A[r, 2] * Z)) for r in range(3)]
    rng = np.random.default_rng(11)
    xyz = rng.uniform(1.0, 5.0, (400, 3))
    pos = np.ascontiguousarray(xyz[:, [1, 0, 2]])
    U, gU, gV, gW = po.particle_sample(cs, dx, pos)
    assert np.allclose(U, xyz @ A.T, atol=1e-12)
    n_total = 500
    s = po.particle_stats(U, gU, gV, gW, n_total=n_total)
    S = 0.5 * (A + A.T)
    assert s.sijsij == pytest.approx(np.sum(S * S) * 400 / n_total, rel=1e-12), (s.sijsij, np.sum(S * S))
    assert s.tke == pytest.approx(np.sum((xyz @ A.T) ** 2) / n_total, rel=1e-12)
    # du/dx and dw/dz are the same at every particle: m3 / m2^1.5 = sign * sqrt(n_total / np)
    f = np.sqrt(n_total / 400)
    assert s.skew_u == pytest.approx(np.sign(A[0, 0]) * f, rel=1e-12)
    assert s.skew_w == pytest.approx(np.sign(A[2, 2]) * f, rel=1e-12)


PINS = {
    "diss_uniform_T": pin_diss_uniform_temperature,
    "diss_T_of_z": pin_diss_temperature_varying_in_z,
    "rms": pin_rms_normalisers,
    "particle_sij": pin_particle_sijsij_linear_field,
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    PINS[name]()


# ------------------------------------------------------- planted mistakes
# (name, original text in gkm_oracle.c, planted text, pins that must fail)
MUTANTS = [
    ("mu_power_076", "g->mu0 * pow((T / g->T0), 1.5) * (g->T0 + 110.4) / (T + 110.4)",
     "g->mu0 * pow((T / g->T0), 0.76)", ["diss_uniform_T", "diss_T_of_z"]),
    ("mu_at_T0", "(g->T0 + 110.4) / (T + 110.4)", "(g->T0 + 110.4) / (g->T0 + 110.4)",
     ["diss_uniform_T", "diss_T_of_z"]),
    ("diss_drops_dvdz", "dvdy * dvdy + dvdz * dvdz +\n", "dvdy * dvdy +\n", ["diss_uniform_T", "diss_T_of_z"]),
    ("prms_no_gamma", "/ (g->gam * p0 * Mt * Mt)", "/ (p0 * Mt * Mt)", ["rms"]),
    ("trms_no_gm1", "((g->gam - 1) * g->T0 * Mt * Mt)", "(g->T0 * Mt * Mt)", ["rms"]),
    ("vrms_no_den0", "(Mt * Mt / den0)", "(Mt * Mt)", ["rms"]),
    ("pstats_swapped_axes", "double dudx = gU[3 * p + 1], dvdx = gV[3 * p + 1]",
     "double dudx = gU[3 * p + 0], dvdx = gV[3 * p + 1]", ["particle_sij"]),
    ("pstats_half_dropped", "0.5 * sq(dwdy + dvdz) + dvdy * dvdy +", "sq(dwdy + dvdz) + dvdy * dvdy +",
     ["particle_sij"]),
]


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    return tmp_path_factory.mktemp("mutants")


@pytest.mark.parametrize("name,orig,planted,must_fail", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_fail_on_planted_mistakes(mutant_dir, name, orig, planted, must_fail):
    src = open(os.path.join(os.path.dirname(po.__file__), "gkm_oracle.c")).read()
    assert src.count(orig) == 1, f"the text to mutate is not in gkm_oracle.c once: {orig!r}"
    d = mutant_dir / name
    d.mkdir()
    shutil.copy(os.path.join(os.path.dirname(po.__file__), "gkm_oracle.h"), d / "gkm_oracle.h")
    (d / "gkm_oracle.c").write_text(src.replace(orig, planted))
    so = d / f"libmut_{name}.so"
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-shared", "-o", str(so), str(d / "gkm_oracle.c"), "-lm"])
    with po.using(str(so)):
        for pin in must_fail:
            with pytest.raises(AssertionError):
                PINS[pin]()
