"""The reference's output files, written from host arrays.

* ``Results/Field/Field_<rank>_<iter>`` -- writeField (update.cpp:1324-1339):
  one text line per interior cell, loops y, x, z, columns
  ``x y z rho u v w p`` (indices %d, values %.15e).
* ``Results/Particles/P_<rank>_<iter>`` -- writeParticleData
  (update.cpp:1341-1369): one text line per particle, tag then 42 values.
* ``Results/Time/T_<iter>`` -- t, dt, t0 (update.cpp:1370-1378).
* ``InitialCondition/Init<a><b><c>`` -- WriteReInit (update.cpp:1381-1422):
  the Init input format (u, v, w, rho, p in MATLAB order), a restart point.

Primitive variables follow generateSplines / c2p (update.cpp:1202-1210,
844-853): u = m/rho, p = (rho (gam-1)) (E/rho - |u|^2/2).
"""
import os

import numpy as np

from . import ic as icm

# P_ line after the tag: a column of the lpt_properties row, or ("pos", a), in
# the order writeParticleData prints them; vectors go out as (x, y, z) =
# members [1], [0], [2], the Hessian as rows x, y, z of the same swap.

def _p_line_columns():
    cols = [("pos", 1), ("pos", 0), ("pos", 2), 0, 1, 2, 3, 4]
    for base in (5, 8, 11, 14, 17, 32, 35, 38):  # gradD, gradU, gradV, gradW, gradP, gradLapU/V/W
        cols += [base + 1, base + 0, base + 2]
    for idx in (4, 3, 5, 1, 0, 2, 7, 6, 8):      # hessP[4], [3], [5], [1], [0], [2], [7], [6], [8]
        cols.append(23 + idx)
    return cols


P_COLUMNS = _p_line_columns()


def primitives(W, gam):
    """rho, u, v, w, p of a padded conservative state (generateSplines order)."""
    den = W[0]
    u = W[1] / den
    v = W[2] / den
    w = W[3] / den
    Vsqr = u * u + v * v + w * w
    p = (den * (gam - 1)) * (W[4] / den - 0.5 * Vsqr)
    return den, u, v, w, p


def _fmt(x):
    return "%.15e" % x


def write_field(results_dir, W, it, gam, rank=0):
    """Results/Field/Field_<rank>_<it> of the padded state W (one GPU: sta = 0)."""
    d = os.path.join(results_dir, "Field")
    os.makedirs(d, exist_ok=True)
    den, u, v, w, p = primitives(W, gam)
    nc = W.shape[-1] - 6
    inner = (slice(3, nc + 3),) * 3
    cols = [a[inner].ravel() for a in (den, u, v, w, p)]
    j, i, k = np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc), indexing="ij")
    path = os.path.join(d, f"Field_{rank}_{it}")
    with open(path, "w") as f:
        for n, (x, y, z) in enumerate(zip(i.ravel(), j.ravel(), k.ravel())):
            f.write("%d %d %d %s %s %s %s %s\n" % (x, y, z, *(_fmt(c[n]) for c in cols)))
    return path


def write_reinit(path, W, gam):
    """InitialCondition/Init<a><b><c> (the Init input format) of the padded state W."""
    den, u, v, w, p = primitives(W, gam)
    nc = W.shape[-1] - 6
    inner = (slice(3, nc + 3),) * 3
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    icm.write_init(path, {"u": u[inner], "v": v[inner], "w": w[inner], "rho": den[inner], "p": p[inner]})
    return path


def write_particles(results_dir, it, pos, props, tags=None, rank=0):
    """Results/Particles/P_<rank>_<it>: pos (np x 3, (y, x, z)) and props (np x 41,
    lpt_properties columns) of the particles in tag order."""
    d = os.path.join(results_dir, "Particles")
    os.makedirs(d, exist_ok=True)
    n = pos.shape[0]
    tags = np.arange(n) if tags is None else tags
    path = os.path.join(d, f"P_{rank}_{it}")
    with open(path, "w") as f:
        for q in range(n):
            vals = [pos[q, c[1]] if isinstance(c, tuple) else props[q, c] for c in P_COLUMNS]
            f.write("%d " % tags[q] + " ".join(_fmt(v) for v in vals) + "\n")
    return path


def write_time(results_dir, it, t, dt, t0):
    """Results/Time/T_<it>: "t dt t0" without a newline."""
    d = os.path.join(results_dir, "Time")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f"T_{it}")
    with open(path, "w") as f:
        f.write("%s %s %s" % (_fmt(t), _fmt(dt), _fmt(t0)))
    return path
