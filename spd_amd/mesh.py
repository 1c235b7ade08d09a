"""Host-side geometry helpers of the simulator facade: the point sets of the
reference (``src/polynomials.py:5-66``) and the initial-condition mesh of
``src/sd_simulator.py:59-72``.  They build small 1-D arrays once at
construction; the quadrature points, the subcell means and everything on the
time-step path are computed by the HIP kernels.
"""
from __future__ import annotations

import numpy as np


def gl01(n):
    """n-point Gauss-Legendre nodes / weights on [0, 1]."""
    if n == 0:
        return np.zeros(0), np.zeros(0)
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1.0) * 1.0 / 2.0, w * 1.0 / 2.0


def solution_points(p):
    i = np.arange(p + 1)
    return (1.0 - np.cos((2 * i + 1) / (2 * (p + 1)) * np.pi)) * 1.0 / 2.0


def flux_points(p):
    x, _ = gl01(p)
    return np.concatenate(([0.0], x, [1.0]))


def ic_mesh(N, p, lim, h, Nghe=1):
    """Subcell lower bounds and widths of the active elements per dimension,
    ``(N_a, p+1)`` arrays, from the reference's element mesh
    (``src/sd_simulator.py:59-72``): elements with ``Nghe`` ghost layers,
    faces ``lim0 + (k + fp) * L_ghosted / N_ghosted - h``."""
    fp = flux_points(p)
    lows, wids = [], []
    for a in range(len(N)):
        Ng = N[a] + 2 * Nghe
        length = (lim[a][1] - lim[a][0]) + 2 * Nghe * h[a]
        mesh = lim[a][0] + (np.arange(Ng)[:, None] + fp[None, :]) * length / Ng - h[a]
        mesh = mesh[Nghe:Ng - Nghe]
        lows.append(np.ascontiguousarray(mesh[:, :-1]))
        wids.append(np.ascontiguousarray(np.diff(mesh, axis=1)))
    return lows, wids
