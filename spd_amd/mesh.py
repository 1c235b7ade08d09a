"""Host-side mesh and initial-condition helpers of the simulator facade.

These run once, on the host, at construction: the user's initial-condition
callable is NumPy code evaluated at quadrature points, exactly as in the
reference (``src/sd_simulator.py:59-90``, ``src/polynomials.py:124-148``).
Nothing here is on the time-step path.
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


def lagrange_matrix(x_to, x_from):
    """L[a, b] = l_b(x_to[a]) (used for ``regular_mesh``)."""
    x_to = np.asarray(x_to, dtype=np.float64)
    x_from = np.asarray(x_from, dtype=np.float64)
    diff_to = x_to[:, None, None] - x_from[None, None, :]
    diff_to = np.broadcast_to(diff_to, (x_to.size, x_from.size, x_from.size)).copy()
    idx = np.arange(x_from.size)
    diff_to[:, idx, idx] = 1.0
    den = x_from[:, None] - x_from[None, :]
    np.fill_diagonal(den, 1.0)
    return diff_to.prod(axis=-1) / den.prod(axis=-1)[None, :]


def subcell_means(init_fct, nvar, N, p, lim, h, Nghe=1):
    """Primitive subcell means of ``init_fct`` on the active elements, element
    layout ``(V, [Nz,] [Ny,] Nx, [nz,] [ny,] nx)``.

    Subcell bounds follow the reference's mesh (elements with ``Nghe`` ghost
    layers, ``lim0 + (k + fp) * L_ghosted / N_ghosted - h``); each subcell mean
    is a 2-point Gauss-Legendre rule per dimension.
    """
    nd = len(N)
    n = p + 1
    fp = flux_points(p)
    xq, wq = gl01(2)
    lows, wids = [], []
    for a in range(nd):
        Ng = N[a] + 2 * Nghe
        length = (lim[a][1] - lim[a][0]) + 2 * Nghe * h[a]
        mesh = lim[a][0] + (np.arange(Ng)[:, None] + fp[None, :]) * length / Ng - h[a]
        mesh = mesh[Nghe:Ng - Nghe]
        lows.append(mesh[:, :-1])
        wids.append(np.diff(mesh, axis=1))
    shapeE = [N[a] for a in range(nd)][::-1]
    full = shapeE + [n] * nd + [2] * nd
    out = np.empty([nvar] + shapeE + [n] * nd)
    pts = np.empty([nd] + full)
    for a in range(nd):
        k = nd - 1 - a
        shp = [1] * (3 * nd)
        shp[k] = N[a]
        shp[nd + k] = n
        qs = [1] * (3 * nd)
        qs[2 * nd + k] = 2
        pts[a] = lows[a].reshape(shp) + wids[a].reshape(shp) * xq.reshape(qs)
    for var in range(nvar):
        f = np.asarray(init_fct(pts, var), dtype=np.float64) * np.ones(full)
        args = [f, list(range(3 * nd))]
        for k in range(nd):
            args += [wq, [2 * nd + k]]
        args.append(list(range(2 * nd)))
        out[var] = np.einsum(*args)
    return out
