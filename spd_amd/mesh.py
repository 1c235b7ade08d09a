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


def ic_mesh(N, p, lim, h, Nghe=1, ghosts=False):
    """Subcell lower bounds and widths of the active elements per dimension,
    ``(N_a, p+1)`` arrays (with ``ghosts``: of all ``N_a + 2 Nghe`` elements),
    from the reference's element mesh (``src/sd_simulator.py:59-72``):
    elements with ``Nghe`` ghost layers, faces
    ``lim0 + (k + fp) * L_ghosted / N_ghosted - h``."""
    fp = flux_points(p)
    lows, wids = [], []
    for a in range(len(N)):
        Ng = N[a] + 2 * Nghe
        length = (lim[a][1] - lim[a][0]) + 2 * Nghe * h[a]
        mesh = lim[a][0] + (np.arange(Ng)[:, None] + fp[None, :]) * length / Ng - h[a]
        if not ghosts:
            mesh = mesh[Nghe:Ng - Nghe]
        lows.append(np.ascontiguousarray(mesh[:, :-1]))
        wids.append(np.ascontiguousarray(np.diff(mesh, axis=1)))
    return lows, wids


def ghost_layer_means(init_fct, N, p, lim, h, nvar, Nghe=1, ngh=2):
    """Primitive subcell means of ``init_fct`` on the ghosted finite-volume
    grid, ``(nvar, Ns_z+2 ngh, Ns_y+2 ngh, Ns_x+2 ngh)``: the 2-point
    Gauss-Legendre means of the element mesh with ``Nghe`` ghost elements per
    side (``src/sd_simulator.py:74-90``, ``src/polynomials.py:124-148``),
    transposed to the FV layout and cut to ``ngh`` ghost layers
    (``src/fv_simulator.py:213-222``).  Evaluated once on the host with NumPy
    (the ghost layers of "ic" / "pressure" sides; the reference evaluates the
    same means with NumPy)."""
    nd = len(N)
    n = p + 1
    lows, wids = ic_mesh(N, p, lim, h, Nghe, ghosts=True)
    xq, wq = gl01(2)
    NE = [lows[a].shape[0] for a in range(nd)]
    full = NE[::-1] + [n] * nd + [2] * nd
    pts = np.empty([nd] + full)
    for a in range(nd):
        k = nd - 1 - a
        shp = [1] * (3 * nd)
        shp[k] = NE[a]
        shp[nd + k] = n
        qs = [1] * (3 * nd)
        qs[2 * nd + k] = 2
        pts[a] = np.broadcast_to(lows[a].reshape(shp) + wids[a].reshape(shp) * xq.reshape(qs), full)
    W = np.empty([nvar] + NE[::-1] + [n] * nd)
    for var in range(nvar):
        f = np.asarray(init_fct(pts, var), dtype=np.float64) * np.ones(full)
        args = [f, list(range(3 * nd))]
        for k in range(nd):
            args += [wq, [2 * nd + k]]
        args.append(list(range(2 * nd)))
        W[var] = np.einsum(*args)
    # element layout -> FV layout, then ngh ghost layers
    perm = [0]
    for k in range(nd):
        perm += [1 + k, 1 + nd + k]
    Wf = np.transpose(W, perm).reshape([nvar] + [NE[nd - 1 - k] * n for k in range(nd)])
    cut = n * Nghe - ngh
    if cut > 0:
        Wf = Wf[(slice(None),) + (slice(cut, -cut),) * nd]
    return np.ascontiguousarray(Wf)
