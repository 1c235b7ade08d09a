"""Exact solution of the 1-D Riemann problem of the ideal-gas Euler equations
(E. F. Toro, *Riemann Solvers and Numerical Methods for Fluid Dynamics*,
3rd ed., Springer 2009, ch. 4: pressure function, Newton iteration for the
star pressure, wave-pattern sampling).  Test infrastructure: the closed form
the Sod shock-tube runs are measured against."""
from __future__ import annotations

import numpy as np


def _f(p, rho, u, P, gamma):
    """Toro (4.6)-(4.7): f_K(p) and its derivative for one side."""
    a = np.sqrt(gamma * P / rho)
    if p > P:  # shock
        A = 2.0 / ((gamma + 1.0) * rho)
        B = (gamma - 1.0) / (gamma + 1.0) * P
        s = np.sqrt(A / (p + B))
        return (p - P) * s, s * (1.0 - 0.5 * (p - P) / (p + B))
    # rarefaction
    r = (p / P) ** ((gamma - 1.0) / (2.0 * gamma))
    return 2.0 * a / (gamma - 1.0) * (r - 1.0), (p / P) ** (-(gamma + 1.0) / (2.0 * gamma)) / (rho * a)


def star_state(L, R, gamma=1.4, tol=1e-14):
    """Star-region pressure and velocity (Newton iteration on f_L + f_R + du)."""
    rL, uL, PL = L
    rR, uR, PR = R
    p = max(1e-12, 0.5 * (PL + PR))
    for _ in range(200):
        fL, dL = _f(p, rL, uL, PL, gamma)
        fR, dR = _f(p, rR, uR, PR, gamma)
        g = fL + fR + (uR - uL)
        step = g / (dL + dR)
        p_new = max(1e-14, p - step)
        if abs(p_new - p) <= tol * 0.5 * (p_new + p):
            p = p_new
            break
        p = p_new
    fL, _ = _f(p, rL, uL, PL, gamma)
    fR, _ = _f(p, rR, uR, PR, gamma)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(x, t, L, R, x0=0.5, gamma=1.4):
    """Primitive state (rho, u, P) at positions ``x`` and time ``t > 0``
    (Toro section 4.5)."""
    rL, uL, PL = L
    rR, uR, PR = R
    ps, us = star_state(L, R, gamma)
    aL = np.sqrt(gamma * PL / rL)
    aR = np.sqrt(gamma * PR / rR)
    g1 = (gamma - 1.0) / (gamma + 1.0)
    out = np.empty((3, np.size(x)))
    for k, xi in enumerate(np.atleast_1d(x)):
        S = (xi - x0) / t
        if S <= us:  # left of the contact
            if ps > PL:  # left shock
                SL = uL - aL * np.sqrt((gamma + 1) / (2 * gamma) * ps / PL + (gamma - 1) / (2 * gamma))
                if S <= SL:
                    out[:, k] = (rL, uL, PL)
                else:
                    out[:, k] = (rL * (ps / PL + g1) / (g1 * ps / PL + 1.0), us, ps)
            else:  # left rarefaction
                asL = aL * (ps / PL) ** ((gamma - 1) / (2 * gamma))
                if S <= uL - aL:
                    out[:, k] = (rL, uL, PL)
                elif S >= us - asL:
                    out[:, k] = (rL * (ps / PL) ** (1.0 / gamma), us, ps)
                else:
                    c = 2.0 / (gamma + 1) + g1 / aL * (uL - S)
                    out[:, k] = (rL * c ** (2.0 / (gamma - 1)),
                                 2.0 / (gamma + 1) * (aL + (gamma - 1) / 2 * uL + S),
                                 PL * c ** (2 * gamma / (gamma - 1)))
        else:  # right of the contact
            if ps > PR:  # right shock
                SR = uR + aR * np.sqrt((gamma + 1) / (2 * gamma) * ps / PR + (gamma - 1) / (2 * gamma))
                if S >= SR:
                    out[:, k] = (rR, uR, PR)
                else:
                    out[:, k] = (rR * (ps / PR + g1) / (g1 * ps / PR + 1.0), us, ps)
            else:  # right rarefaction
                asR = aR * (ps / PR) ** ((gamma - 1) / (2 * gamma))
                if S >= uR + aR:
                    out[:, k] = (rR, uR, PR)
                elif S <= us + asR:
                    out[:, k] = (rR * (ps / PR) ** (1.0 / gamma), us, ps)
                else:
                    c = 2.0 / (gamma + 1) - g1 / aR * (uR - S)
                    out[:, k] = (rR * c ** (2.0 / (gamma - 1)),
                                 2.0 / (gamma + 1) * (-aR + (gamma - 1) / 2 * uR + S),
                                 PR * c ** (2 * gamma / (gamma - 1)))
    return out
