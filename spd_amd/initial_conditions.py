"""Initial-condition callables in the simulator's convention.

Every callable has the signature ``init_fct(xyz, case) -> ndarray`` where
``xyz`` is an array of shape ``(ndim, ...)`` holding point coordinates and
``case`` is the index of the primitive variable requested:
``0`` density, ``1..ndim`` velocity components, ``ndim+1`` pressure.
This is the convention of ``src/initial_conditions.py:6-13`` and
``src/initial_conditions_{1d,2d,3d}.py`` in the reference; the simulator
averages the callable over every subcell with a 2-point Gauss-Legendre rule
per dimension (``src/polynomials.py:124-148``).

The callables here accept NumPy arrays or torch tensors (the simulator
evaluates them on the GPU at the quadrature points it computes there);
user callables written for NumPy are evaluated on the host instead.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "sine_wave", "step_function", "sod_shock_tube", "sod",
    "kelvin_helmholtz", "KH_instability", "isentropic_vortex",
    "four_quadrant", "random_four_quadrant_states", "taylor_green",
    "subcell_noise",
]


def _nd(xyz):
    return xyz.shape[0]


class _TorchNS:
    """The few NumPy functions the callables use, on torch tensors."""

    def __init__(self):
        import torch
        self.t = torch
        self.pi = np.pi

    def where(self, c, a, b):
        # Python numbers become FP64 tensors (torch.where would otherwise
        # produce the default FP32 dtype for two scalar operands)
        def f64(v):
            if isinstance(v, self.t.Tensor):
                return v
            return self.t.tensor(v, dtype=self.t.float64, device=c.device)
        return self.t.where(c, f64(a), f64(b))

    def sin(self, x):
        return self.t.sin(x)

    def cos(self, x):
        return self.t.cos(x)

    def exp(self, x):
        return self.t.exp(x)

    def abs(self, x):
        return self.t.abs(x)

    def round(self, x):
        return self.t.round(x)

    def ones_like(self, x):
        return self.t.ones_like(x)

    def zeros_like(self, x):
        return self.t.zeros_like(x)


_TORCH_NS = None


def _xp(x):
    """Array namespace of ``x``: NumPy, or the torch adapter for tensors."""
    global _TORCH_NS
    if type(x).__module__.startswith("torch"):
        if _TORCH_NS is None:
            _TORCH_NS = _TorchNS()
        return _TORCH_NS
    return np


def sine_wave(xyz: np.ndarray, case: int, A=0.125, vx=1.0, vy=1.0, vz=1.0, P=1.0):
    """Density sine wave advected at constant velocity and pressure.

    Follows ``src/initial_conditions_3d.py:25-56`` (and the 1-D/2-D variants it
    dispatches to): rho = 1 + A sin(2 pi (x [+ y [+ z]])), v = const, P = const.
    In 1-D only ``vx`` is used, in 2-D ``vx, vy``.
    """
    xp = _xp(xyz)
    nd = _nd(xyz)
    x = xyz[0]
    vel = (vx, vy, vz)[:nd]
    if case == 0:
        s = xyz[0]
        for d in range(1, nd):
            s = s + xyz[d]
        return 1.0 + A * xp.sin(2 * np.pi * s)
    if 1 <= case <= nd:
        return vel[case - 1] * xp.ones_like(x)
    if case == nd + 1:
        return P * xp.ones_like(x)
    return xp.ones_like(x)


def step_function(xyz: np.ndarray, case: int, vx=1.0, vy=1.0, vz=1.0, P=1.0):
    """Density box (2 inside |x_d - 0.5| < 0.25 for all d, 1 outside).

    Follows ``src/initial_conditions_3d.py:5-23`` and its 1-D/2-D variants.
    """
    xp = _xp(xyz)
    nd = _nd(xyz)
    x = xyz[0]
    vel = (vx, vy, vz)[:nd]
    if case == 0:
        inside = xp.abs(xyz[0] - 0.5) < 0.25
        for d in range(1, nd):
            inside = inside & (xp.abs(xyz[d] - 0.5) < 0.25)
        return xp.where(inside, 2.0, 1.0)
    if 1 <= case <= nd:
        return vel[case - 1] * xp.ones_like(x)
    if case == nd + 1:
        return P * xp.ones_like(x)
    return xp.ones_like(x)


def sod_shock_tube(xyz: np.ndarray, case: int):
    """1-D Sod tube: (rho, u, P) = (1, 0, 1) for x < 0.5, (0.125, 0, 0.1) else.

    Same states as ``src/initial_conditions_1d.py:31-40``.
    """
    xp = _xp(xyz)
    x = xyz[0]
    if case == 0:
        return xp.where(x < 0.5, 1.0, 0.125)
    if case == _nd(xyz) + 1:
        return xp.where(x < 0.5, 1.0, 0.1)
    return xp.zeros_like(x)


sod = sod_shock_tube


def kelvin_helmholtz(xyz: np.ndarray, case: int, rho_in=2.0, rho_out=1.0,
                     v_in=0.5, v_out=-0.5, P=2.5, A=0.01):
    """Two-layer shear flow on [0,1]^2.

    rho = rho_in and vx = v_in for |y - 0.5| < 0.25, else rho_out / v_out;
    vy = A sin(4 pi x); P constant.  (The seeded velocity noise of the
    benchmark is added to the subcell means afterwards, see ``subcell_noise``.)
    """
    xp = _xp(xyz)
    x, y = xyz[0], xyz[1]
    inner = xp.abs(y - 0.5) < 0.25
    if case == 0:
        return xp.where(inner, rho_in, rho_out)
    if case == 1:
        return xp.where(inner, v_in, v_out)
    if case == 2:
        return A * xp.sin(4 * np.pi * x)
    if case == 3:
        return P * xp.ones_like(x)
    return xp.zeros_like(x)


def KH_instability(xyz: np.ndarray, case: int):
    """The reference's own Kelvin-Helmholtz set-up (Gaussian-localised vy),
    ``src/initial_conditions_2d.py:40-52``."""
    xp = _xp(xyz)
    y = xyz[1]
    w0 = 0.1
    sigma = 0.05 / np.sqrt(2)
    if case == 0:
        return xp.where(y < 0.25, 1.0, xp.where(y < 0.75, 2.0, 1.0))
    if case == 1:
        return xp.where(y < 0.25, -0.5, xp.where(y < 0.75, 0.5, -0.5))
    if case == 2:
        return w0 * xp.sin(4 * np.pi * xyz[0]) * (
            xp.exp(-(y - 0.25) ** 2 / (2 * sigma ** 2))
            + xp.exp(-(y - 0.75) ** 2 / (2 * sigma ** 2)))
    if case == 3:
        return 2.5 * xp.ones_like(y)
    return xp.zeros_like(xyz[0])


def isentropic_vortex(xyz: np.ndarray, case: int, beta=5.0, gamma=1.4,
                      u0=1.0, v0=1.0, x0=0.0, y0=0.0, t=0.0,
                      period=(10.0, 10.0)):
    """Shu's isentropic vortex (exact solution of 2-D Euler).

    Background rho = 1, (u, v) = (u0, v0), P = 1; perturbations
    du = -beta/(2 pi) e^{(1-r^2)/2} (y - yc), dv = beta/(2 pi) e^{(1-r^2)/2} (x - xc),
    T = 1 - (gamma-1) beta^2 / (8 gamma pi^2) e^{1-r^2},
    rho = T^{1/(gamma-1)}, P = rho T.  With ``t > 0`` the vortex centre is
    advected to (x0 + u0 t, y0 + v0 t), wrapped into the periodic box.
    """
    xp = _xp(xyz)
    x, y = xyz[0], xyz[1]
    Lx, Ly = period
    xc = x0 + u0 * t
    yc = y0 + v0 * t
    dx = x - xc
    dy = y - yc
    dx = dx - Lx * xp.round(dx / Lx)
    dy = dy - Ly * xp.round(dy / Ly)
    r2 = dx * dx + dy * dy
    f = xp.exp(0.5 * (1.0 - r2))
    if case == 1:
        return u0 - beta / (2 * np.pi) * f * dy
    if case == 2:
        return v0 + beta / (2 * np.pi) * f * dx
    T = 1.0 - (gamma - 1.0) * beta * beta / (8.0 * gamma * np.pi * np.pi) * f * f
    rho = T ** (1.0 / (gamma - 1.0))
    if case == 0:
        return rho
    if case == 3:
        return rho * T
    return xp.zeros_like(x)


def random_four_quadrant_states(rng: np.random.Generator):
    """Draw the four quadrant states of a 2-D Riemann problem.

    Returns an array ``(4, 4)`` indexed ``[quadrant, var]`` with quadrants
    ordered (x<.5,y<.5), (x>=.5,y<.5), (x<.5,y>=.5), (x>=.5,y>=.5) and variables
    (rho, vx, vy, P): rho ~ U[0.1, 2], v ~ U[-1, 1], P ~ U[0.05, 1.5].
    """
    s = np.empty((4, 4))
    for q in range(4):
        s[q, 0] = rng.uniform(0.1, 2.0)
        s[q, 1] = rng.uniform(-1.0, 1.0)
        s[q, 2] = rng.uniform(-1.0, 1.0)
        s[q, 3] = rng.uniform(0.05, 1.5)
    return s


def four_quadrant(states, xc=0.5, yc=0.5):
    """IC callable of a four-quadrant 2-D Riemann problem with the given states."""
    states = np.asarray(states, dtype=np.float64)

    def fct(xyz: np.ndarray, case: int):
        x, y = xyz[0], xyz[1]
        if _xp(xyz) is not np:
            import torch
            q = (x >= xc).long() + 2 * (y >= yc).long()
            return torch.as_tensor(states[:, case], dtype=torch.float64, device=x.device)[q]
        q = (x >= xc).astype(np.int64) + 2 * (y >= yc).astype(np.int64)
        return states[:, case][q]

    return fct


def taylor_green(xyz: np.ndarray, case: int, P0=100.0):
    """Compressible Taylor-Green flow on [0, 2 pi]^3.

    rho = 1; u = sin x cos y cos z, v = -cos x sin y cos z, w = 0;
    P = P0 + (cos 2x + cos 2y)(cos 2z + 2)/16.
    """
    xp = _xp(xyz)
    x, y, z = xyz[0], xyz[1], xyz[2]
    if case == 0:
        return xp.ones_like(x)
    if case == 1:
        return xp.sin(x) * xp.cos(y) * xp.cos(z)
    if case == 2:
        return -xp.cos(x) * xp.sin(y) * xp.cos(z)
    if case == 3:
        return xp.zeros_like(x)
    if case == 4:
        return P0 + (xp.cos(2 * x) + xp.cos(2 * y)) * (xp.cos(2 * z) + 2.0) / 16.0
    return xp.zeros_like(x)


def subcell_noise(N, p, variables, amplitude, seed=0, nvar=None):
    """Seeded uniform noise on subcell means, in the simulator's element layout.

    ``N`` is the tuple of element counts (x first), ``p`` the order.  The draw is
    ``numpy.random.default_rng(seed).uniform(-amplitude, amplitude,
    size=(len(variables),) + subcell_grid)`` where ``subcell_grid`` is the
    finite-volume grid ``(..., Ny*(p+1), Nx*(p+1))``; it is then reshaped to
    the element layout ``(var, [Nz,] Ny, Nx, [nz,] ny, nx)`` and scattered into
    an array of ``nvar`` variables (zeros for variables not listed).
    """
    ndim = len(N)
    n = p + 1
    if nvar is None:
        nvar = ndim + 2
    rng = np.random.default_rng(seed)
    grid = tuple(N[d] * n for d in range(ndim))[::-1]
    draw = rng.uniform(-amplitude, amplitude, size=(len(variables),) + grid)
    out = np.zeros((nvar,) + grid)
    for k, v in enumerate(variables):
        out[v] = draw[k]
    # finite-volume layout -> element layout
    Ns = list(N)[::-1]
    shape = [nvar]
    for d in range(ndim):
        shape += [Ns[d], n]
    out = out.reshape(shape)
    if ndim == 1:
        return out
    if ndim == 2:
        return np.ascontiguousarray(np.transpose(out, (0, 1, 3, 2, 4)))
    return np.ascontiguousarray(np.transpose(out, (0, 1, 3, 5, 2, 4, 6)))
