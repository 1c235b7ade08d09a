"""The differentiable torch restatement (oracle/spd_torch_ref.py, the checker of
the HIP adjoint) against the NumPy oracle and against its own finite
differences (CPU).

Pinning: with the oracle's face coefficients of each substep, one torch step
equals one oracle step (which the golden vectors pin) to 1e-11 norm-wise
relative per field; its autograd gradient of <c, step(U)> matches central
differences on random directions.  The derivative checks start from the
fixture's state plus seeded noise: the step is piecewise smooth (limiter
cases, |v|, HLLC state selection) and the fixtures' exactly uniform regions
and zero velocities sit on those kinks, where no derivative exists.
"""
import numpy as np
import pytest
import torch

from fixtures import load, make_ic, noise_of, relerr_fields, sim_kwargs

NAMES = ["sod_p3", "sod_p1_llf", "sod_p3_lhllc_moncen", "sine1d_p3_FBG", "sine2d_p3_SD",
         "sine2d_p1_FV", "kh16_p3", "kh16_p3_moncen", "quad12_p2", "quad12_p3_lhllc_moncen",
         "step12_p2_refl", "vortex12_p5", "step3d_p2", "tg4_p3"]


def noisy(U0, seed=3, amp=1e-2):
    """U0 plus seeded noise of relative amplitude ``amp`` per field (momenta
    scaled by the larger of their own and the density's magnitude)."""
    rng = np.random.default_rng(seed)
    V = U0.shape[0]
    out = U0.copy()
    rho = np.max(np.abs(U0[0]))
    for v in range(V):
        sc = np.max(np.abs(U0[v]))
        if 1 <= v <= V - 2:
            sc = max(sc, rho)
        out[v] = U0[v] + amp * sc * rng.standard_normal(U0[v].shape)
    return out


def oracle_step(name, perturb=False):
    from oracle.spd_oracle import OracleSimulator
    fx = load(name)
    cfg = fx["cfg"]
    o = OracleSimulator(make_ic(cfg["ic"]), noise=noise_of(cfg), **sim_kwargs(cfg))
    if perturb:
        o.U_sp = noisy(o.U_sp)
        o.W_cv = o.prims(o.sp_to_cv(o.U_sp))
    o.compute_dt()
    U0, dt = o.U_sp.copy(), o.dt
    fb = o.update == "FV" and o.FB
    o.captures = [] if fb else None
    o.perform_update()
    aff = None
    if fb:
        aff = [[torch.as_tensor(r[f"aff_{'xyz'[a]}"]) for a in range(o.ndim)]
               for r in o.captures]
    return o, U0, dt, aff


@pytest.mark.parametrize("name", NAMES)
def test_forward_equals_oracle(name):
    from oracle.spd_torch_ref import FrozenStep
    o, U0, dt, aff = oracle_step(name)
    U1 = FrozenStep(o)(torch.as_tensor(U0), dt, aff)
    assert relerr_fields(U1.numpy(), o.U_sp) < 1e-11


@pytest.mark.parametrize("name", ["sod_p3", "sod_p3_lhllc_moncen", "quad12_p2", "step3d_p2"])
def test_autograd_vs_finite_differences(name):
    from oracle.spd_torch_ref import FrozenStep
    o, U0, dt, aff = oracle_step(name, perturb=True)
    fs = FrozenStep(o)
    rng = np.random.default_rng(1)
    c = torch.as_tensor(rng.standard_normal(U0.shape))
    U = torch.as_tensor(U0).clone().requires_grad_(True)
    (fs(U, dt, aff) * c).sum().backward()
    g = U.grad
    scale = torch.as_tensor(np.max(np.abs(U0.reshape(U0.shape[0], -1)), axis=1))
    scale = scale.reshape((-1,) + (1,) * (U0.ndim - 1))
    for k in range(3):
        d = torch.as_tensor(rng.standard_normal(U0.shape)) * scale
        eps = 1e-8
        with torch.no_grad():
            jp = (fs(torch.as_tensor(U0) + eps * d, dt, aff) * c).sum()
            jm = (fs(torch.as_tensor(U0) - eps * d, dt, aff) * c).sum()
        fd = float((jp - jm) / (2 * eps))
        ad = float((g * d).sum())
        assert abs(fd - ad) <= 1e-5 * abs(ad), (k, fd, ad)
