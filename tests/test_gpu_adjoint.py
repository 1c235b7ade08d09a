"""GPU: the hand-written adjoint of the step (spd_adj_step / spd_adj_backward,
through spd_amd.adjoint) against torch autograd through the differentiable
FP64 restatement (oracle/spd_torch_ref.py, on the CPU) and against central
finite differences of the frozen-flag step on the GPU.

The map differentiated is the step with its blending coefficients and dt
frozen (recorded by the forward pass).  States are the fixtures' plus seeded
noise (the exactly uniform regions of the fixtures sit on the kinks of the
limited scheme, where no derivative exists).

Bars: vector-Jacobian products within 1e-10 (max-norm relative) of torch
autograd; the FP32-bulk form's within 1e-5 (relative L1) of the FP64 form's; directional derivatives within 1e-5 relative of central
differences; the checkpointed multi-step gradient equal to the unsegmented
one to 1e-13 and within 1e-5 of central differences on 16 random directions.
The limited scheme has kinks (limiter cases, HLLC state selection, min/max)
densely around shock states: a kink within the difference step spoils that
difference, so each direction is differenced with steps 1e-8 and 1e-9 of the
field scale and the closer of the two is held to the bar.
"""
import numpy as np
import pytest
import torch

from fixtures import load, make_ic, noise_of, relerr_fields, sim_kwargs
from test_torch_ref import noisy

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")

NAMES = ["sod_p3", "sod_p1_llf", "sod_p3_lhllc_moncen", "sine1d_p3_FBG", "sine2d_p3_SD",
         "sine2d_p1_FV", "kh16_p3", "kh16_p3_moncen", "quad12_p2", "quad12_p3_lhllc_moncen",
         "step12_p2_refl", "vortex12_p5", "step3d_p2", "step3d_p2_moncen", "tg4_p3"]


def setup(name, seed=3):
    from oracle.spd_oracle import OracleSimulator
    from spd_amd.adjoint import Adjoint
    from spd_amd.simulator import SDADER_Simulator
    fx = load(name)
    cfg = fx["cfg"]
    s = SDADER_Simulator(init_fct=make_ic(cfg["ic"]), device="cuda:0", verbose=False,
                         **sim_kwargs(cfg))
    o = OracleSimulator(make_ic(cfg["ic"]), noise=noise_of(cfg), **sim_kwargs(cfg))
    U0 = noisy(o.U_sp, seed=seed)
    o.U_sp = U0.copy()
    o.W_cv = o.prims(o.sp_to_cv(U0))
    o.compute_dt()
    return s, o, Adjoint(s), U0, o.dt


def frozen_coefficients(s, o, th):
    from oracle.spd_torch_ref import face_coefficients
    Ng = [s.N[d] * (s.p + 1) + 4 for d in s.dims][::-1]
    thc = th.cpu().numpy().reshape([o.T] + Ng)
    return [face_coefficients(o, thc[i]) for i in range(o.T)]


def field_scale(U0):
    sc = np.max(np.abs(U0.reshape(U0.shape[0], -1)), axis=1)
    sc[1:-1] = max(np.max(sc[1:-1]), sc[0])
    return sc.reshape((-1,) + (1,) * (U0.ndim - 1))


@pytest.mark.parametrize("name", NAMES)
def test_vjp_vs_autograd(name):
    from oracle.spd_torch_ref import FrozenStep
    print(f"[gpu adjoint] vjp {name}", flush=True)
    s, o, adj, U0, dt = setup(name)
    Ud = torch.as_tensor(U0, device=DEV)
    dtd = torch.full((1,), dt, dtype=torch.float64, device=DEV)
    if adj.fb:
        U1, th = adj.step(Ud, dtd, record_theta=True)
        aff = frozen_coefficients(s, o, th)
    else:
        U1, aff = adj.step(Ud, dtd), None
    c = np.random.default_rng(7).standard_normal(U0.shape)
    G = adj.step_vjp(Ud, dtd, torch.as_tensor(c, device=DEV)).cpu().numpy()
    U = torch.as_tensor(U0).clone().requires_grad_(True)
    U1c = FrozenStep(o)(U, dt, aff)
    assert relerr_fields(U1.cpu().numpy(), U1c.detach().numpy()) < 1e-12
    (U1c * torch.as_tensor(c)).sum().backward()
    ref = U.grad.numpy()
    err = float(np.max(np.abs(G - ref)) / np.max(np.abs(ref)))
    print(f"  max-norm relative difference to autograd: {err:.2e}", flush=True)
    assert err < 1e-10, err


@pytest.mark.parametrize("name", ["sod_p3", "kh16_p3", "quad12_p3_lhllc_moncen", "step12_p2_refl",
                                  "sine2d_p3_SD", "step3d_p2", "tg4_p3"])
def test_vjp_vs_finite_differences(name):
    print(f"[gpu adjoint] finite differences {name}", flush=True)
    s, o, adj, U0, dt = setup(name)
    Ud = torch.as_tensor(U0, device=DEV)
    dtd = torch.full((1,), dt, dtype=torch.float64, device=DEV)
    th = adj.step(Ud, dtd, record_theta=True)[1] if adj.fb else None
    rng = np.random.default_rng(11)
    c = torch.as_tensor(rng.standard_normal(U0.shape), device=DEV)
    G = adj.step_vjp(Ud, dtd, c, theta=th)
    scale = torch.as_tensor(field_scale(U0), device=DEV)
    for k in range(4):
        d = torch.as_tensor(rng.standard_normal(U0.shape), device=DEV) * scale
        ad = float((G * d).sum())
        errs = []
        for eps in (1e-8, 1e-9):
            dj = float(((adj.step(Ud + eps * d, dtd, theta=th)
                         - adj.step(Ud - eps * d, dtd, theta=th)) * c).sum())
            errs.append(abs(dj / (2 * eps) - ad) / abs(ad))
        assert min(errs) <= 1e-5, (k, errs, ad)


def kinetic_energy_points(U, ndim):
    """|m|^2 / (2 rho) at every solution point."""
    return 0.5 * sum(U[1 + a] ** 2 for a in range(ndim)) / U[0]


def kinetic_energy(U, ndim):
    return kinetic_energy_points(U, ndim).sum()


def cons_from_density(rho, W0, gamma, ndim):
    """Conservative state from the density field and fixed velocities/pressure."""
    vel = [W0[1 + a] for a in range(ndim)]
    E = W0[ndim + 1] / (gamma - 1) + 0.5 * rho * sum(v * v for v in vel)
    return torch.stack([rho] + [rho * v for v in vel] + [E])


@pytest.mark.parametrize("name,nsteps,k", [("sod_p3", 6, 4), ("kh16_p3", 7, 3),
                                           ("quad12_p2", 5, 2), ("tg4_p3", 3, 2)])
def test_checkpointed_gradient_of_kinetic_energy(name, nsteps, k):
    print(f"[gpu adjoint] run {name} nsteps={nsteps} k={k}", flush=True)
    s, o, adj, U0, dt = setup(name)
    nd, gam = s.ndim, s.gamma
    W0 = torch.as_tensor(o.prims(U0), device=DEV)
    grads = []
    for kk in (k, nsteps):
        rho = W0[0].clone().requires_grad_(True)
        UN = adj.run(cons_from_density(rho, W0, gam, nd).contiguous(), nsteps, checkpoint_every=kk)
        kinetic_energy(UN, nd).backward()
        grads.append(rho.grad.clone())
    g = grads[0]
    assert float((g - grads[1]).abs().max() / g.abs().max()) < 1e-13
    # the same with the checkpoints and interval states in host memory
    from spd_amd.adjoint import Adjoint
    adj_h = Adjoint(s, offload=True)
    rho = W0[0].clone().requires_grad_(True)
    UN = adj_h.run(cons_from_density(rho, W0, gam, nd).contiguous(), nsteps, checkpoint_every=k)
    kinetic_energy(UN, nd).backward()
    assert float((g - rho.grad).abs().max() / g.abs().max()) < 1e-13
    # central differences of the same frozen map: the forward's dt and
    # coefficients of every step replayed
    with torch.no_grad():
        U = cons_from_density(W0[0], W0, gam, nd).contiguous()
        _, _, dts = adj._forward(U, nsteps, nsteps, None)
        thetas = []
        for st in range(nsteps):
            if adj.fb:
                U, th = adj.step(U, dts[st:st + 1], record_theta=True)
            else:
                U, th = adj.step(U, dts[st:st + 1]), None
            thetas.append(th)

        def final(rho):
            U = cons_from_density(rho, W0, gam, nd).contiguous()
            for st in range(nsteps):
                U = adj.step(U, dts[st:st + 1], theta=thetas[st])
            return kinetic_energy_points(U, nd)

        def dJ(d, eps):
            # pointwise difference first, then the sum (the sum of the
            # energies would carry round-off of the order of the difference)
            return float((final(W0[0] + eps * d) - final(W0[0] - eps * d)).sum())

        rng = np.random.default_rng(5)
        rs = float(W0[0].abs().max())
        worst = 0.0
        for _ in range(16):
            d = torch.as_tensor(rng.standard_normal(tuple(W0[0].shape)), device=DEV) * rs
            ad = float((g * d).sum())
            errs = [abs(dJ(d, eps) / (2 * eps) - ad) / abs(ad) for eps in (1e-8, 1e-9)]
            worst = max(worst, min(errs))
        print(f"  worst relative difference to central differences: {worst:.2e}", flush=True)
        assert worst <= 1e-5, worst


@pytest.mark.parametrize("name", ["sod_p3", "kh16_p3", "step3d_p2"])
def test_fp32_form_vjp_close_to_fp64(name):
    """The FP32-bulk form's VJP (FP64 reverse sweep over the form's FP32
    predictor iterates) against the FP64 form's at the same state and
    cotangent: relative L1 difference below 1e-5."""
    from spd_amd.adjoint import Adjoint
    from spd_amd.simulator import SDADER_Simulator
    print(f"[gpu adjoint] fp32 vs fp64 {name}", flush=True)
    s64, o, adj64, U0, dt = setup(name)
    cfg = load(name)["cfg"]
    s32 = SDADER_Simulator(init_fct=make_ic(cfg["ic"]), device="cuda:0", verbose=False,
                           precision="fp32", **sim_kwargs(cfg))
    adj32 = Adjoint(s32)
    Ud = torch.as_tensor(U0, device=DEV)
    dtd = torch.full((1,), dt, dtype=torch.float64, device=DEV)
    c = torch.as_tensor(np.random.default_rng(7).standard_normal(U0.shape), device=DEV)
    g64 = adj64.step_vjp(Ud, dtd, c)
    g32 = adj32.step_vjp(Ud, dtd, c)
    rel = float((g32 - g64).abs().sum() / g64.abs().sum())
    print(f"  relative L1 difference fp32 vs fp64: {rel:.2e}", flush=True)
    assert rel < 1e-5, rel


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["kh16_p3", "quad12_p2", "step12_p2_refl", "vortex12_p5",
                                  "sine2d_p3_SD", "tg4_p3"])
def test_taped_step_equals_product_step(name, precision):
    """The product step of the 2-D path writes the first Picard iterate in
    compact form (its one-node increment) when another iteration follows, and
    the next iteration forms every node from U_sp and that increment; the
    taped step of the derivative keeps the full iterates.  Both form the same
    values with the same operations in the same order, so one step from the
    same state and dt agrees bit for bit."""
    from spd_amd.adjoint import Adjoint
    from spd_amd.simulator import SDADER_Simulator
    print(f"[gpu taped step] {name} {precision}", flush=True)
    cfg = load(name)["cfg"]
    s = SDADER_Simulator(init_fct=make_ic(cfg["ic"]), device="cuda:0", verbose=False,
                         precision=precision, **sim_kwargs(cfg))
    adj = Adjoint(s)
    U0 = s.dm.device_value("U_sp").clone().contiguous()
    dt = adj.dt_of(U0)
    taped = adj.step(U0, dt)
    U = U0.clone()
    adj._advance(U, dt)
    torch.cuda.synchronize()
    assert torch.isfinite(U).all()
    assert torch.equal(taped, U), float((taped - U).abs().max())


@pytest.mark.parametrize("name", ["kh16_p3", "quad12_p2"])
def test_adjoint_leaves_simulator_dt_alone(name):
    """The adjoint runs its steps on the simulator's context; the maximum wave
    speed that context records for the simulator's next dt (after an FV step
    the simulator takes dt from it, not from dm.W_cv) must survive every
    adjoint call: perform_update, then adjoint calls from another state, then
    compute_dt gives the dt of a simulator that made no adjoint call."""
    from spd_amd.adjoint import Adjoint
    from spd_amd.simulator import SDADER_Simulator
    print(f"[adjoint keeps sim dt] {name}", flush=True)
    fx = load(name)
    cfg = fx["cfg"]
    sims = []
    for _ in range(2):
        s = SDADER_Simulator(init_fct=make_ic(cfg["ic"]), device="cuda:0", verbose=False,
                             **sim_kwargs(cfg))
        nz = noise_of(cfg)
        if nz is not None:
            s.perturb_initial_condition(nz)
        s.compute_dt()
        s.perform_update()
        sims.append(s)
    a, b = sims
    adj = Adjoint(b)
    U = torch.as_tensor(noisy(np.asarray(b.dm.U_sp), seed=7), device=DEV)
    dt = adj.dt_of(U)
    adj.step_vjp(U, dt, torch.ones_like(U))
    adj.run(U.clone().requires_grad_(True), 3, checkpoint_every=2).sum().backward()
    assert a.compute_dt() == b.compute_dt()
    a.perform_update()
    b.perform_update()
    assert a.compute_dt() == b.compute_dt()
    np.testing.assert_array_equal(a.dm.U_sp, b.dm.U_sp)
