"""Initial condition and snapshot formats on the device: quadrature points and
subcell means (HIP kernels) against the reference's golden initial states and
the host quadrature; regular_mesh and the element-layout transforms against
host evaluations of the same formulas (src/sd_simulator.py:59-131,
src/polynomials.py:124-148).

Bars: initial subcell means within 1e-12 of the reference's (the golden
fixtures; sin/exp of the device and of NumPy may differ by an ulp) and
bit-identical between a torch-evaluated callable and the same callable
evaluated on the host copy of the points; transforms within 1e-13 of the
host contractions.
"""
import numpy as np
import pytest

from fixtures import golden_names, load, make_ic, noise_of, relerr_fields, sim_kwargs

pytestmark = pytest.mark.gpu


def make(name, init_fct=None, **kw):
    from spd_amd.simulator import SDADER_Simulator
    fx = load(name)
    cfg = fx["cfg"]
    s = SDADER_Simulator(init_fct=init_fct or make_ic(cfg["ic"]), device="cuda:0",
                         verbose=False, **sim_kwargs(cfg), **kw)
    return fx, s


@pytest.mark.parametrize("name", golden_names())
def test_initial_state_vs_golden(name):
    print(f"[ic] {name}", flush=True)
    fx, s = make(name)
    nz = noise_of(fx["cfg"])
    if nz is not None:  # the fixture's initial means carry the seeded noise
        s.perturb_initial_condition(nz)
    assert relerr_fields(s.W_init_cv, fx["W_init_cv"]) < 1e-12
    assert relerr_fields(s.dm.U_sp, fx["U_sp0"]) < 1e-12


@pytest.mark.parametrize("name", ["kh16_p3", "sod_p3", "tg4_p3", "quad12_p4"])
def test_numpy_only_callable_same_means(name):
    """A callable that rejects tensors (NumPy-only user code) is evaluated on
    the host copy of the device quadrature points; the means are then the same
    as from the torch evaluation whenever the callable is exact arithmetic."""
    fx = load(name)
    f = make_ic(fx["cfg"]["ic"])

    def numpy_only(xyz, case):
        if not isinstance(xyz, np.ndarray):
            raise TypeError("NumPy arrays only")
        return f(xyz, case)

    _, a = make(name)
    _, b = make(name, init_fct=numpy_only)
    if fx["cfg"]["ic"]["kind"] in ("sod", "quad", "step"):
        np.testing.assert_array_equal(a.W_init_cv, b.W_init_cv)
    else:
        assert relerr_fields(a.W_init_cv, b.W_init_cv) < 1e-14


def host_regular(s, W):
    from oracle.spd_oracle import lagrange, solution_pts
    p = s.p
    x = np.arange(p + 2) / (p + 1)
    x = 0.5 * (x[1:] + x[:-1])
    m = lagrange(x, solution_pts(p))
    nd = s.ndim
    out = np.asarray(W, dtype=np.float64)
    for a in range(nd):
        ax = out.ndim - 1 - a
        out = np.moveaxis(np.tensordot(out, m, axes=([ax], [1])), -1, ax)
    return out


@pytest.mark.parametrize("name", ["kh16_p3", "sod_p3", "tg4_p3", "vortex12_p5"])
def test_regular_mesh_and_transforms(name):
    import torch
    fx, s = make(name)
    s.perform_iterations(1)
    W = np.asarray(s.dm.W_cv)
    ref = host_regular(s, W)
    got = s.regular_mesh(W)
    assert isinstance(got, np.ndarray)
    assert relerr_fields(got, ref) < 1e-13
    got_dev = s.regular_mesh(torch.from_numpy(W).to("cuda:0"))
    assert isinstance(got_dev, torch.Tensor)
    np.testing.assert_array_equal(got_dev.cpu().numpy(), got)
    # leading sizes other than the number of variables: one row, and 2 x V rows
    one = s.regular_mesh(W[:1])
    np.testing.assert_array_equal(one, got[:1])
    two = s.compute_cv_from_sp(np.stack([W, 2 * W]))
    np.testing.assert_array_equal(two[0], s.compute_cv_from_sp(W))
    with pytest.raises(ValueError):
        s.compute_cv_from_sp(W[..., :1])
