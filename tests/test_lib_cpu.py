"""CPU checks of the C-ABI library: it loads, exports every function declared
in include/spd.h, its config struct matches the ctypes mirror, and its host-side
operator matrices and subcell geometry agree with the oracle (no GPU calls)."""
import os
import re

import numpy as np
import pytest

from fixtures import REPO

HEADER = os.path.join(REPO, "include", "spd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spd_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from spd_amd import _lib
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names


def test_abi_version_and_struct_size():
    import ctypes
    from spd_amd import _lib
    lib = _lib.load()
    assert lib.spd_abi_version() == 3
    assert lib.spd_config_size() == ctypes.sizeof(_lib.SpdConfig)


@pytest.mark.parametrize("p", range(1, 8))
def test_host_operators_match_oracle(p):
    from spd_amd import _lib
    from oracle.spd_oracle import Operators
    h = _lib.host_operators(p)
    o = Operators(p, p)
    for key in ("sp_to_fp", "dfp_to_sp", "sp_to_cv", "cv_to_sp", "invader", "w_tp", "x_tp"):
        ref = getattr(o, key)
        assert np.max(np.abs(h[key] - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref))), key


@pytest.mark.parametrize("p", range(1, 8))
def test_host_operator_invariants(p):
    """Closed-form checks: interpolation rows sum to 1, derivative rows to 0,
    exactness on monomials, sp<->cv inverse, ADER identities."""
    from spd_amd import _lib
    from spd_amd.mesh import solution_points, flux_points
    h = _lib.host_operators(p)
    sp, fp = solution_points(p), flux_points(p)
    assert np.allclose(h["sp_to_fp"].sum(1), 1.0, atol=1e-13)
    assert np.allclose(h["dfp_to_sp"].sum(1), 0.0, atol=1e-11)
    for k in range(p + 1):        # degree <= p is reproduced exactly
        assert np.allclose(h["sp_to_fp"] @ sp ** k, fp ** k, atol=1e-13)
    for k in range(p + 2):        # derivative of degree <= p+1 from the flux points
        d = k * sp ** (k - 1) if k else np.zeros_like(sp)
        assert np.allclose(h["dfp_to_sp"] @ fp ** k, d, atol=1e-10)
    for k in range(p + 1):        # subcell means of monomials
        means = (fp[1:] ** (k + 1) - fp[:-1] ** (k + 1)) / ((k + 1) * np.diff(fp))
        assert np.allclose(h["sp_to_cv"] @ sp ** k, means, atol=1e-13)
    assert np.allclose(h["sp_to_cv"] @ h["cv_to_sp"], np.eye(p + 1), atol=1e-12)
    # invader: sum_p invader[n, p] = x_n (exact for a constant time derivative)
    assert np.allclose(h["invader"].sum(1), h["x_tp"], atol=1e-13)
    assert abs(h["w_tp"].sum() - 1.0) < 1e-14


@pytest.mark.parametrize("N,h,p", [(10, 0.1, 3), (7, 0.3, 2), (100, 0.01, 3), (12, 10 / 12, 5)])
def test_host_geometry_matches_oracle(N, h, p):
    from spd_amd import _lib
    from oracle.spd_oracle import OracleSimulator
    import types
    g = _lib.host_geometry(N, h, p)
    ref = types.SimpleNamespace(ndim=1, N=[N], n=p + 1, h=[h], ngh=2)
    from oracle.spd_oracle import Operators
    ref.ops = Operators(p, p)
    OracleSimulator._geometry(ref)
    assert np.array_equal(g["faces"], ref.faces[0])
    assert np.array_equal(g["centers"], ref.centers[0])
    assert np.array_equal(g["h_fp"], ref.h_fp[0])
    assert np.array_equal(g["h_cv"], ref.h_cv[0])


def test_facade_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from spd_amd.simulator import SDADER_Simulator
    with pytest.raises(RuntimeError):
        SDADER_Simulator(p=1, N=(4,))


def test_unsupported_options_raise():
    from spd_amd.simulator import SDADER_Simulator
    for kw in (dict(viscosity=True), dict(WB=True), dict(slope_limiter="vanleer"),
               dict(riemann_solver_sd="roe"), dict(predictor=False)):
        with pytest.raises(NotImplementedError):
            SDADER_Simulator(p=1, N=(4,), **kw)


def test_subcell_means_match_oracle_ic():
    """The facade's host IC pipeline equals the oracle's (same quadrature)."""
    from fixtures import load, make_ic, sim_kwargs
    from spd_amd.mesh import subcell_means
    from oracle.spd_oracle import OracleSimulator
    for name in ("kh16_p3", "sod_p3", "vortex12_p5", "tg4_p3"):
        fx = load(name)
        kw = sim_kwargs(fx["cfg"])
        o = OracleSimulator(make_ic(fx["cfg"]["ic"]), **kw)
        nd = len(kw["N"])
        lims = [kw.get(k, (0, 1)) for k in ("xlim", "ylim", "zlim")][:nd]
        h = [(lims[a][1] - lims[a][0]) / kw["N"][a] for a in range(nd)]
        W = subcell_means(make_ic(fx["cfg"]["ic"]), nd + 2, list(kw["N"]), kw["p"], lims, h)
        assert np.array_equal(W, o.W_init_cv), name
