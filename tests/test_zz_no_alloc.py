"""The allocation-free dt (spd_compute_dt_static, include/spd.h) through the
C ABI against spd_compute_dt on a context: the same kernels over the same
W_cv, so the same dt bit for bit (the bar for this reduction), on two calls
in a row (the second with a smaller input, so the static scratch holds the
first call's larger contents), for 1-D, 2-D and 3-D inputs; inputs beyond
the array's limit are rejected before anything runs."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
LIMIT = 1 << 25  # solution points the static scratch covers


def _sim(**kw):
    from spd_amd import initial_conditions as ic
    from spd_amd.simulator import SDADER_Simulator
    return SDADER_Simulator(init_fct=kw.pop("init_fct", ic.sine_wave), device="cuda:0",
                            verbose=False, **kw)


def _both(s, W):
    from spd_amd import _lib
    lib = _lib.load()
    st = _lib.stream_handle(torch.device("cuda:0"))
    a = torch.full((1,), -1.0, dtype=torch.float64, device="cuda:0")
    b = torch.full((1,), -2.0, dtype=torch.float64, device="cuda:0")
    _lib.check(lib.spd_compute_dt(s._ctx, _lib.ptr(W), _lib.ptr(a), st), "spd_compute_dt")
    _lib.check(lib.spd_compute_dt_static(ctypes.byref(s._cfg), _lib.ptr(W), _lib.ptr(b), st),
               "spd_compute_dt_static")
    torch.cuda.synchronize()
    return float(a.item()), float(b.item())


def test_static_dt_matches_context_dt():
    from spd_amd import initial_conditions as ic
    big = _sim(init_fct=ic.kelvin_helmholtz, p=3, N=(256, 256), update="FV", FB=True)
    small = _sim(p=2, N=(12, 10))
    sod = _sim(init_fct=ic.sod_shock_tube, p=3, N=(100,), BC=(("gradfree", "gradfree"),))
    tg = _sim(init_fct=ic.taylor_green, p=3, N=(6, 5, 4), xlim=(0, 6.3), ylim=(0, 6.3),
              zlim=(0, 6.3))
    for s in (big, small, sod, tg, big):
        print(f"[no-alloc dt] ndim={s.ndim} N={[s.N[d] for d in s.dims]} p={s.p}", flush=True)
        W = s.dm.device_value("W_cv").contiguous()
        # a non-trivial state: scale the velocities by a seeded factor field
        g = torch.Generator(device="cuda:0").manual_seed(s.ndim)
        W = W.clone()
        W[1:s.ndim + 1] *= 1 + torch.rand(W[1:s.ndim + 1].shape, generator=g, device="cuda:0",
                                          dtype=torch.float64)
        a, b = _both(s, W.contiguous())
        assert np.isfinite(a) and a > 0
        assert a == b, (a, b)


def test_static_dt_rejects_oversized_input():
    from spd_amd import _lib
    s = _sim(p=1, N=(4, 4))
    cfg = _lib.SpdConfig.from_buffer_copy(s._cfg)
    cfg.N[0], cfg.N[1], cfg.p = 2048, 2048, 3   # 2^26 solution points
    W = torch.zeros(4, dtype=torch.float64, device="cuda:0")
    dt = torch.full((1,), 7.0, dtype=torch.float64, device="cuda:0")
    lib = _lib.load()
    rc = lib.spd_compute_dt_static(ctypes.byref(cfg), _lib.ptr(W), _lib.ptr(dt),
                                   _lib.stream_handle(torch.device("cuda:0")))
    assert rc == -1
    assert str(LIMIT) in lib.spd_last_error().decode()
    torch.cuda.synchronize()
    assert float(dt.item()) == 7.0
    cfg.N[0], cfg.N[1] = 2048, 1024              # exactly the limit: accepted shape
    assert 2048 * 1024 * 16 == LIMIT
