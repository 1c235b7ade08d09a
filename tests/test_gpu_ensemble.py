"""Batched ensemble (one HIP graph launch advances every member by a chunk of
steps; members spread over capture branches; launch shapes autotuned) against
the members run one by one with the host-driven loop.

The same kernels run on the same data in the same order within each member,
and the tuned launch shapes only regroup independent elements into blocks, so
the bar is bit-identical states, times and step counts.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def members(n, seed, sizes, orders):
    from spd_amd.ensemble import riemann_member_kwargs, riemann_members
    return [riemann_member_kwargs(m) for m in riemann_members(n, seed, sizes, orders)]


def solo(kw, **run):
    from spd_amd.simulator import SDADER_Simulator
    s = SDADER_Simulator(device="cuda:0", device_loop=False, **kw)
    if "t_end" in run:
        s.perform_time_evolution(run["t_end"])
    else:
        s.perform_iterations(run["steps"])
    return s


@pytest.mark.parametrize("run", [{"t_end": 0.03}, {"steps": 7}])
def test_ensemble_matches_solo_runs(run):
    from spd_amd.ensemble import Ensemble
    kws = members(6, 3, (8, 12, 16), (2, 3, 5))
    print(f"[ensemble] {run} sizes {[k['N'] for k in kws]} p {[k['p'] for k in kws]}",
          flush=True)
    ens = Ensemble(kws, device="cuda:0", graph_steps=4, branches=3, tune=True)
    ens.run(**run)
    torch.cuda.synchronize()
    assert ens.graph_builds >= 1 and ens.replays >= 1
    for i, kw in enumerate(kws):
        ref = solo(kw, **run)
        s = ens.sims[i]
        assert s.n_step == ref.n_step, i
        assert s.time == ref.time, i
        np.testing.assert_array_equal(s.dm.U_sp, ref.dm.U_sp)
        np.testing.assert_array_equal(s.dm.W_cv, ref.dm.W_cv)


def test_autotuner_shapes_do_not_change_results():
    """Every candidate launch shape of both keys gives the same step."""
    import ctypes
    from spd_amd import _lib
    from spd_amd.simulator import SDADER_Simulator
    kw = members(1, 5, (12,), (3,))[0]
    ref = SDADER_Simulator(device="cuda:0", **kw)
    ref.perform_iterations(2)
    from spd_amd.ensemble import TUNED
    for name, key in _lib.LAUNCH.items():
        if name not in TUNED:
            continue
        s0 = SDADER_Simulator(device="cuda:0", **kw)
        cur, mx = ctypes.c_int(), ctypes.c_int()
        _lib.check(s0._lib.spd_get_launch(s0._ctx, key, ctypes.byref(cur), ctypes.byref(mx)))
        for v in sorted({1, 2, mx.value // 2, mx.value} - {0}):
            print(f"[launch] {name}={v} (max {mx.value})", flush=True)
            s = SDADER_Simulator(device="cuda:0", **kw)
            _lib.check(s._lib.spd_set_launch(s._ctx, key, v))
            s.perform_iterations(2)
            np.testing.assert_array_equal(s.dm.U_sp, ref.dm.U_sp)
        assert s0._lib.spd_set_launch(s0._ctx, key, mx.value + 1) < 0
