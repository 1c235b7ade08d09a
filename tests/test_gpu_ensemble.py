"""Batched ensemble (one HIP graph launch advances every member by a chunk of
steps; members of one order share one launch per stage of the step, or keep
their own launches spread over capture branches with autotuned shapes)
against the members run one by one with the host-driven loop, and against
the CPU oracle.

The same kernels run on the same data in the same order within each member
(batched launches run each member's blocks with the member's arguments; the
tuned launch shapes only regroup independent elements into blocks), so the
bar is bit-identical states, times and step counts against the solo runs with
the same substep form; a member is non-finite only where the CPU oracle is
too, and the members' states are held to the oracle at 1e-12.
The non-finite members of the small-member benchmark draw are those the
reference scheme itself drives to non-finite values (the census that
tests/golden/make_golden.py recorded from the reference).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def members(n, seed, sizes, orders, precision="fp64"):
    from spd_amd.ensemble import riemann_member_kwargs, riemann_members
    return [riemann_member_kwargs(m, precision) for m in riemann_members(n, seed, sizes, orders)]


def solo(kw, form=0, **run):
    from spd_amd import _lib
    from spd_amd.simulator import SDADER_Simulator
    s = SDADER_Simulator(device="cuda:0", device_loop=False, **kw)
    if form:
        _lib.check(s._lib.spd_set_launch(s._ctx, _lib.LAUNCH["fb_form"], form), "spd_set_launch")
    if "t_end" in run:
        s.perform_time_evolution(run["t_end"])
    else:
        s.perform_iterations(run["steps"])
    return s


@pytest.mark.parametrize("batched", [True, False])
@pytest.mark.parametrize("run", [{"t_end": 0.03}, {"steps": 7}])
def test_ensemble_matches_solo_runs(run, batched):
    from spd_amd.ensemble import Ensemble
    kws = members(6, 3, (8, 12, 16), (2, 3, 5))
    print(f"[ensemble batched={batched}] {run} sizes {[k['N'] for k in kws]} "
          f"p {[k['p'] for k in kws]}", flush=True)
    ens = Ensemble(kws, device="cuda:0", graph_steps=4, branches=3, tune=True, batched=batched)
    if batched:
        assert sum(len(g) for g in ens.groups) == len(kws) and len(ens.groups) == 3
    ens.run(**run)
    torch.cuda.synchronize()
    assert ens.graph_builds >= 1 and ens.replays >= 1
    for i, kw in enumerate(kws):
        ref = solo(kw, form=2 if batched else 0, **run)
        s = ens.sims[i]
        assert s.n_step == ref.n_step, i
        assert s.time == ref.time, i
        np.testing.assert_array_equal(s.dm.U_sp, ref.dm.U_sp)
        np.testing.assert_array_equal(s.dm.W_cv, ref.dm.W_cv)
    # one member against the oracle: the largest order among the members
    # that stay physical (one member of this draw collapses: its dt falls
    # below 1e-14 and the run stops early, as the reference's does; on such
    # a state round-off is amplified without bound)
    from fixtures import relerr_fields
    from oracle.spd_oracle import OracleSimulator

    def oracle_run(j):
        kw = dict(kws[j])
        init = kw.pop("init_fct")
        for key in ("precision", "verbose"):
            kw.pop(key, None)
        o = OracleSimulator(init, **kw)
        with np.errstate(all="ignore"):
            if "t_end" in run:
                o.perform_time_evolution(run["t_end"])
            else:
                for _ in range(run["steps"]):
                    o.compute_dt()
                    o.perform_update()
        return o

    # a member with a non-finite state is one the reference scheme also
    # drives to non-finite values
    for j, s in enumerate(ens.sims):
        if not np.all(np.isfinite(s.dm.U_sp)):
            assert not np.all(np.isfinite(oracle_run(j).U_sp)), j

    def physical(j):
        s = ens.sims[j]
        if "t_end" in run:
            done = s.time == run["t_end"]
        else:
            done = s.n_step == run["steps"] and s.dm.dt > 1e-10
        return done and float(np.min(s.dm.U_sp[0])) > 0.0
    ok = [j for j in range(len(kws)) if physical(j)]
    assert ok
    i = max(ok, key=lambda j: (kws[j]["p"], kws[j]["N"]))
    o = oracle_run(i)
    assert relerr_fields(ens.sims[i].dm.U_sp, o.U_sp) < 1e-12, relerr_fields(ens.sims[i].dm.U_sp,
                                                                              o.U_sp)


def test_batched_fp32_members_match_solo_runs():
    """The batched kernels of the FP32-bulk form: members of three orders,
    7 steps, against the members run alone with the same substep form (bit
    for bit)."""
    from spd_amd.ensemble import Ensemble
    kws = members(6, 3, (8, 12, 16), (2, 3, 5), precision="fp32")
    print("[ensemble batched fp32]", flush=True)
    ens = Ensemble(kws, device="cuda:0", graph_steps=4, branches=3, batched=True)
    assert len(ens.groups) == 3 and not ens._solo
    ens.run(steps=7)
    torch.cuda.synchronize()
    for i, kw in enumerate(kws):
        ref = solo(kw, form=2, steps=7)
        s = ens.sims[i]
        # (a member that the FP32 rounding drives to non-finite values does so
        # in both runs, at the same step: NaN positions and time included)
        assert s.n_step == ref.n_step, i
        np.testing.assert_array_equal(np.float64(s.time), np.float64(ref.time))
        np.testing.assert_array_equal(s.dm.U_sp, ref.dm.U_sp)


def test_nonfinite_members_are_the_reference_schemes():
    """The small-member benchmark draw (64 members, N in {16, 32, 64}, p in
    {2..5}), 12 batched steps: the members with a non-finite state are
    exactly those whose reference run was non-finite by step 12."""
    import json
    import os
    from fixtures import GOLDEN
    from spd_amd.ensemble import Ensemble
    with open(os.path.join(GOLDEN, "ensemble_small_nonfinite.json")) as f:
        census = json.load(f)
    d = census["draw"]
    steps = 12
    kws = members(d["n"], d["seed"], tuple(d["sizes"]), tuple(d["orders"]))
    print(f"[ensemble census] {len(kws)} members, {steps} steps", flush=True)
    ens = Ensemble(kws, device="cuda:0", graph_steps=4, branches=8, tune=False)
    with np.errstate(all="ignore"):
        ens.run(steps=steps, finish=False)
    torch.cuda.synchronize()
    got = {i for i, s in enumerate(ens.sims)
           if not bool(torch.isfinite(s.dm.device_value("U_sp")).all())}
    want = {i for i, k in enumerate(census["first_nonfinite_step"])
            if k is not None and k <= steps}
    assert want, "the draw should hold members that blow up within the steps"
    assert got == want, (sorted(got), sorted(want))


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


@pytest.mark.parametrize("bc,seed", [("gradfree", 24), ("reflective", 23), ("periodic", 25)])
def test_batched_members_vs_oracle(bc, seed):
    """Members of fixture size (8 to 16 elements per side, p = 2..5, four
    quadrant states), with each kind of side, advanced by the batched
    launches (several members per order, so every batched kernel runs a grid
    of several members' blocks) in lock-step with the CPU oracle: before each
    of 3 steps every member takes the oracle's state (and its subcell-mean
    primitives, from which dt follows), then one batched step of all members
    is held to one oracle step at 1e-12 (norm-wise relative per field) for
    every member whose oracle state stays finite with positive density.
    (Run freely, the fully troubled members amplify the two implementations'
    round-off by ~100 per step on these states, as the reference's own
    scheme would.)  The draws' seeds avoid states on which one step already
    amplifies the round-off of identical inputs past the bar (some periodic
    four-quadrant members, troubled in every cell, reach 1e-11 in one step
    in solo runs as well as batched ones)."""
    from fixtures import relerr_fields
    from oracle.spd_oracle import OracleSimulator
    from spd_amd.ensemble import Ensemble, riemann_members
    specs = riemann_members(10, seed, (8, 12, 16), (2, 3, 4, 5))
    kws = []
    for m in members(10, seed, (8, 12, 16), (2, 3, 4, 5)):
        m["BC"] = ((bc, bc), (bc, bc))
        kws.append(m)
    print(f"[batched vs oracle {bc}] N {[k['N'] for k in kws]} p {[k['p'] for k in kws]}",
          flush=True)
    ens = Ensemble(kws, device="cuda:0", graph_steps=1, branches=4, batched=True)
    assert len(ens.groups) == len({k["p"] for k in kws}) and not ens._solo
    orc = []
    for kw in kws:
        kwo = dict(kw)
        init = kwo.pop("init_fct")
        for key in ("precision", "verbose"):
            kwo.pop(key, None)
        orc.append(OracleSimulator(init, **kwo))
    dev = torch.device("cuda:0")
    live = set(range(len(kws)))
    checked = 0
    for k in range(3):
        for j, (s, o) in enumerate(zip(ens.sims, orc)):
            s.init_sim()
            s.dm.device_value("U_sp").copy_(torch.as_tensor(o.U_sp, device=dev))
            s.dm.set_device_value("W_cv", torch.as_tensor(o.W_cv, device=dev).contiguous())
            s._dt_from_step = False
        with np.errstate(all="ignore"):
            ens.run(steps=1, finish=False)
            for o in orc:
                o.compute_dt()
                o.perform_update()
        torch.cuda.synchronize()
        for j in sorted(live):
            o = orc[j]
            if not np.all(np.isfinite(o.U_sp)) or float(np.min(o.U_sp[0])) <= 0.0:
                live.discard(j)
                continue
            got = ens.sims[j].dm.device_value("U_sp").cpu().numpy()
            err = relerr_fields(got, o.U_sp)
            assert err < 1e-12, (k, j, specs[j]["N"], specs[j]["p"], err)
            checked += 1
    assert checked >= 12, checked
