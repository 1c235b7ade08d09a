"""Members of the seeded four-quadrant ensemble that the reference scheme
itself drives to non-finite values.

``tests/golden/make_golden.py`` ran the reference on every member of the
small-member draw (``riemann_members(64, seed=0, sizes=(16, 32, 64))``: N in
{16, 32, 64}, p in {2..5}, outflow sides, FV + fallback, HLLC, minmod, CFL
0.8) for 50 steps of compute_dt + perform_update and recorded each member's
first non-finite step (``ensemble_small_nonfinite.json``); for members #30
(N = 16, p = 4) and #45 (N = 16, p = 5) it recorded every step up to and
including that step (``ens30_p4_blowup``, ``ens45_p5_blowup``).

CPU: the oracle follows the recorded steps and goes non-finite at the same
step.  GPU: the HIP step follows the oracle and the recorded states step by
step, and is non-finite at the same step as both.  The bar is 1e-12 (states,
norm-wise relative per field, and dt) while the recorded state is physical
(positive density).  A blow-up passes through finite non-physical states
(member #45 after step 6: negative density, |E| ~ 1e74) in which the scheme
amplifies round-off by orders of magnitude per step (the oracle's 6e-15
difference from the reference after step 5 becomes 1.7e-9 after step 6); such
states, and the dt computed from them, are held to 1e-6.
"""
import json
import os

import numpy as np
import pytest

from fixtures import GOLDEN, golden_names, load, make_ic, relerr_fields, sim_kwargs

TOL = 1e-12
TOL_NONPHYSICAL = 1e-6
NAMES = golden_names(blowup=True)
CENSUS = os.path.join(GOLDEN, "ensemble_small_nonfinite.json")


def _physical(U):
    return bool(np.all(np.isfinite(U)) and np.min(U[0]) > 0.0)


def _oracle(fx):
    from oracle.spd_oracle import OracleSimulator
    cfg = fx["cfg"]
    return OracleSimulator(make_ic(cfg["ic"]), **sim_kwargs(cfg))


def test_census_matches_draw():
    """The census covers the draw the ensemble benchmark and tests use."""
    from spd_amd.ensemble import riemann_members
    with open(CENSUS) as f:
        c = json.load(f)
    d = c["draw"]
    specs = riemann_members(d["n"], seed=d["seed"], sizes=tuple(d["sizes"]))
    assert [{"N": s["N"], "p": s["p"]} for s in specs] == c["members"]
    first = c["first_nonfinite_step"]
    assert len(first) == d["n"]
    assert first[30] == int(load("ens30_p4_blowup")["nsteps"])
    assert first[45] == int(load("ens45_p5_blowup")["nsteps"])


@pytest.mark.parametrize("name", NAMES)
def test_oracle_follows_blowup(name):
    fx = load(name)
    o = _oracle(fx)
    nsteps = int(fx["nsteps"])
    assert relerr_fields(o.U_sp, fx["U_sp0"]) < TOL
    tol = TOL
    with np.errstate(all="ignore"):
        for k in range(1, nsteps + 1):
            o.compute_dt()
            assert abs(o.dt - float(fx[f"dt_{k}"])) <= tol * float(fx[f"dt_{k}"]), k
            o.perform_update()
            ref = fx[f"U_sp_{k}"]
            tol = TOL if _physical(ref) else TOL_NONPHYSICAL
            if k < nsteps:
                assert np.all(np.isfinite(ref)) and np.all(np.isfinite(o.U_sp)), k
                assert relerr_fields(o.U_sp, ref) < tol, (k, relerr_fields(o.U_sp, ref))
    assert not np.all(np.isfinite(fx[f"U_sp_{nsteps}"]))
    assert not np.all(np.isfinite(o.U_sp))


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_follows_blowup(name):
    from spd_amd.simulator import SDADER_Simulator
    print(f"[gpu blowup] {name}", flush=True)
    fx = load(name)
    cfg = fx["cfg"]
    s = SDADER_Simulator(init_fct=make_ic(cfg["ic"]), device="cuda:0", verbose=False,
                         **sim_kwargs(cfg))
    o = _oracle(fx)
    nsteps = int(fx["nsteps"])
    tol = TOL
    with np.errstate(all="ignore"):
        for k in range(1, nsteps + 1):
            s.compute_dt()
            o.compute_dt()
            assert abs(s.dm.dt - float(fx[f"dt_{k}"])) <= tol * float(fx[f"dt_{k}"]), k
            s.perform_update()
            o.perform_update()
            U = s.dm.U_sp
            ref = fx[f"U_sp_{k}"]
            tol = TOL if _physical(ref) else TOL_NONPHYSICAL
            if k < nsteps:
                assert np.all(np.isfinite(U)), k
                assert relerr_fields(U, o.U_sp) < tol, ("oracle", k, relerr_fields(U, o.U_sp))
                assert relerr_fields(U, ref) < tol, ("golden", k, relerr_fields(U, ref))
    assert not np.all(np.isfinite(s.dm.U_sp))
