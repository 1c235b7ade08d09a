"""Shared helpers of the test-suite: loading the golden fixtures and building
the oracle / simulator configurations they describe."""
from __future__ import annotations

import glob
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
if REPO not in sys.path:
    sys.path.insert(0, REPO)

from spd_amd import initial_conditions as ic  # noqa: E402


def golden_names(blowup=False):
    """Names of the golden fixtures.  The fixtures of runs that the reference
    scheme drives to non-finite values (names ending in ``_blowup``) are
    listed only with ``blowup=True`` (and then only they): their last step
    is compared for non-finiteness, not to a tolerance."""
    names = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
    return [n for n in names if n.endswith("_blowup") == blowup]


def load(name):
    fx = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    fx["cfg"] = json.loads(str(fx["cfg"]))
    return fx


def make_ic(desc):
    kind = desc["kind"]
    if kind == "sine":
        return lambda xyz, case: ic.sine_wave(xyz, case, **desc.get("kw", {}))
    if kind == "step":
        return lambda xyz, case: ic.step_function(xyz, case, **desc.get("kw", {}))
    if kind == "sod":
        return ic.sod_shock_tube
    if kind == "kh":
        return ic.kelvin_helmholtz
    if kind == "vortex":
        return lambda xyz, case: ic.isentropic_vortex(xyz, case, **desc.get("kw", {}))
    if kind == "quad":
        return ic.four_quadrant(np.array(desc["states"]))
    if kind == "tg":
        return ic.taylor_green
    raise ValueError(kind)


def sim_kwargs(cfg):
    """Constructor keywords of a fixture (tuples restored, host-only keys dropped)."""
    kw = dict(cfg["kw"])
    kw.pop("use_cupy", None)
    kw.pop("verbose", None)
    kw["N"] = tuple(kw["N"])
    if "BC" in kw:
        kw["BC"] = tuple(tuple(b) for b in kw["BC"])
    for k in ("xlim", "ylim", "zlim"):
        if k in kw:
            kw[k] = tuple(kw[k])
    return kw


def noise_of(cfg):
    nz = cfg.get("noise")
    if not nz:
        return None
    kw = cfg["kw"]
    return ic.subcell_noise(tuple(kw["N"]), kw["p"], nz["vars"], nz["amp"], seed=nz["seed"])


def relerr(a, b):
    """Norm-wise relative difference max|a-b| / max|b| (per array)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    scale = np.max(np.abs(b)) if b.size else 0.0
    if scale == 0.0:
        scale = 1.0
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


def relerr_fields(a, b):
    """Largest norm-wise relative difference over the fields of a state
    (leading axis = variable: density, ndim momenta/velocities, energy/pressure).

    Density and energy are scaled by their own maximum; the momentum (or
    velocity) components share one scale, the largest magnitude over all of
    them, so that a component that is identically zero in the initial
    condition is measured against the flow's velocity scale, not its own
    round-off.
    """
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    nv = a.shape[0]
    vel_scale = max(float(np.max(np.abs(b[1:nv - 1]))) if nv > 2 else 0.0, 1e-300)
    worst = 0.0
    for v in range(nv):
        if 1 <= v <= nv - 2:
            e = float(np.max(np.abs(a[v] - b[v]))) / vel_scale if a[v].size else 0.0
        else:
            e = relerr(a[v], b[v])
        worst = max(worst, e)
    return worst


def flag_mismatch(got, ref, margin, tie=1e-10):
    """Number of flag disagreements that are NOT within ``tie`` (relative) of a
    detection threshold according to the oracle's ``margin`` array."""
    got = np.asarray(got).reshape(np.shape(ref))
    margin = np.asarray(margin).reshape(np.shape(ref))
    diff = got != ref
    return int(np.sum(diff & (margin >= tie))), int(np.sum(diff))
