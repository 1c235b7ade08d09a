"""Write the golden FP64 vectors of the SD-ADER step from the reference itself.

This script is the only place that runs the reference (pure Python + NumPy,
read-only under /root/reference/src).  It is run by hand in the development
container, never by the tests, and writes small ``.npz`` fixtures next to
itself.  Each fixture records, for one configuration:

* ``cfg``           -- the constructor keywords (JSON) and the IC description;
* ``W_init_cv``, ``U_sp0``, ``U_cv0``, ``dt0``  -- the IC pipeline output
  (``post_init`` + ``compute_dt``, ``src/sd_simulator.py:74-90,214-227``);
* ``U_sp_k``, ``U_cv_k``, ``W_cv_k``, ``dt_k``, ``time_k`` after step ``k``
  of a continuous run (``compute_dt`` + ``perform_update``,
  ``src/sdader_simulator.py:361-378``);
* for the first steps, per FV substep ``i``: the troubled-cell flags,
  ``theta``, the affected-face coefficients, the high-order and fallback face
  fluxes, the candidate state and the primitive state at the start of the
  substep (captured right before ``correct_fluxes``,
  ``src/sdader_simulator.py:325-349``).

The functions these vectors pin (``detect_troubles``, ``MUSCL_Hancock_fluxes``,
``fv_store_BC`` at non-periodic sides, the composite ``perform_update`` and the
IC pipeline) are the ones no closed form or invariant fixes.

Usage:  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_SRC = "/root/reference/src"

import numpy as np  # noqa: E402

sys.path.insert(0, REPO)
from spd_amd import initial_conditions as ic  # noqa: E402

sys.path.insert(0, REF_SRC)
from sdader_simulator import SDADER_Simulator  # noqa: E402


class Capturing(SDADER_Simulator):
    """Reference simulator that records the FV substep internals."""

    def correct_fluxes(self):
        if getattr(self, "capture", False):
            # Flags and theta for every captured substep; the face fluxes and
            # the candidate state only for the first substep of the first
            # captured step (and never in 3-D), to keep the fixtures small.
            full = self.capture_full and not self.captures and self.ndim < 3
            rec = {
                "troubles": np.array(self.dm.troubles, dtype=np.float64),
                "theta": np.array(self.dm.theta),
            }
            for dim in self.dims:
                rec[f"aff_{dim}"] = np.array(getattr(self.dm, f"affected_faces_{dim}"))
            if full:
                rec["Ucand"] = np.array(self.dm.U_new)
                rec["Wold"] = np.array(self.dm.W_cv)
                for dim in self.dims:
                    rec[f"FHO_{dim}"] = np.array(self.F_faces[dim])
                    rec[f"FFB_{dim}"] = np.array(self.F_faces_FB[dim])
            self.captures.append(rec)
        super().correct_fluxes()


def apply_noise(s, dW):
    """Add subcell-mean noise to the primitive IC and redo the IC tail
    (``src/sd_simulator.py:84-87``) and ``compute_dt``."""
    s.dm.W_cv = s.dm.W_cv + dW
    s.W_init_cv = s.dm.W_cv.copy()
    s.dm.W_sp = s.compute_sp_from_cv(s.dm.W_cv)
    s.dm.U_sp = s.compute_conservatives(s.dm.W_sp)
    s.dm.U_cv = s.compute_conservatives(s.dm.W_cv)
    s.compute_dt()


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


def run_case(name, kw, desc, steps, capture_steps=(), record_steps=None,
             t_end=None, noise=None):
    kw = dict(kw)
    kw["use_cupy"] = False
    kw["verbose"] = False
    s = Capturing(init_fct=make_ic(desc), **kw)
    if noise is not None:
        dW = ic.subcell_noise(kw["N"], kw["p"], noise["vars"], noise["amp"],
                              seed=noise["seed"])
        apply_noise(s, dW)
    out = {
        "cfg": np.array(json.dumps({"kw": {k: v for k, v in kw.items()},
                                    "ic": desc, "noise": noise,
                                    "t_end": t_end})),
        "W_init_cv": np.array(s.W_init_cv),
        "U_sp0": np.array(s.dm.U_sp),
        "dt0": np.array(float(s.dm.dt)),
    }
    if record_steps is None:
        record_steps = range(1, steps + 1)
    record_steps = set(record_steps)
    s.init_sim()
    s.captures = []
    capture_steps = tuple(capture_steps)
    for k in range(1, steps + 1):
        s.capture = k in capture_steps
        s.capture_full = bool(capture_steps) and k == capture_steps[0]
        s.captures = []
        s.compute_dt()
        s.perform_update()
        out[f"dt_{k}"] = np.array(float(s.dm.dt))
        if k in record_steps:
            out[f"U_sp_{k}"] = np.array(s.dm.U_sp)
            if k == max(record_steps):
                out[f"U_cv_{k}"] = np.array(s.dm.U_cv)
                out[f"W_cv_{k}"] = np.array(s.dm.W_cv)
            out[f"time_{k}"] = np.array(float(s.time))
        for i, rec in enumerate(s.captures):
            for key, val in rec.items():
                out[f"cap{k}_{i}_{key}"] = val
    s.capture = False
    out["nsteps"] = np.array(steps)
    if t_end is not None:
        s2 = Capturing(init_fct=make_ic(desc), **kw)
        if noise is not None:
            apply_noise(s2, dW)
        s2.perform_time_evolution(t_end)
        out["final_U_sp"] = np.array(s2.dm.U_sp)
        out["final_U_cv"] = np.array(s2.dm.U_cv)
        out["final_W_cv"] = np.array(s2.dm.W_cv)
        out["final_time"] = np.array(float(s2.time))
        out["final_n_step"] = np.array(int(s2.n_step))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print(f"{name}: {os.path.getsize(path)/1024:.0f} KiB", flush=True)


def _census_member(args):
    """First step after which the member's U_sp holds a non-finite value
    (fixed steps of compute_dt + perform_update, no t_end clamp), or None."""
    spec, steps = args
    kw = ensemble_member_kw(spec)
    kw.update(use_cupy=False, verbose=False)
    s = SDADER_Simulator(init_fct=ic.four_quadrant(np.asarray(spec["states"])), **kw)
    s.init_sim()
    with np.errstate(all="ignore"):
        for k in range(1, steps + 1):
            s.compute_dt()
            s.perform_update()
            if not np.all(np.isfinite(s.dm.U_sp)):
                return k
    return None


def ensemble_member_kw(spec):
    """Keywords of one member of the four-quadrant ensemble (outflow sides,
    FV + fallback, HLLC, minmod, CFL 0.8)."""
    N = spec["N"]
    return dict(p=spec["p"], N=(N, N), BC=GF2, update="FV", FB=True,
                riemann_solver_sd="hllc", riemann_solver_fv="hllc", slope_limiter="minmod",
                cfl_coeff=0.8, gamma=1.4)


def census(n=64, seed=0, sizes=(16, 32, 64), steps=50, procs=6):
    """Run every member of the seeded small-member ensemble draw through the
    reference for ``steps`` steps and write, per member, the first step whose
    state is non-finite (the reference scheme's own blow-ups on these random
    four-quadrant states) to ``ensemble_small_nonfinite.json``."""
    from multiprocessing import Pool
    from spd_amd.ensemble import riemann_members
    specs = riemann_members(n, seed=seed, sizes=sizes)
    with Pool(procs) as pool:
        first = pool.map(_census_member, [(sp, steps) for sp in specs], chunksize=1)
    out = {"draw": {"n": n, "seed": seed, "sizes": list(sizes), "orders": [2, 3, 4, 5]},
           "steps": steps, "members": [{"N": sp["N"], "p": sp["p"]} for sp in specs],
           "first_nonfinite_step": first}
    path = os.path.join(HERE, "ensemble_small_nonfinite.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"census: {sum(x is not None for x in first)} of {n} members non-finite "
          f"within {steps} steps -> {path}", flush=True)
    return first


HLLC = dict(riemann_solver_sd="hllc", riemann_solver_fv="hllc")
FVFB = dict(update="FV", FB=True, slope_limiter="minmod", cfl_coeff=0.8)
GF1 = (("gradfree", "gradfree"),)
GF2 = (("gradfree", "gradfree"), ("gradfree", "gradfree"))
PER2 = (("periodic", "periodic"), ("periodic", "periodic"))


def main():
    only = set(sys.argv[1:])

    def want(n):
        return not only or n in only

    # Members of the small-member ensemble draw that the reference scheme
    # drives to non-finite values: the census of all 64, and two members
    # recorded step by step up to and including their first non-finite step.
    if "census" in only:
        census()
    for idx, name in ((30, "ens30_p4_blowup"), (45, "ens45_p5_blowup")):
        if name in only:
            from spd_amd.ensemble import riemann_members
            spec = riemann_members(64, seed=0, sizes=(16, 32, 64))[idx]
            k = _census_member((spec, 50))
            assert k is not None, (idx, "stays finite")
            with np.errstate(all="ignore"):
                run_case(name, ensemble_member_kw(spec),
                         {"kind": "quad", "states": np.asarray(spec["states"]).tolist()},
                         steps=k)
    if only and only <= {"census", "ens30_p4_blowup", "ens45_p5_blowup"}:
        return

    # Sod shock tube: the 1-D CPU configuration (outflow = "gradfree").
    if want("sod_p3"):
        run_case("sod_p3", dict(p=3, N=(100,), BC=GF1, **HLLC, **FVFB),
                 {"kind": "sod"}, steps=50, capture_steps=(1, 2),
                 record_steps=(1, 2, 3, 4, 5, 20, 50), t_end=0.2)
    # p=1 (no SED), LLF faces everywhere, relative NAD.
    if want("sod_p1_llf"):
        run_case("sod_p1_llf", dict(p=1, N=(50,), BC=GF1,
                                    riemann_solver_sd="llf", riemann_solver_fv="llf",
                                    **FVFB),
                 {"kind": "sod"}, steps=6, capture_steps=(1, 2, 3))
    # p=2, "delta" NAD, no blending, reflective walls.
    if want("sod_p2_delta"):
        run_case("sod_p2_delta", dict(p=2, N=(40,), BC=(("reflective", "reflective"),),
                                      NAD="delta", blending=False, **HLLC, **FVFB),
                 {"kind": "sod"}, steps=6, capture_steps=(1, 2))
    # Kelvin-Helmholtz (benchmark set-up) at 16^2 and 32^2, periodic.
    kh_noise = {"vars": [1, 2], "amp": 1e-3, "seed": 0}
    if want("kh16_p3"):
        run_case("kh16_p3", dict(p=3, N=(16, 16), BC=PER2, **HLLC, **FVFB),
                 {"kind": "kh"}, steps=3, capture_steps=(1,), noise=kh_noise)
    # Four-quadrant Riemann problems, outflow, p = 2..5.
    rng = np.random.default_rng(1234)
    for p in (2, 3, 4, 5):
        states = ic.random_four_quadrant_states(rng)
        name = f"quad12_p{p}"
        if want(name):
            run_case(name, dict(p=p, N=(12, 12), BC=GF2, **HLLC, **FVFB),
                     {"kind": "quad", "states": states.tolist()}, steps=2,
                     capture_steps=(1,))
    # Isentropic vortex, p=5, periodic on [-5,5]^2.
    if want("vortex12_p5"):
        run_case("vortex12_p5", dict(p=5, N=(12, 12), BC=PER2, xlim=(-5, 5),
                                     ylim=(-5, 5), **HLLC, **FVFB),
                 {"kind": "vortex"}, steps=2, capture_steps=(1,))
    # Mixed boundaries: periodic x, reflective y, LLF faces, p=2.
    if want("step12_p2_refl"):
        run_case("step12_p2_refl",
                 dict(p=2, N=(12, 10), BC=(("periodic", "periodic"),
                                           ("reflective", "reflective")),
                      riemann_solver_sd="llf", riemann_solver_fv="llf", **FVFB),
                 {"kind": "step", "kw": {"vx": 1.0, "vy": 0.4, "P": 1.0}},
                 steps=3, capture_steps=(1, 2, 3))
    # Mixed boundaries: gradfree x, periodic y, HLLC, p=3.
    if want("quad10_p3_mixed"):
        states = ic.random_four_quadrant_states(np.random.default_rng(77))
        run_case("quad10_p3_mixed",
                 dict(p=3, N=(10, 8), BC=(("gradfree", "gradfree"),
                                          ("periodic", "periodic")), **HLLC, **FVFB),
                 {"kind": "quad", "states": states.tolist()}, steps=2,
                 capture_steps=(1, 2))
    # The reference's own invariance set-up (test/test_preserve.py): sine wave,
    # N=6, run to t=1 in SD, FV, FV+FB and FV+FB+godunov modes.
    for ndim in (1, 2):
        for p in (1, 3):
            for mode in ("SD", "FV", "FB", "FBG"):
                name = f"sine{ndim}d_p{p}_{mode}"
                if not want(name) and not want("sine"):
                    continue
                kw = dict(p=p, N=(6,) * ndim)
                if mode == "SD":
                    kw["update"] = "SD"
                elif mode == "FV":
                    kw["update"] = "FV"
                else:
                    kw.update(update="FV", FB=True, godunov=(mode == "FBG"))
                run_case(name, kw, {"kind": "sine"}, steps=3,
                         capture_steps=(1,) if mode in ("FB", "FBG") else (),
                         t_end=1.0)
    # Low-Mach HLLC (lhllc) and the moncen limiter: 1-D Sod to t_end, 2-D
    # four-quadrant outflow, 2-D Kelvin-Helmholtz (periodic), 3-D step.
    LM = dict(riemann_solver_sd="lhllc", riemann_solver_fv="lhllc")
    MC = dict(FVFB, slope_limiter="moncen")
    if want("sod_p3_lhllc_moncen"):
        run_case("sod_p3_lhllc_moncen", dict(p=3, N=(60,), BC=GF1, **LM, **MC),
                 {"kind": "sod"}, steps=20, capture_steps=(1, 2),
                 record_steps=(1, 2, 3, 10, 20), t_end=0.2)
    if want("quad12_p3_lhllc_moncen"):
        states = ic.random_four_quadrant_states(np.random.default_rng(4321))
        run_case("quad12_p3_lhllc_moncen", dict(p=3, N=(12, 12), BC=GF2, **LM, **MC),
                 {"kind": "quad", "states": states.tolist()}, steps=2, capture_steps=(1,))
    if want("kh16_p3_moncen"):
        run_case("kh16_p3_moncen", dict(p=3, N=(16, 16), BC=PER2, riemann_solver_sd="hllc",
                                        riemann_solver_fv="lhllc", **MC),
                 {"kind": "kh"}, steps=3, capture_steps=(1,), noise=kh_noise)
    if want("step3d_p2_moncen"):
        run_case("step3d_p2_moncen", dict(p=2, N=(5, 4, 4),
                                          BC=GF2 + (("gradfree", "gradfree"),), **LM, **MC),
                 {"kind": "step", "kw": {"vx": 0.5, "vy": -0.3, "vz": 0.2}},
                 steps=1, capture_steps=(1,))
    # Sides whose boundary values are stored at construction: "ic" keeps the
    # initial trace / ghost layers outside, "pressure" also overwrites the
    # boundary element's own trace with it (src/sd_boundary.py:50-56,
    # src/fv_simulator.py:213-253).
    if want("sine1d_p3_ic"):
        run_case("sine1d_p3_ic", dict(p=3, N=(12,), BC=(("ic", "ic"),), **HLLC, **FVFB),
                 {"kind": "sine"}, steps=5, capture_steps=(1, 2), t_end=0.5)
    if want("sod_p2_pressure_ic"):
        run_case("sod_p2_pressure_ic", dict(p=2, N=(30,), BC=(("pressure", "ic"),), **HLLC,
                                            **FVFB),
                 {"kind": "sod"}, steps=5, capture_steps=(1, 2), t_end=0.4)
    if want("quad10_p3_ic_pressure"):
        states = ic.random_four_quadrant_states(np.random.default_rng(99))
        run_case("quad10_p3_ic_pressure",
                 dict(p=3, N=(10, 8), BC=(("ic", "pressure"), ("gradfree", "ic")), **HLLC, **FVFB),
                 {"kind": "quad", "states": states.tolist()}, steps=3, capture_steps=(1, 2, 3))
    if want("step12_p2_pressure"):
        run_case("step12_p2_pressure",
                 dict(p=2, N=(12, 10), BC=(("pressure", "ic"), ("pressure", "reflective")),
                      riemann_solver_sd="llf", riemann_solver_fv="llf", **FVFB),
                 {"kind": "step", "kw": {"vx": 1.0, "vy": 0.4, "P": 1.0}},
                 steps=3, capture_steps=(1, 2, 3))
    if want("step3d_p2_ic"):
        run_case("step3d_p2_ic", dict(p=2, N=(5, 4, 4),
                                      BC=(("ic", "gradfree"), ("gradfree", "pressure"),
                                          ("ic", "ic")), **HLLC, **FVFB),
                 {"kind": "step", "kw": {"vx": 0.5, "vy": -0.3, "vz": 0.2}},
                 steps=2, capture_steps=(1, 2))
    # 3-D Taylor-Green, 4^3, p=3, periodic on [0, 2 pi]^3.
    if want("tg4_p3"):
        twopi = 2 * np.pi
        run_case("tg4_p3", dict(p=3, N=(4, 4, 4), BC=PER2 + (("periodic", "periodic"),),
                                xlim=(0, twopi), ylim=(0, twopi), zlim=(0, twopi),
                                **HLLC, **FVFB),
                 {"kind": "tg"}, steps=1, capture_steps=(1,),
                 noise={"vars": [0], "amp": 1e-2, "seed": 0})
    # 3-D step function with outflow sides, p=2, exercises the 3-D blending.
    if want("step3d_p2"):
        run_case("step3d_p2", dict(p=2, N=(5, 4, 4),
                                   BC=GF2 + (("gradfree", "gradfree"),), **HLLC, **FVFB),
                 {"kind": "step", "kw": {"vx": 0.5, "vy": -0.3, "vz": 0.2}},
                 steps=1, capture_steps=(1,))


if __name__ == "__main__":
    main()
