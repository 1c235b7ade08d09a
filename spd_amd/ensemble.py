"""Ensembles of uneven SD-ADER problems advanced together on one GPU.

Each member is a :class:`SDADER_Simulator` with its own library context and
its own device-side time loop (dt, t_end clamp, dt < 1e-14 stop and step
count on the device).  A chunk of steps of *every* member is captured into
one HIP graph: the members are spread over a few capture branches (streams
forked from the capture stream and joined back), balanced by the estimated
work of each member (degrees of freedom x remaining steps), so one graph
launch advances the whole ensemble by ``graph_steps`` steps with no host
synchronisation inside.  Members that reach their end keep replaying as no-ops
(every step kernel tests the member's device skip flag first); the graph is
rebuilt without them once the work left in the graph has shrunk.

Launch shapes the kernels take at run time (elements per block of the
predictor update, of the 2-D high-order flux kernel and of the 2-D fused
high-order flux + candidate kernel) are chosen per member configuration by
:class:`Autotuner`, which times the candidate shapes on the member itself
(state saved and restored around the trial steps) and caches the winner per
(ndim, N, p, precision); batched launches take each member's own shape.

The members' numbers are those of the members run alone: the same kernels run
on the same data in the same order within each member.
"""
from __future__ import annotations

import ctypes
import time as _time

import numpy as np
import torch

from . import _lib
from . import initial_conditions as ic
from .simulator import SDADER_Simulator


def riemann_members(n=64, seed=0, sizes=(64, 128, 256, 384), orders=(2, 3, 4, 5)):
    """Member specifications of the 2-D four-quadrant Riemann ensemble: N and
    p drawn uniformly from ``sizes`` and ``orders`` and the quadrant states from
    :func:`initial_conditions.random_four_quadrant_states`, all from one seeded
    generator.  Returns a list of dicts (``N``, ``p``, ``states``)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(rng.choice(sizes))
        p = int(rng.choice(orders))
        out.append({"N": N, "p": p, "states": ic.random_four_quadrant_states(rng)})
    return out


def riemann_member_kwargs(spec, precision="fp64"):
    """Simulator keywords of one ensemble member: [0,1]^2, outflow sides,
    FV corrector with the fallback, HLLC faces, minmod, CFL 0.8, gamma 1.4."""
    N = spec["N"]
    return dict(init_fct=ic.four_quadrant(spec["states"]), p=spec["p"], N=(N, N),
                BC=(("gradfree", "gradfree"), ("gradfree", "gradfree")), update="FV", FB=True,
                riemann_solver_sd="hllc", riemann_solver_fv="hllc", slope_limiter="minmod",
                cfl_coeff=0.8, gamma=1.4, precision=precision, verbose=False)


def _dof(sim):
    d = 1
    for dim in sim.dims:
        d *= sim.N[dim] * (sim.p + 1)
    return d


class _Saved:
    """Device copy of a member's state and host scalars (for trial steps)."""

    def __init__(self, sim):
        self.U = sim.dm.device_value("U_sp").clone()
        self.scalars = (sim.time, sim.n_step, sim.dm.dt, sim._dt_from_step)

    def restore(self, sim):
        sim.dm.device_value("U_sp").copy_(self.U)
        sim.time, sim.n_step, sim.dm.dt, sim._dt_from_step = self.scalars


# Launch keys the autotuner sets: the ones that only regroup independent work
# into blocks (bit-identical results).  The form of the 2-D fallback kernel
# ("fb_form") changes the evaluation order of a few expressions and stays fixed.
TUNED = ("pred_epb", "ho_epb", "hoc_epb")


class Autotuner:
    """Per-configuration choice of the run-time launch shapes.

    For each key the candidates are the maximum and its halvings down to 1;
    the keys are tuned one after the other (coordinate descent), each candidate
    timed over ``steps`` steps after one warm-up step with HIP events on the
    member's stream.  Results are cached per (ndim, N, p, precision)."""

    def __init__(self, steps=2, max_candidates=5):
        self.steps = steps
        self.max_candidates = max_candidates
        self.cache = {}

    @staticmethod
    def key(sim):
        return (sim.ndim, tuple(sim.N[d] for d in sim.dims), sim.p, sim.precision)

    def _time_steps(self, sim, n):
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record()
        sim.advance(n)
        stop.record()
        stop.synchronize()
        return start.elapsed_time(stop) / n

    def tune(self, sim):
        k = self.key(sim)
        if k in self.cache:
            self.apply(sim, self.cache[k])
            return self.cache[k]
        lib, ctx = sim._lib, sim._ctx
        saved = _Saved(sim)
        sim.dm.switch_to(_lib_device())
        choice = {}
        for name, key in _lib.LAUNCH.items():
            if name not in TUNED or (name in ("ho_epb", "hoc_epb") and sim.ndim != 2):
                continue
            cur, mx = ctypes.c_int(), ctypes.c_int()
            _lib.check(lib.spd_get_launch(ctx, key, ctypes.byref(cur), ctypes.byref(mx)))
            cands = []
            v = mx.value
            while v >= 1 and len(cands) < self.max_candidates:
                cands.append(v)
                v //= 2
            if cur.value not in cands:
                cands.append(cur.value)
            best, best_t = cur.value, None
            for c in cands:
                _lib.check(lib.spd_set_launch(ctx, key, c), "spd_set_launch")
                sim.advance(1)
                t = self._time_steps(sim, self.steps)
                saved.restore(sim)
                if best_t is None or t < best_t:
                    best, best_t = c, t
            _lib.check(lib.spd_set_launch(ctx, key, best), "spd_set_launch")
            choice[name] = best
        saved.restore(sim)
        self.cache[k] = choice
        return choice

    @staticmethod
    def apply(sim, choice):
        for name, v in choice.items():
            _lib.check(sim._lib.spd_set_launch(sim._ctx, _lib.LAUNCH[name], int(v)),
                       "spd_set_launch")


def _lib_device():
    from .simulator import DEVICE
    return DEVICE


def _batch_key(sim, max_dof):
    """Members with equal keys can share batched launches (spd_batch_create):
    2-D, FV with the fallback, and the same order, precision, Riemann solvers
    and kinds of sides; None for a member that cannot be batched or whose
    degrees of freedom reach ``max_dof`` (such a member's own launches
    already fill the GPU)."""
    if sim.ndim != 2 or sim.update != "FV" or not sim.FB or _dof(sim) >= max_dof:
        return None
    sides = [k for d in sim.dims for k in sim.BC[d]]
    stored = any(k in ("ic", "pressure") for k in sides)
    gate = any(k != "periodic" for k in sides)
    return (sim.p, sim.precision, sim.riemann_solver_sd_name, sim.riemann_solver_fv_name,
            stored, gate)


class Ensemble:
    """A set of simulators advanced together (see the module docstring).

    ``members``: list of keyword dicts for :class:`SDADER_Simulator` (each with
    its ``init_fct``).  ``graph_steps``: steps of every member per graph launch.
    ``branches``: capture branches (streams) the launch groups are spread over.
    ``batched``: members that can share launches (2-D, FV with the fallback,
    same order, precision, Riemann solvers and kinds of sides, fewer than
    ``batch_max_dof`` degrees of freedom) are advanced by one launch per stage
    of the step for the whole group (spd_batch_run_steps; the substep form is
    set to 2 on those members, whose results then equal their own runs with
    that form bit for bit); the others, or all members without ``batched``,
    by their own launches.  ``tune``: run the autotuner on every member
    configuration that keeps its own launches (batched launches honour each
    member's shapes, set with spd_set_launch)."""

    def __init__(self, members, device=None, graph_steps=8, branches=4, tune=True,
                 tuner=None, batched=True, batch_max_dof=1 << 20):
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.sims = [SDADER_Simulator(device=self.device, **kw) for kw in members]
        self.graph_steps = int(graph_steps)
        self.branches = max(1, int(branches))
        self.batched = bool(batched)
        keys = [_batch_key(s, batch_max_dof) if self.batched else None for s in self.sims]
        groups = {}
        for i, k in enumerate(keys):
            if k is not None:
                groups.setdefault(k, []).append(i)
        self.groups = list(groups.values())
        self._solo = [i for i, k in enumerate(keys) if k is None]
        for g in self.groups:
            for i in g:
                s = self.sims[i]
                _lib.check(s._lib.spd_set_launch(s._ctx, _lib.LAUNCH["fb_form"], 2),
                           "spd_set_launch")
        self.tuner = tuner if tuner is not None else (Autotuner() if tune else None)
        self.tuning = {}
        if self.tuner is not None:
            # members in batched groups keep the default shapes: a shape timed
            # on a member alone (more, smaller blocks for a small grid) is not
            # the best inside a batched launch whose members fill the GPU
            # together (small-member ensemble: 4.9x with tuned members, 5.5x
            # without)
            for i in self._solo:
                s = self.sims[i]
                self.tuning[Autotuner.key(s)] = self.tuner.tune(s)
        self._streams = [torch.cuda.Stream(device=self.device) for _ in range(self.branches)]
        self._graph = None
        self._batches = {}   # member indices -> (state pointers, spd_batch handle)
        self.graph_builds = 0
        self.build_seconds = 0.0
        self.replays = 0

    def __del__(self):
        try:
            for _, h in self._batches.values():
                _lib.load().spd_batch_destroy(h)
        except Exception:
            pass

    def _batch(self, members):
        """The spd_batch of a group, made again when a member's state tensor
        has moved."""
        key = tuple(members)
        U = [self.sims[i].dm.device_value("U_sp") for i in members]
        ptrs = tuple(u.data_ptr() for u in U)
        old = self._batches.get(key)
        if old is not None and old[0] == ptrs:
            return old[1]
        lib = _lib.load()
        if old is not None:
            lib.spd_batch_destroy(old[1])
        n = len(members)
        ctxs = (ctypes.c_void_p * n)(*[self.sims[i]._ctx for i in members])
        states = (ctypes.c_void_p * n)(*ptrs)
        h = ctypes.c_void_p()
        _lib.check(lib.spd_batch_create(ctxs, states, n, ctypes.byref(h)), "spd_batch_create")
        self._batches[key] = (ptrs, h)
        return h

    # -- helpers ------------------------------------------------------------------
    @property
    def dof(self):
        return [_dof(s) for s in self.sims]

    def _status(self, s):
        t, n, done, dt = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_int(), ctypes.c_double()
        _lib.check(s._lib.spd_run_status(s._ctx, ctypes.byref(t), ctypes.byref(n),
                                         ctypes.byref(done), ctypes.byref(dt),
                                         _lib.stream_handle(self.device)), "spd_run_status")
        return t.value, int(n.value), bool(done.value), dt.value

    def _assign(self, active, cost):
        """Longest-processing-time-first assignment of members to branches."""
        loads = [0.0] * self.branches
        groups = [[] for _ in range(self.branches)]
        for i in sorted(active, key=lambda i: -cost[i]):
            b = int(np.argmin(loads))
            groups[b].append(i)
            loads[b] += cost[i]
        return groups

    def _units(self, active):
        """Launch units of the active members: each batched group whose
        members are all active as one unit (a finished member's blocks leave
        at once), each other member alone."""
        act = set(active)
        units = []
        for g in self.groups:
            if any(i in act for i in g):
                units.append(("batch", tuple(g)))
        grouped = {i for g in self.groups for i in g}
        units += [("solo", (i,)) for i in active if i not in grouped]
        return units

    def _graph_key(self, active):
        units = self._units(active)
        return (tuple(units), self.graph_steps, self.branches,
                tuple(self.sims[i].dm.device_value("U_sp").data_ptr()
                      for _, ms in units for i in ms))

    def _build_graph(self, active, cost):
        t0 = _time.perf_counter()
        key = self._graph_key(active)
        self._graph = None
        units = self._units(active)
        handles = {u: self._batch(u[1]) for u in units if u[0] == "batch"}
        ucost = [sum(cost[i] for i in ms) for _, ms in units]
        branches = self._assign(range(len(units)), ucost)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.device(self.device), torch.cuda.graph(g):
            cap = torch.cuda.current_stream()
            for b, idx in enumerate(branches):
                if not idx:
                    continue
                st = self._streams[b]
                st.wait_stream(cap)
                with torch.cuda.stream(st):
                    for j in idx:
                        kind, ms = units[j]
                        sh = ctypes.c_void_p(st.cuda_stream)
                        if kind == "batch":
                            _lib.check(_lib.load().spd_batch_run_steps(
                                handles[units[j]], self.graph_steps, sh), "spd_batch_run_steps")
                            continue
                        s = self.sims[ms[0]]
                        U = s.dm.device_value("U_sp")
                        _lib.check(s._lib.spd_run_steps(s._ctx, _lib.ptr(U), self.graph_steps, sh),
                                   "spd_run_steps")
            for b, idx in enumerate(branches):
                if idx:
                    cap.wait_stream(self._streams[b])
        self.graph_builds += 1
        self._graph = (key, g)
        self.build_seconds += _time.perf_counter() - t0

    # -- runs -----------------------------------------------------------------------
    def run(self, t_end=None, steps=None, rebuild_fraction=0.5, finish=True):
        """Advance every member to ``t_end`` (with the reference's clamp and
        stop rule) or by ``steps`` steps, batched.  The graph is rebuilt
        without finished members once their share of the graph's estimated
        work exceeds ``rebuild_fraction``; a graph whose members, buffers and
        shape are unchanged is replayed again by the next call.  With
        ``finish`` the members end as ``perform_time_evolution`` ends them
        (snapshot arrays, host location); without, they stay on the device."""
        if (t_end is None) == (steps is None):
            raise ValueError("give exactly one of t_end and steps")
        sims = self.sims
        stream = _lib.stream_handle(self.device)
        remaining = []
        for s in sims:
            s.init_sim()
            s.compute_dt()
            if t_end is not None:
                _lib.check(s._lib.spd_run_begin(s._ctx, float(s.time), float(t_end),
                                                int(s.n_step), -1, 1, stream), "spd_run_begin")
                remaining.append(max(1.0, (t_end - s.time) / max(s.dm.dt, 1e-300)))
            else:
                _lib.check(s._lib.spd_run_begin(s._ctx, 0.0, 0.0, int(s.n_step), int(steps), 0,
                                                stream), "spd_run_begin")
                remaining.append(float(steps))
        torch.cuda.synchronize(self.device)
        dof = self.dof
        cost = [dof[i] * remaining[i] for i in range(len(sims))]
        active = [i for i in range(len(sims)) if remaining[i] > 0]
        done = [False] * len(sims)
        if t_end is not None:
            for i, s in enumerate(sims):
                if not s.time < t_end:
                    done[i] = True
        active = [i for i in active if not done[i]]
        graph_cost = sum(cost[i] for i in active)
        while active:
            if self._graph is None or self._graph[0] != self._graph_key(active):
                graph_cost = sum(cost[i] for i in active)
                self._build_graph(active, cost)
            self._graph[1].replay()
            self.replays += 1
            torch.cuda.synchronize(self.device)
            for i in active:
                t, n, d, dt = self._status(sims[i])
                s = sims[i]
                s.time, s.n_step, s.dm.dt = t, n, dt
                done[i] = d
            still = [i for i in active if not done[i]]
            finished_cost = sum(cost[i] for i in active if done[i])
            if still and finished_cost > rebuild_fraction * graph_cost:
                active = still                       # rebuild without the finished ones
            elif not still:
                active = []
        for s in sims:
            if s.update == "FV":
                s._dt_from_step = True
            if finish:
                s.end_sim()
        return self

    def run_sequential(self, t_end=None, steps=None):
        """The members one after the other with the host-driven loop of the
        reference (src/sdader_simulator.py:394-413): the plain baseline."""
        if (t_end is None) == (steps is None):
            raise ValueError("give exactly one of t_end and steps")
        for s in self.sims:
            loop = s.device_loop
            s.device_loop = False
            try:
                if t_end is not None:
                    s.perform_time_evolution(t_end)
                else:
                    s.perform_iterations(steps)
            finally:
                s.device_loop = loop
        return self


def timed(fn):
    """Wall time of ``fn()`` between two device synchronisations."""
    torch.cuda.synchronize()
    t0 = _time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return _time.perf_counter() - t0
