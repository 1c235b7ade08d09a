"""Reverse-mode derivative (adjoint) of SD-ADER steps behind ``torch.autograd``.

The step is differentiated with its trouble flags, blending coefficients and
dt frozen: they are taken from the forward pass and held fixed (the flags and
the CFL dt are piecewise-constant / non-smooth functions of the state and have
no useful derivative).  With them fixed, one step of
``SDADER_Simulator.perform_update`` (``src/sdader_simulator.py:361-378``) is a
smooth map of ``U_sp`` almost everywhere, and the library's hand-written
reverse sweep (``spd_adj_step`` / ``spd_adj_backward``, ``include/spd.h``)
returns its vector-Jacobian product on the device.

``Adjoint.run`` advances ``nsteps`` steps as a ``torch.autograd.Function``:
the forward pass keeps the state every ``checkpoint_every`` steps and the dt
of every step; the backward pass re-runs each interval from its checkpoint
(the step is deterministic, so the recomputed flags are the forward pass's),
then walks the interval backwards, recording one step's tape (Picard
iterates, FV substep states, blending coefficients) and running its reverse
sweep.  Memory: ``nsteps / k + k`` states plus one step's tape; with
``offload`` the states live in host memory and two at a time on the device.

Both forms are differentiated; the reverse sweep is FP64 in both.  In the
FP32-bulk form the tape holds the form's FP32 predictor iterates, so the sweep
linearises the FP64 formulas at the FP32 form's states.
"""
from __future__ import annotations

import contextlib
import functools
import sys

import torch

from . import _lib


def _own_wave_max(fn):
    """Run ``fn`` with the adjoint's own recorded maximum wave speed swapped
    into the shared context (and the simulator's swapped back afterwards), so
    that the steps the adjoint runs never change the dt the simulator's next
    ``compute_dt`` takes from the context.  Re-entrant: only the outermost
    call swaps."""
    @functools.wraps(fn)
    def wrapper(self, *args, **kwargs):
        with self._wave_max():
            return fn(self, *args, **kwargs)
    return wrapper


class Adjoint:
    """Adjoint of the steps of one ``SDADER_Simulator``: it runs on the
    simulator's context (options, mesh, operator matrices and scratch; the
    tape and the reverse sweep's buffers are added to it on first use).  The
    simulator's state is not touched: neither its arrays nor the maximum wave
    speed its context records for the simulator's next dt."""

    def __init__(self, sim, offload=False, verbose=False):
        self.sim = sim
        self.verbose = bool(verbose)  # one line per checkpoint interval (stderr)
        # keep checkpoints and the recomputed states of an interval in host
        # memory (the differentiation configuration's 128^3 grid: 30 states of
        # 5.4 GB beside a 204 GiB context)
        self.offload = bool(offload)
        self.lib = _lib.load()
        self.device = sim.device
        self.fb = sim.update == "FV" and bool(sim.FB)
        self.n_theta = int(self.lib.spd_adj_theta_count(sim._ctx))
        self._cmax = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._depth = 0

    @contextlib.contextmanager
    def _wave_max(self):
        if self._depth == 0:
            _lib.check(self.lib.spd_swap_wave_max(self.sim._ctx, _lib.ptr(self._cmax),
                                                  self._stream), "spd_swap_wave_max")
        self._depth += 1
        try:
            yield
        finally:
            self._depth -= 1
            if self._depth == 0:
                _lib.check(self.lib.spd_swap_wave_max(self.sim._ctx, _lib.ptr(self._cmax),
                                                      self._stream), "spd_swap_wave_max")

    # -- one step ------------------------------------------------------------------
    @property
    def _stream(self):
        return _lib.stream_handle(self.device)

    def _check(self, U):
        shape = tuple(self.sim._shape_sd())
        if tuple(U.shape) != shape or U.dtype != torch.float64 or not U.is_cuda:
            raise ValueError(f"expected a float64 device tensor of shape {shape}, "
                             f"got {tuple(U.shape)} {U.dtype} {U.device}")
        return U.contiguous()

    @_own_wave_max
    def dt_of(self, U_sp, out=None):
        """CFL dt of a state (``src/sd_simulator.py:214-227``): from the
        subcell-mean primitives of ``U_sp``.  Returns a 1-element device tensor."""
        U_sp = self._check(U_sp)
        W = self.sim._pointwise("spd_cons_to_prims", self.sim._tensor_op("spd_sp_to_cv", U_sp))
        out = torch.empty(1, dtype=torch.float64, device=self.device) if out is None else out
        _lib.check(self.lib.spd_compute_dt(self.sim._ctx, _lib.ptr(W.contiguous()), _lib.ptr(out),
                                           self._stream), "spd_compute_dt")
        return out

    @_own_wave_max
    def step(self, U_sp, dt, theta=None, record_theta=False):
        """One step out of place, recording the tape.  ``dt``: 1-element device
        tensor.  ``theta`` (optional): blending coefficients of every substep
        (as returned with ``record_theta``) to use instead of the detection.
        Returns ``U_out`` or ``(U_out, theta)``."""
        U_sp = self._check(U_sp)
        out = torch.empty_like(U_sp)
        th_out = None
        if record_theta:
            if not self.fb:
                raise ValueError("blending coefficients exist only with update='FV', FB=True")
            th_out = torch.empty(self.n_theta, dtype=torch.float64, device=self.device)
        th_in = None if theta is None else theta.contiguous()
        if th_in is not None and th_in.numel() != self.n_theta:
            raise ValueError(f"theta must hold {self.n_theta} values")
        _lib.check(self.lib.spd_adj_step(self.sim._ctx, _lib.ptr(U_sp), _lib.ptr(out),
                                         _lib.ptr(dt), _lib.ptr(th_in), _lib.ptr(th_out),
                                         self._stream), "spd_adj_step")
        return (out, th_out) if record_theta else out

    @_own_wave_max
    def step_vjp(self, U_sp, dt, U_bar, theta=None):
        """Vector-Jacobian product of one step at ``U_sp`` (flags frozen as
        detected at ``U_sp``, or ``theta``): returns ``U_bar^T d step / d U_sp``."""
        U_sp = self._check(U_sp)
        self.step(U_sp, dt, theta=theta)
        Ub = self._check(U_bar).clone()
        _lib.check(self.lib.spd_adj_backward(self.sim._ctx, _lib.ptr(U_sp), _lib.ptr(dt),
                                             _lib.ptr(Ub), self._stream), "spd_adj_backward")
        return Ub

    # -- many steps ----------------------------------------------------------------
    def _advance(self, U, dt):
        _lib.check(self.lib.spd_step(self.sim._ctx, _lib.ptr(U), None, None, _lib.ptr(dt), 0,
                                     self._stream), "spd_step")

    def _keep(self, U, slot=None):
        """A copy of the state U for later: on the device, or with ``offload``
        in host memory (``slot``, a host tensor of the same shape, is reused
        when given).  Host memory that the GPU box lets a process pin is a few
        tens of GB, less than the 30 states of 5.4 GB the differentiation
        configuration keeps, so the copies go through pageable memory."""
        if not self.offload:
            return U.clone()
        if slot is None:
            slot = torch.empty(U.shape, dtype=U.dtype)
        slot.copy_(U)
        return slot

    def _fetch(self, V, out=None):
        """The kept state V on the device (``out`` reused when given)."""
        if not self.offload:
            return V if out is None else out.copy_(V)
        if out is None:
            return V.to(self.device)
        return out.copy_(V)

    @_own_wave_max
    def _forward(self, U0, nsteps, k, dt):
        """Run nsteps from U0; returns (U_N, checkpoints, dts)."""
        U = self._check(U0).clone()
        dts = torch.empty(max(nsteps, 1), dtype=torch.float64, device=self.device)
        ckpt = {}
        for s in range(nsteps):
            if s % k == 0:
                ckpt[s] = self._keep(U)
                if self.verbose and s:
                    print(f"adjoint: forward step {s} of {nsteps}", file=sys.stderr, flush=True)
            d = dts[s:s + 1]
            if dt is not None:
                d.fill_(float(dt))
            elif s == 0 or self.sim.update == "SD":
                # the SD update never refreshes W_cv, so its dt stays the first
                # one (src/sdader_simulator.py:376)
                if s == 0:
                    self.dt_of(U, out=d)
                else:
                    d.copy_(dts[0:1])
            else:
                _lib.check(self.lib.spd_compute_dt(self.sim._ctx, None, _lib.ptr(d),
                                                   self._stream), "spd_compute_dt")
            self._advance(U, d)
        return U, ckpt, dts

    @_own_wave_max
    def _backward(self, U_bar, nsteps, k, ckpt, dts):
        Ub = self._check(U_bar).clone()
        # the reverse sweep's buffers are allocated on first use: hand torch's
        # cached blocks back to the device first
        torch.cuda.empty_cache()
        starts = sorted(ckpt)
        slots = []   # host copies of an interval's states (offload), reused
        Ud = None    # device copy of a kept state (offload)
        for j0 in reversed(starts):
            j1 = min(j0 + k, nsteps)
            states = [ckpt[j0]]
            U = self._fetch(ckpt[j0]).clone() if not self.offload else ckpt[j0].to(self.device)
            for s in range(j0, j1 - 1):
                self._advance(U, dts[s:s + 1])
                i = s - j0
                states.append(self._keep(U, slots[i] if i < len(slots) else None))
                if self.offload and i >= len(slots):
                    slots.append(states[-1])
            del U
            for s in range(j1 - 1, j0 - 1, -1):
                if self.offload:
                    Ud = self._fetch(states[s - j0], Ud)
                    Ub = self.step_vjp(Ud, dts[s:s + 1], Ub)
                else:
                    Ub = self.step_vjp(states[s - j0], dts[s:s + 1], Ub)
            if self.verbose:
                torch.cuda.synchronize(self.device)
                print(f"adjoint: steps {j0}..{j1 - 1} reversed", file=sys.stderr, flush=True)
        return Ub

    def run(self, U0, nsteps, checkpoint_every=20, dt=None):
        """``nsteps`` steps from ``U0`` (element layout, FP64, on the device),
        differentiable with respect to ``U0`` through ``torch.autograd``.
        ``dt``: a fixed step, or None for the CFL dt of each state (frozen in
        the derivative).  Returns the final state."""
        if nsteps < 0 or checkpoint_every < 1:
            raise ValueError("nsteps >= 0 and checkpoint_every >= 1 required")
        return _Run.apply(U0, self, int(nsteps), int(checkpoint_every), dt)


class _Run(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, adj, nsteps, k, dt):
        with torch.no_grad():
            U, ckpt, dts = adj._forward(U0.detach(), nsteps, k, dt)
        ctx.adj, ctx.nsteps, ctx.k = adj, nsteps, k
        ctx.ckpt, ctx.dts = ckpt, dts
        return U

    @staticmethod
    def backward(ctx, U_bar):
        with torch.no_grad():
            g = ctx.adj._backward(U_bar.contiguous(), ctx.nsteps, ctx.k, ctx.ckpt, ctx.dts)
        ctx.ckpt = None
        return g, None, None, None, None
