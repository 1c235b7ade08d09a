"""Reference-compatible simulator facade over the HIP step.

:class:`SDADER_Simulator` keeps the constructor keywords, the run methods
(``perform_update``, ``perform_iterations``, ``perform_time_evolution``,
``compute_dt``, ``init_sim``, ``end_sim``) and the snapshot attributes
(``dm.U_sp``, ``dm.U_cv``, ``dm.W_cv``, ``dm.dt``, ``W_init_cv``, ``time``,
``n_step``) of the reference's ``SDADER_Simulator``
(``src/sdader_simulator.py:14-413`` and its bases ``src/sd_simulator.py``,
``src/fv_simulator.py``, ``src/simulator.py``).  Every array operation of the
step runs in ``libspd_hip.so``; the host keeps only scalars (time, dt) and the
one-off initial-condition evaluation.

Array location follows the reference's data manager
(``src/data_management.py:32-68``): between ``init_sim`` and ``end_sim`` the
``dm`` arrays are device tensors (``torch.float64`` on the GPU); outside a run
they are host NumPy arrays, converted lazily on access.
"""
from __future__ import annotations

import ctypes
from timeit import default_timer as timer

import numpy as np
import torch

from . import _lib
from . import initial_conditions as _ic
from .mesh import ghost_layer_means, ic_mesh, solution_points

HOST, DEVICE = "host", "device"


class DataManager:
    """Attribute store whose arrays live on the host or on the GPU.

    Mirrors ``GPUDataManager`` (``src/data_management.py:32-68``): each
    attribute remembers where its current value lives and is converted on
    first access from the other location.  Scalars are stored as-is.
    """

    def __init__(self, device):
        object.__setattr__(self, "_vals", {})
        object.__setattr__(self, "_where", {})
        object.__setattr__(self, "_loc", HOST)
        object.__setattr__(self, "device", device)

    @property
    def current_location(self):
        return self._loc

    def _convert(self, v, loc):
        if loc == DEVICE and isinstance(v, np.ndarray):
            return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(self.device)
        if loc == HOST and isinstance(v, torch.Tensor):
            return v.detach().cpu().numpy()
        return v

    def __getattr__(self, key):
        vals = object.__getattribute__(self, "_vals")
        if key not in vals:
            raise AttributeError(key)
        where = object.__getattribute__(self, "_where")
        loc = object.__getattribute__(self, "_loc")
        if where[key] != loc:
            vals[key] = self._convert(vals[key], loc)
            where[key] = loc
        return vals[key]

    def __setattr__(self, key, value):
        self._vals[key] = self._convert(value, self._loc)
        self._where[key] = self._loc

    def set_device_value(self, key, tensor):
        """Store a device tensor as the current value regardless of location."""
        self._vals[key] = tensor
        self._where[key] = DEVICE

    def device_value(self, key):
        """Current value of ``key`` as a device tensor (converted if needed)."""
        if self._where[key] != DEVICE:
            self._vals[key] = self._convert(self._vals[key], DEVICE)
            self._where[key] = DEVICE
        return self._vals[key]

    def switch_to(self, loc):
        object.__setattr__(self, "_loc", loc)

    def asnumpy(self, arr):
        if isinstance(arr, torch.Tensor):
            return arr.detach().cpu().numpy()
        return np.asarray(arr)

    @property
    def xp(self):
        return np if self._loc == HOST else torch


def _bc_pair(side_pair):
    out = []
    for side in side_pair:
        if side not in _lib.BC:
            if side == "eq":
                raise NotImplementedError(f"boundary type {side!r} is not implemented")
            raise ValueError(f"Undetermined boundary type {side!r}")
        out.append(_lib.BC[side])
    return out


class SDADER_Simulator:
    """SD-ADER simulator with an optional a-posteriori MUSCL-Hancock fallback.

    Keywords and defaults are those of the reference
    (``src/sdader_simulator.py:15-27``, ``src/sd_simulator.py:20-26``,
    ``src/fv_simulator.py:11-17``, ``src/simulator.py:10-36``).  Options the
    reference offers beyond this path (``viscosity``, ``potential``, ``WB``,
    ``isothermal``, ``"eq"`` sides, ``predictor=False``, ``m != p``,
    ``Nghc != 2``) raise ``NotImplementedError``.  ``device`` selects the GPU (default: current).

    ``perform_time_evolution`` and ``perform_iterations`` run the time loop on
    the device when ``device_loop`` is true: dt, the t_end clamp, the dt < 1e-14
    stop, ``time`` and ``n_step`` are computed by kernels, and chunks of
    ``graph_steps`` steps are captured once in a HIP graph and replayed (0: no
    graph, the chunk's kernels are enqueued directly), with one host
    synchronisation per chunk.  The numbers are those of the host loop.

    ``precision="fp32"`` selects the FP32-bulk form: predictor, interpolation,
    Riemann solvers and MUSCL-Hancock in FP32; the state, the conservative
    update accumulation, the pressure recovery from total energy, trouble
    detection and dt stay FP64.  ``"fp64"`` (default) is the reference's
    arithmetic.
    """

    def __init__(self, init_fct=_ic.sine_wave, eq_fct=_ic.sine_wave, p=1, m=-1, N=(32, 32),
                 Nghe=1, Nghc=2, xlim=(0, 1), ylim=(0, 1), zlim=(0, 1), gamma=1.4,
                 beta=2.0 / 3, nu=1e-4, cfl_coeff=0.8, min_c2=1e-10, isothermal=False,
                 viscosity=False, potential=False, WB=False, use_cupy=True,
                 BC=(("periodic", "periodic"),) * 3, verbose=True,
                 riemann_solver_sd="llf", update="SD", riemann_solver_fv="llf",
                 slope_limiter="minmod", predictor=True, FB=False, tolerance=1e-5,
                 SED=True, NAD="", PAD=True, blending=True, min_rho=1e-10, max_rho=1e10,
                 min_P=1e-10, godunov=False, device=None, device_loop=True, graph_steps=16,
                 precision="fp64"):
        if m == -1:
            m = p
        for flag, name in ((isothermal, "isothermal"), (viscosity, "viscosity"),
                           (potential, "potential"), (WB, "WB")):
            if flag:
                raise NotImplementedError(f"{name}=True is not implemented")
        if not predictor:
            raise NotImplementedError("predictor=False is not implemented")
        if slope_limiter not in _lib.LIMITER:
            raise NotImplementedError(f"slope limiter {slope_limiter!r} is not implemented")
        for rs in (riemann_solver_sd, riemann_solver_fv):
            if rs not in _lib.RS:
                raise NotImplementedError(f"Riemann solver {rs!r} is not implemented")
        if update not in _lib.UPDATE:
            raise ValueError(f"update must be 'SD' or 'FV', got {update!r}")
        if NAD not in ("", "delta"):
            raise ValueError(f"NAD must be '' or 'delta', got {NAD!r}")
        if m != p:
            raise NotImplementedError("m != p is not implemented")
        if Nghc != 2:
            raise NotImplementedError("Nghc != 2 is not implemented")
        if precision not in ("fp64", "fp32"):
            raise ValueError(f"precision must be 'fp64' or 'fp32', got {precision!r}")
        self.precision = precision
        if not torch.cuda.is_available():
            raise RuntimeError("SDADER_Simulator needs a ROCm GPU (torch.cuda is unavailable)")

        self.init_fct = init_fct
        self.eq_fct = eq_fct
        self.p, self.m = p, m
        ndim = len(N)
        self.ndim = ndim
        assert len(BC) >= ndim
        self.BC, self.idims, self.dims = {}, {}, {}
        for idim, dim in enumerate("xyz"[:ndim]):
            self.idims[idim] = dim
            self.dims[dim] = idim
            self.BC[dim] = tuple(BC[idim])
            setattr(self, f"N{dim}", N[idim])
        self.Y, self.Z = ndim > 1, ndim > 2
        self.xlim, self.ylim, self.zlim = tuple(xlim), tuple(ylim), tuple(zlim)
        self.time = 0.0
        self.n_step = 0
        self.Nghe, self.Nghc = Nghe, Nghc
        self.gamma, self.beta, self.nu = gamma, beta, nu
        self.cfl_coeff, self.min_c2 = cfl_coeff, min_c2
        self.isothermal, self.viscosity, self.potential, self.WB = False, False, False, False
        self.verbose = verbose
        self.use_cupy = use_cupy
        self.update, self.FB = update, FB
        self.riemann_solver_sd_name, self.riemann_solver_fv_name = riemann_solver_sd, riemann_solver_fv
        self.slope_limiter_name, self.predictor = slope_limiter, predictor
        self.tolerance, self.SED, self.NAD, self.PAD = tolerance, SED, NAD, PAD
        self.blending, self.godunov = blending, godunov
        self.min_rho, self.max_rho, self.min_P = min_rho, max_rho, min_P
        self.rank = 0
        self.lim, self.len, self.h, self.N, self.n = {}, {}, {}, {}, {}
        self.h_min = 1e10
        for dim in self.dims:
            self.lim[dim] = tuple(getattr(self, f"{dim}lim"))
            self.len[dim] = self.lim[dim][1] - self.lim[dim][0]
            self.N[dim] = getattr(self, f"N{dim}")
            self.h[dim] = self.len[dim] / self.N[dim]
            setattr(self, f"d{dim}", self.h[dim])
            self.h_min = min(self.h_min, self.h[dim])
            self.n[dim] = p + 1
        self._d_ = 0
        self.vels = np.arange(ndim) + 1
        self._p_ = ndim + 1
        self.nvar = ndim + 2
        self.nader = m + 1
        self.variables = ["$\\rho$"] + [f"v{d}" for d in self.dims] + ["P"]

        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.dm = DataManager(self.device)
        self._lib = _lib.load()
        cfg = _lib.SpdConfig()
        cfg.ndim = ndim
        for a in range(3):
            cfg.N[a] = int(N[a]) if a < ndim else 1
            lim = (xlim, ylim, zlim)[a]
            cfg.lim[a][0], cfg.lim[a][1] = float(lim[0]), float(lim[1])
            if a < ndim:
                cfg.bc[a][0], cfg.bc[a][1] = _bc_pair(BC[a])
        cfg.p, cfg.m = p, m
        cfg.gamma, cfg.cfl_coeff, cfg.min_c2 = gamma, cfl_coeff, min_c2
        cfg.update = _lib.UPDATE[update]
        cfg.FB = int(bool(FB))
        cfg.riemann_sd, cfg.riemann_fv = _lib.RS[riemann_solver_sd], _lib.RS[riemann_solver_fv]
        cfg.tolerance = tolerance
        cfg.NAD = 1 if NAD == "delta" else 0
        cfg.SED, cfg.PAD, cfg.blending = int(bool(SED)), int(bool(PAD)), int(bool(blending))
        cfg.min_rho, cfg.max_rho, cfg.min_P = min_rho, max_rho, min_P
        cfg.godunov = int(bool(godunov))
        cfg.precision = _lib.PRECISION[precision]
        cfg.slope_limiter = _lib.LIMITER[slope_limiter]
        self._cfg = cfg
        ctx = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.spd_ctx_create(ctypes.byref(cfg), ctypes.byref(ctx)),
                       "spd_ctx_create")
        self._ctx = ctx
        self._dt_dev = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._dt_from_step = False   # the library holds the max wave speed of the state
        self._tap = None
        self.device_loop, self.graph_steps = bool(device_loop), int(graph_steps)
        self._graph = None           # (U_sp pointer, steps, torch.cuda.CUDAGraph)
        self._profiling = False

        # operator matrices (host copies, as built by the library)
        n, nf, T = p + 1, p + 2, m + 1
        buf = (ctypes.c_double * (nf * n * 2 + 3 * n * n + T * T + 2 * T))()
        cnt = _lib.check(self._lib.spd_get_operators(self._ctx, buf, len(buf)))
        ops = np.frombuffer(buf, dtype=np.float64, count=cnt).copy()
        o = 0

        def take(r, c):
            nonlocal o
            v = ops[o:o + r * c].reshape(r, c) if c else ops[o:o + r]
            o += r * c if c else r
            return v
        self.sp_to_fp = take(nf, n)
        self.dfp_to_sp = take(n, nf)
        self.sp_to_cv = take(n, n)
        self.cv_to_sp = take(n, n)
        self.invader = take(T, T)
        self.w_tp = take(T, 0)
        self.x_tp = take(T, 0)
        self.x_sp = solution_points(p)
        from .mesh import flux_points
        self.x_fp = flux_points(p)
        self.sp = {d: self.x_sp for d in self.dims}
        self.fp = {d: self.x_fp for d in self.dims}

        self.post_init()
        self._store_boundary_values()
        self.compute_dt()

    # -- lifetime ---------------------------------------------------------------
    def __del__(self):
        self._graph = None
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            try:
                torch.cuda.synchronize(self.device)
                self._lib.spd_ctx_destroy(ctx)
            except Exception:
                pass
            self._ctx = None

    @property
    def _stream(self):
        return _lib.stream_handle(self.device)

    def _shape_sd(self, nvar=None):
        nd, n = self.ndim, self.p + 1
        E = [self.N[d] for d in self.dims][::-1]
        return [self.nvar if nvar is None else nvar] + E + [n] * nd

    def _empty(self, nvar=None):
        return torch.empty(self._shape_sd(nvar), dtype=torch.float64, device=self.device)

    # -- initial condition (src/sd_simulator.py:59-90) -------------------------------
    def post_init(self):
        """Primitive subcell means of ``init_fct`` (2-point Gauss-Legendre rule
        per dimension) computed on the device: the quadrature points come from
        ``spd_ic_points``, the callable is evaluated on them as torch tensors
        (the callables of :mod:`initial_conditions` accept tensors; a callable
        written for NumPy is evaluated on a host copy of the points instead) and
        ``spd_ic_mean`` reduces the values per subcell."""
        dims = list(self.dims)
        lows, wids = ic_mesh([self.N[d] for d in dims], self.p, [self.lim[d] for d in dims],
                             [self.h[d] for d in dims], self.Nghe)
        dev = self.device
        lw = torch.from_numpy(np.concatenate([a.ravel() for a in lows])).to(dev)
        ww = torch.from_numpy(np.concatenate([a.ravel() for a in wids])).to(dev)
        nd = self.ndim
        full = [self.N[d] for d in dims][::-1] + [self.p + 1] * nd + [2] * nd
        pts = torch.empty([nd] + full, dtype=torch.float64, device=dev)
        _lib.check(self._lib.spd_ic_points(self._ctx, _lib.ptr(lw), _lib.ptr(ww), _lib.ptr(pts),
                                           self._stream), "spd_ic_points")
        W = self._empty()
        host_pts = None
        for var in range(self.nvar):
            vals = None
            if host_pts is None:
                try:
                    vals = self.init_fct(pts, var)
                except (TypeError, RuntimeError, AttributeError):
                    host_pts = pts.cpu().numpy()
            if host_pts is not None:
                vals = np.asarray(self.init_fct(host_pts, var), dtype=np.float64)
            vals = torch.as_tensor(vals, dtype=torch.float64, device=dev).expand(full).contiguous()
            _lib.check(self._lib.spd_ic_mean(self._ctx, _lib.ptr(vals), _lib.ptr(W[var]),
                                             self._stream), "spd_ic_mean")
        self._set_initial_primitives(W)

    def _store_boundary_values(self):
        """The values "ic" / "pressure" sides keep, from the initial condition
        as built at construction (init_sd_Boundaries / init_fv_Boundaries,
        src/sdader_simulator.py:50-55,415-423, src/fv_simulator.py:213-222):
        the library forms the element traces of the initial U_sp; the ghost
        layers of the initial primitives come from ``mesh.ghost_layer_means``.
        A later perturbation of the initial condition does not change them,
        as in the reference."""
        if not any(k in ("ic", "pressure") for d in self.dims for k in self.BC[d]):
            return
        dims = list(self.dims)
        Wg = ghost_layer_means(self.init_fct, [self.N[d] for d in dims], self.p,
                               [self.lim[d] for d in dims], [self.h[d] for d in dims], self.nvar,
                               self.Nghe, self.Nghc)
        Wd = torch.from_numpy(Wg).to(self.device)
        _lib.check(self._lib.spd_set_boundary_values(self._ctx,
                                                     _lib.ptr(self.dm.device_value("U_sp")),
                                                     _lib.ptr(Wd), self._stream),
                   "spd_set_boundary_values")
        torch.cuda.synchronize(self.device)

    def _set_initial_primitives(self, W_cv):
        if isinstance(W_cv, torch.Tensor):
            Wd = W_cv.to(self.device, dtype=torch.float64).contiguous()
            self.W_init_cv = Wd.cpu().numpy()
        else:
            self.W_init_cv = np.array(W_cv, dtype=np.float64)
            Wd = torch.from_numpy(self.W_init_cv).to(self.device)
        self.dm.set_device_value("W_cv", Wd)
        W_sp = self._tensor_op("spd_cv_to_sp", Wd)
        self.dm.set_device_value("W_sp", W_sp)
        self.dm.set_device_value("U_sp", self._pointwise("spd_prims_to_cons", W_sp))
        self.dm.set_device_value("U_cv", self._pointwise("spd_prims_to_cons", Wd))
        self._dt_from_step = False

    def perturb_initial_condition(self, dW_cv):
        """Add ``dW_cv`` (element layout) to the primitive subcell means of the
        initial condition and redo ``W_sp = cv_to_sp(W_cv)``, ``U_sp = cons(W_sp)``,
        ``U_cv = cons(W_cv)`` and ``compute_dt`` (how the seeded velocity noise of
        the Kelvin-Helmholtz benchmark enters; not a reference method)."""
        W = self.dm.asnumpy(self.dm.device_value("W_cv")) + np.asarray(dW_cv)
        self._set_initial_primitives(W)
        self.compute_dt()

    # -- library wrappers -----------------------------------------------------------
    def _tensor_op(self, name, src):
        """Element-layout transform of an array (..., E.., L..) with any number
        of leading rows; the kernels take V rows per call, so other counts are
        processed in zero-padded chunks of V rows."""
        src = self._as_device(src)
        per = 1
        for d in self.dims:
            per *= self.N[d] * (self.p + 1)
        if src.numel() % per != 0 or list(src.shape[-2 * self.ndim:]) != self._shape_sd()[1:]:
            raise ValueError(f"array of shape {tuple(src.shape)} is not in the element layout "
                             f"{tuple(self._shape_sd()[1:])}")
        rows = src.numel() // per
        flat = src.reshape(rows, per)
        V = self.nvar
        nchunk = (rows + V - 1) // V
        buf = torch.zeros(nchunk * V, per, dtype=torch.float64, device=self.device)
        buf[:rows] = flat
        out = torch.empty_like(buf)
        for k in range(nchunk):
            _lib.check(getattr(self._lib, name)(self._ctx, _lib.ptr(buf[k * V:(k + 1) * V]),
                                                _lib.ptr(out[k * V:(k + 1) * V]), self._stream),
                       name)
        return out[:rows].reshape(src.shape)

    def _pointwise(self, name, src):
        src = self._as_device(src)
        if src.dim() < 1 or src.shape[0] != self.nvar:
            raise ValueError(f"expected {self.nvar} variables along the first axis, "
                             f"got shape {tuple(src.shape)}")
        out = torch.empty_like(src)
        count = src.numel() // src.shape[0]
        _lib.check(getattr(self._lib, name)(self._ctx, _lib.ptr(src), _lib.ptr(out), count,
                                            self._stream), name)
        return out

    def _as_device(self, x):
        if isinstance(x, torch.Tensor):
            return x.to(self.device, torch.float64).contiguous()
        return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(self.device)

    @staticmethod
    def _like_input(x, out):
        return out if isinstance(x, torch.Tensor) else out.cpu().numpy()

    def compute_cv_from_sp(self, M_sp):
        """Subcell means from solution points (``src/sd_simulator.py:186-188``)."""
        return self._like_input(M_sp, self._tensor_op("spd_sp_to_cv", M_sp))

    def compute_sp_from_cv(self, M_cv):
        """Solution points from subcell means (``src/sd_simulator.py:183-185``)."""
        return self._like_input(M_cv, self._tensor_op("spd_cv_to_sp", M_cv))

    def compute_primitives(self, U, **kwargs):
        """Conservative -> primitive (``src/hydro.py:48-86``)."""
        return self._like_input(U, self._pointwise("spd_cons_to_prims", U))

    def compute_conservatives(self, W, **kwargs):
        """Primitive -> conservative (``src/hydro.py:88-125``)."""
        return self._like_input(W, self._pointwise("spd_prims_to_cons", W))

    # -- layouts (src/sd_simulator.py:92-131) ------------------------------------------
    def transpose_to_fv(self, M):
        nd = self.ndim
        lead = M.ndim - 2 * nd
        E = M.shape[lead:lead + nd]
        L = M.shape[lead + nd:]
        perm = list(range(lead))
        for k in range(nd):
            perm += [lead + k, lead + nd + k]
        shape = list(M.shape[:lead]) + [E[k] * L[k] for k in range(nd)]
        if isinstance(M, torch.Tensor):
            return M.permute(*perm).reshape(shape)
        return np.transpose(M, perm).reshape(shape)

    def transpose_to_sd(self, M):
        nd = self.ndim
        lead = M.ndim - nd
        shape = list(M.shape[:lead])
        for k in range(nd):
            shape += [self.N[self.idims[nd - 1 - k]], self.p + 1]
        perm = list(range(lead)) + [lead + 2 * k for k in range(nd)] + \
            [lead + 2 * k + 1 for k in range(nd)]
        if isinstance(M, torch.Tensor):
            return M.reshape(shape).permute(*perm)
        return np.transpose(M.reshape(shape), perm)

    def regular_mesh(self, W):
        """Interpolate element-layout values to a regular subcell mesh
        (``src/sd_simulator.py:92-102``)."""
        p = self.p
        if p <= 1:
            return W
        return self._like_input(W, self._tensor_op("spd_sp_to_regular", W))

    def crop(self, M, ngh=1):
        return M[(slice(None),) + (slice(ngh, -ngh),) * self.ndim + (Ellipsis,)]

    def domain_size(self):
        return [self.N[dim] * self.n[dim] for dim in self.dims]

    def regular_faces(self):
        return [np.linspace(self.lim[d][0], self.lim[d][1], self.N[d] * self.n[d] + 1)
                for d in self.dims]

    def regular_centers(self):
        return [np.linspace(self.lim[d][0], self.lim[d][1], self.N[d] * self.n[d])
                for d in self.dims]

    # -- time step (src/sd_simulator.py:214-227, src/sdader_simulator.py:361-413) ---------
    def compute_dt(self):
        """dt from the subcell-mean primitives ``dm.W_cv`` (device reduction)."""
        if self._dt_from_step:
            Wp = None
        else:
            Wd = self.dm.device_value("W_cv").contiguous()
            Wp = _lib.ptr(Wd)
        _lib.check(self._lib.spd_compute_dt(self._ctx, Wp, _lib.ptr(self._dt_dev), self._stream),
                   "spd_compute_dt")
        self.dm.dt = float(self._dt_dev.item())
        return self.dm.dt

    def _step(self, snapshot):
        self._dt_dev.fill_(float(self.dm.dt))
        U_sp = self.dm.device_value("U_sp")
        if snapshot and self.update == "FV":
            U_cv = self._empty()
            W_cv = self._empty()
        else:
            U_cv = W_cv = None
        _lib.check(self._lib.spd_step(self._ctx, _lib.ptr(U_sp), _lib.ptr(U_cv), _lib.ptr(W_cv),
                                      _lib.ptr(self._dt_dev), int(bool(snapshot)), self._stream),
                   "spd_step")
        if self.update == "FV":
            if snapshot:
                self.dm.set_device_value("U_cv", U_cv)
                self.dm.set_device_value("W_cv", W_cv)
            self._dt_from_step = True
        self.n_step += 1
        self.time += self.dm.dt

    def advance(self, nsteps=1):
        """``nsteps`` of (compute_dt, step) without snapshot writes and without a
        t_end clamp -- the inner loop of ``perform_iterations`` without
        ``init_sim``/``end_sim``.  ``dm.U_cv``/``dm.W_cv`` are left stale until
        ``end_sim`` (as in the reference, where they are recomputed there)."""
        for _ in range(nsteps):
            self.compute_dt()
            self._step(snapshot=False)

    def advance_device(self, nsteps=1):
        """``nsteps`` steps of the device-side loop (dt from the recorded wave
        speed, step, time += dt on the device; no t_end clamp), in chunks of
        ``graph_steps`` replayed from a HIP graph, one host synchronisation
        per chunk -- ``advance`` without a host round trip per step, and
        without ``init_sim``/``end_sim``."""
        if nsteps > 0:
            self._run_on_device(0.0, nsteps, 0)

    def perform_update(self) -> bool:
        """One step; ``dm.U_sp``, ``dm.U_cv``, ``dm.W_cv``, ``time`` and ``n_step``
        are updated as in ``src/sdader_simulator.py:361-378``."""
        self._step(snapshot=True)
        return True

    def init_sim(self):
        self.dm.switch_to(DEVICE)
        self.execution_time = -timer()

    def end_sim(self):
        torch.cuda.synchronize(self.device)
        self.execution_time += timer()
        U_cv = self._tensor_op("spd_sp_to_cv", self.dm.device_value("U_sp"))
        self.dm.set_device_value("U_cv", U_cv)
        self.dm.set_device_value("W_cv", self._pointwise("spd_cons_to_prims", U_cv))
        self._dt_from_step = False
        self.dm.switch_to(HOST)
        if self.rank == 0:
            print(f"t={self.time}, steps taken {self.n_step}, time taken {self.execution_time}")

    def prepare_graph(self, k=None):
        """Capture the device loop's chunk of ``k`` (default ``graph_steps``)
        steps in a HIP graph without running it, and make ``k`` the chunk
        size; the first chunk of a run otherwise captures its graph first."""
        k = self.graph_steps if k is None else int(k)
        if k <= 0:
            raise ValueError("the chunk of a graph holds at least one step")
        self.graph_steps = k
        U_sp = self.dm.device_value("U_sp")
        key = (U_sp.data_ptr(), k)
        if self._graph is None or self._graph[0] != key:
            self._graph = None
            g = torch.cuda.CUDAGraph()
            with torch.cuda.device(self.device), torch.cuda.graph(g):
                _lib.check(self._lib.spd_run_steps(self._ctx, _lib.ptr(U_sp), k,
                                                   self._stream), "spd_run_steps")
            self._graph = (key, g)
        return self._graph[1]

    def _run_chunk(self, U_sp, k):
        """Enqueue ``k`` device-loop steps: replay of a captured graph of
        ``graph_steps`` steps when ``k`` equals it, else direct launches."""
        s = self._stream
        if k == self.graph_steps and k > 0 and not self._profiling:
            self.prepare_graph(k).replay()
        else:
            _lib.check(self._lib.spd_run_steps(self._ctx, _lib.ptr(U_sp), k, s), "spd_run_steps")

    def _run_on_device(self, t_end, max_steps, clamp):
        """Device-side time loop (``src/sdader_simulator.py:394-413``); returns the
        final dt.  ``max_steps`` < 0 means no step budget."""
        self.compute_dt()                       # records the max wave speed of the state
        U_sp = self.dm.device_value("U_sp")
        _lib.check(self._lib.spd_run_begin(self._ctx, float(self.time), float(t_end),
                                           int(self.n_step), int(max_steps), int(clamp),
                                           self._stream), "spd_run_begin")
        chunk = self.graph_steps if self.graph_steps > 0 else 16
        t, n, done, dt = (ctypes.c_double(), ctypes.c_longlong(), ctypes.c_int(),
                          ctypes.c_double())
        left = max_steps
        while True:
            k = chunk if left < 0 else min(chunk, left)
            if k <= 0:
                break
            before = self.n_step
            self._run_chunk(U_sp, k)
            _lib.check(self._lib.spd_run_status(self._ctx, ctypes.byref(t), ctypes.byref(n),
                                                ctypes.byref(done), ctypes.byref(dt),
                                                self._stream), "spd_run_status")
            self.time, self.n_step, self.dm.dt = t.value, int(n.value), dt.value
            if left >= 0:
                left -= self.n_step - before
            if self.verbose and self.rank == 0 and self.n_step // 100 > before // 100:
                print(f"Time step #{self.n_step} (t = {self.time})", end="\r")
            if done.value:
                break
        if self.update == "FV":
            self._dt_from_step = True
        return self.dm.dt

    def perform_iterations(self, n_step: int) -> None:
        """``src/sdader_simulator.py:394-399``."""
        self.init_sim()
        if self.device_loop and n_step > 0:
            self._run_on_device(0.0, n_step, 0)
        else:
            for _ in range(n_step):
                self.compute_dt()
                self._step(snapshot=False)
        self.end_sim()

    def perform_time_evolution(self, t_end: float, nsteps=0) -> None:
        """``src/sdader_simulator.py:401-413``: dt clamped to reach ``t_end``; the
        loop stops if dt < 1e-14."""
        self.init_sim()
        if self.device_loop:
            if self.time < t_end:
                dt = self._run_on_device(t_end, -1, 1)
                if dt < 1e-14:
                    print(f"dt={dt}")
            self.end_sim()
            return
        while self.time < t_end:
            if not self.n_step % 100 and self.rank == 0 and self.verbose:
                print(f"Time step #{self.n_step} (t = {self.time})", end="\r")
            self.compute_dt()
            if self.time + self.dm.dt >= t_end:
                self.dm.dt = t_end - self.time
            if self.dm.dt < 1e-14:
                print(f"dt={self.dm.dt}")
                break
            self._step(snapshot=False)
        self.end_sim()

    # -- parity instrumentation -------------------------------------------------------
    def perform_update_with_flags(self):
        """``perform_update`` that also returns, for every FV substep, the
        troubled-cell flags, theta (ghosted grid) and face coefficients as NumPy
        arrays shaped like the reference's ``dm.troubles``, ``dm.theta`` and
        ``dm.affected_faces_{x,y,z}`` (a list of dicts, one per substep)."""
        nd, T = self.ndim, self.m + 1
        Ns = [self.N[d] * (self.p + 1) for d in self.dims]
        tr = torch.empty([T, 1] + Ns[::-1], dtype=torch.float64, device=self.device)
        th = torch.empty([T, 1] + [s + 4 for s in Ns[::-1]], dtype=torch.float64,
                         device=self.device)
        aff = []
        for a in range(nd):
            shp = [s + (1 if b == a else 0) for b, s in enumerate(Ns)][::-1]
            aff.append(torch.empty([T, 1] + shp, dtype=torch.float64, device=self.device))
        affp = [_lib.ptr(t) for t in aff] + [None] * (3 - nd)
        _lib.check(self._lib.spd_set_debug_tap(self._ctx, _lib.TAP_ALL, _lib.ptr(tr),
                                               _lib.ptr(th), *affp), "spd_set_debug_tap")
        try:
            self.perform_update()
        finally:
            self._lib.spd_set_debug_tap(self._ctx, -1, None, None, None, None, None)
        torch.cuda.synchronize(self.device)
        tr, th = tr.cpu().numpy(), th.cpu().numpy()
        affn = [t.cpu().numpy() for t in aff]
        out = []
        for i in range(T):
            rec = {"troubles": tr[i], "theta": th[i]}
            for a in range(nd):
                rec[f"aff_{'xyz'[a]}"] = affn[a][i]
            out.append(rec)
        return out

    def troubled_fraction_next_step(self):
        """Advance one step (compute_dt + step) and return the fraction of
        troubled subcells averaged over its FV substeps (computed on the GPU)."""
        if not (self.update == "FV" and self.FB):
            return 0.0
        T = self.m + 1
        Ns = [self.N[d] * (self.p + 1) for d in self.dims]
        tr = torch.empty([T] + Ns[::-1], dtype=torch.float64, device=self.device)
        _lib.check(self._lib.spd_set_debug_tap(self._ctx, _lib.TAP_ALL, _lib.ptr(tr), None,
                                               None, None, None), "spd_set_debug_tap")
        try:
            self.compute_dt()
            self._step(snapshot=False)
        finally:
            self._lib.spd_set_debug_tap(self._ctx, -1, None, None, None, None, None)
        return float(tr.mean().item())

    def profile_begin(self, max_launches=100000):
        """Start HIP-event timing of every kernel launch of the following steps."""
        _lib.check(self._lib.spd_profile_begin(self._ctx, int(max_launches)), "spd_profile_begin")
        self._profiling = True

    def profile_end(self):
        """Stop timing; return {kernel: (total_ms, launches)}."""
        k = self._lib.spd_kernel_class_count()
        ms = (ctypes.c_double * k)()
        cnt = (ctypes.c_longlong * k)()
        _lib.check(self._lib.spd_profile_end(self._ctx, ms, cnt, k), "spd_profile_end")
        self._profiling = False
        out = {}
        for i in range(k):
            if cnt[i]:
                out[self._lib.spd_kernel_class_name(i).decode()] = (float(ms[i]), int(cnt[i]))
        return out

    @property
    def scratch_bytes(self):
        return int(self._lib.spd_ctx_scratch_bytes(self._ctx))
