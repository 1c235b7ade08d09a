"""ctypes binding of ``libspd_hip.so`` (declarations in ``include/spd.h``).

The library is loaded after ``torch`` so that both share the process's single
HIP runtime (``libamdhip64.so.7``).  A missing library raises ``RuntimeError``:
there is no fallback implementation of the step.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime first)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspd_hip.so")

RS = {"llf": 0, "hllc": 1, "lhllc": 2}
LIMITER = {"minmod": 0, "moncen": 1}
BC = {"periodic": 0, "gradfree": 1, "reflective": 2, "ic": 3, "pressure": 4}
UPDATE = {"SD": 0, "FV": 1}
PRECISION = {"fp64": 0, "fp32": 1}
TAP_ALL = -2
LAUNCH = {"pred_epb": 0, "ho_epb": 1, "fb_form": 2, "hoc_epb": 3}

EXPORTS = [
    "spd_last_error", "spd_abi_version", "spd_config_size", "spd_ctx_create", "spd_ctx_destroy",
    "spd_ctx_scratch_bytes", "spd_get_operators", "spd_compute_dt", "spd_step",
    "spd_sp_to_cv", "spd_cv_to_sp", "spd_cons_to_prims", "spd_prims_to_cons",
    "spd_set_debug_tap", "spd_host_operators", "spd_host_geometry",
    "spd_profile_begin", "spd_profile_end", "spd_kernel_class_count", "spd_kernel_class_name",
    "spd_run_begin", "spd_run_steps", "spd_run_status", "spd_set_launch", "spd_get_launch",
    "spd_ic_points", "spd_ic_mean", "spd_sp_to_regular",
    "spd_adj_step", "spd_adj_backward", "spd_adj_theta_count", "spd_swap_wave_max",
    "spd_compute_dt_static", "spd_set_boundary_values",
    "spd_batch_create", "spd_batch_run_steps", "spd_batch_destroy",
]


class SpdConfig(ctypes.Structure):
    """Mirror of ``struct spd_config``."""
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("N", ctypes.c_int * 3),
        ("p", ctypes.c_int),
        ("m", ctypes.c_int),
        ("lim", (ctypes.c_double * 2) * 3),
        ("gamma", ctypes.c_double),
        ("cfl_coeff", ctypes.c_double),
        ("min_c2", ctypes.c_double),
        ("update", ctypes.c_int),
        ("FB", ctypes.c_int),
        ("riemann_sd", ctypes.c_int),
        ("riemann_fv", ctypes.c_int),
        ("bc", (ctypes.c_int * 2) * 3),
        ("tolerance", ctypes.c_double),
        ("NAD", ctypes.c_int),
        ("SED", ctypes.c_int),
        ("PAD", ctypes.c_int),
        ("blending", ctypes.c_int),
        ("min_rho", ctypes.c_double),
        ("max_rho", ctypes.c_double),
        ("min_P", ctypes.c_double),
        ("godunov", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("slope_limiter", ctypes.c_int),
    ]


_lib = None


def load():
    """Load the library (once) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m spd_amd.build` "
            "(the SD-ADER step has no implementation other than the HIP kernels)")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, i, ll = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
    L.spd_last_error.restype = ctypes.c_char_p
    L.spd_last_error.argtypes = []
    L.spd_abi_version.restype = i
    L.spd_config_size.restype = ctypes.c_size_t
    L.spd_config_size.argtypes = []
    if L.spd_config_size() != ctypes.sizeof(SpdConfig):
        raise RuntimeError("spd_config layout mismatch between spd.h and the ctypes mirror")
    L.spd_ctx_create.argtypes = [ctypes.POINTER(SpdConfig), ctypes.POINTER(vp)]
    L.spd_ctx_create.restype = i
    L.spd_ctx_destroy.argtypes = [vp]
    L.spd_ctx_destroy.restype = i
    L.spd_ctx_scratch_bytes.argtypes = [vp]
    L.spd_ctx_scratch_bytes.restype = ctypes.c_size_t
    L.spd_get_operators.argtypes = [vp, ctypes.POINTER(ctypes.c_double), i]
    L.spd_get_operators.restype = i
    L.spd_compute_dt.argtypes = [vp, dp, dp, vp]
    L.spd_compute_dt.restype = i
    L.spd_compute_dt_static.argtypes = [ctypes.POINTER(SpdConfig), dp, dp, vp]
    L.spd_compute_dt_static.restype = i
    L.spd_set_boundary_values.argtypes = [vp, dp, dp, vp]
    L.spd_set_boundary_values.restype = i
    L.spd_swap_wave_max.argtypes = [vp, dp, vp]
    L.spd_swap_wave_max.restype = i
    L.spd_step.argtypes = [vp, dp, dp, dp, dp, i, vp]
    L.spd_step.restype = i
    for name in ("spd_sp_to_cv", "spd_cv_to_sp"):
        getattr(L, name).argtypes = [vp, dp, dp, vp]
        getattr(L, name).restype = i
    for name in ("spd_cons_to_prims", "spd_prims_to_cons"):
        getattr(L, name).argtypes = [vp, dp, dp, ll, vp]
        getattr(L, name).restype = i
    L.spd_set_debug_tap.argtypes = [vp, i, dp, dp, dp, dp, dp]
    L.spd_set_debug_tap.restype = i
    L.spd_host_operators.argtypes = [i, i, ctypes.POINTER(ctypes.c_double), i]
    L.spd_host_operators.restype = i
    L.spd_host_geometry.argtypes = [i, ctypes.c_double, i, ctypes.POINTER(ctypes.c_double), i]
    L.spd_host_geometry.restype = i
    L.spd_profile_begin.argtypes = [vp, i]
    L.spd_profile_begin.restype = i
    L.spd_profile_end.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_longlong), i]
    L.spd_profile_end.restype = i
    L.spd_kernel_class_count.argtypes = []
    L.spd_kernel_class_count.restype = i
    L.spd_kernel_class_name.argtypes = [i]
    L.spd_kernel_class_name.restype = ctypes.c_char_p
    L.spd_run_begin.argtypes = [vp, ctypes.c_double, ctypes.c_double, ll, ll, i, vp]
    L.spd_run_begin.restype = i
    L.spd_run_steps.argtypes = [vp, dp, i, vp]
    L.spd_run_steps.restype = i
    L.spd_run_status.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ll),
                                 ctypes.POINTER(i), ctypes.POINTER(ctypes.c_double), vp]
    L.spd_run_status.restype = i
    L.spd_ic_points.argtypes = [vp, dp, dp, dp, vp]
    L.spd_ic_points.restype = i
    L.spd_ic_mean.argtypes = [vp, dp, dp, vp]
    L.spd_ic_mean.restype = i
    L.spd_sp_to_regular.argtypes = [vp, dp, dp, vp]
    L.spd_sp_to_regular.restype = i
    L.spd_set_launch.argtypes = [vp, i, i]
    L.spd_set_launch.restype = i
    L.spd_get_launch.argtypes = [vp, i, ctypes.POINTER(i), ctypes.POINTER(i)]
    L.spd_get_launch.restype = i
    L.spd_adj_step.argtypes = [vp, dp, dp, dp, dp, dp, vp]
    L.spd_adj_step.restype = i
    L.spd_adj_backward.argtypes = [vp, dp, dp, dp, vp]
    L.spd_adj_backward.restype = i
    L.spd_adj_theta_count.argtypes = [vp]
    L.spd_adj_theta_count.restype = ctypes.c_size_t
    L.spd_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(dp), i, ctypes.POINTER(vp)]
    L.spd_batch_create.restype = i
    L.spd_batch_run_steps.argtypes = [vp, i, vp]
    L.spd_batch_run_steps.restype = i
    L.spd_batch_destroy.argtypes = [vp]
    L.spd_batch_destroy.restype = i
    _lib = L
    return L


def host_operators(p, m=None):
    """Operator matrices built by the library's host code (no GPU needed):
    dict with sp_to_fp, dfp_to_sp, sp_to_cv, cv_to_sp, invader, w_tp, x_tp."""
    import numpy as np
    m = p if m is None else m
    n, nf, T = p + 1, p + 2, m + 1
    buf = (ctypes.c_double * (2 * nf * n + 2 * n * n + T * T + 2 * T))()
    cnt = check(load().spd_host_operators(p, m, buf, len(buf)), "spd_host_operators")
    v = np.frombuffer(buf, dtype=np.float64, count=cnt).copy()
    out, o = {}, 0
    for key, shape in (("sp_to_fp", (nf, n)), ("dfp_to_sp", (n, nf)), ("sp_to_cv", (n, n)),
                       ("cv_to_sp", (n, n)), ("invader", (T, T)), ("w_tp", (T,)),
                       ("x_tp", (T,))):
        size = int(np.prod(shape))
        out[key] = v[o:o + size].reshape(shape)
        o += size
    return out


def host_geometry(N, h, p):
    """Ghosted subcell geometry of one dimension (no GPU needed): dict with
    faces, centers, h_fp, h_cv."""
    import numpy as np
    Ns = N * (p + 1)
    total = (Ns + 5) + (Ns + 4) + (Ns + 4) + (Ns + 3)
    buf = (ctypes.c_double * total)()
    check(load().spd_host_geometry(N, float(h), p, buf, total), "spd_host_geometry")
    v = np.frombuffer(buf, dtype=np.float64).copy()
    out, o = {}, 0
    for key, size in (("faces", Ns + 5), ("centers", Ns + 4), ("h_fp", Ns + 4), ("h_cv", Ns + 3)):
        out[key] = v[o:o + size]
        o += size
    return out


def check(rc, what="spd call"):
    if rc < 0:
        msg = load().spd_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")
    return rc


def ptr(t):
    """Device pointer of a contiguous float64 torch tensor (or None)."""
    if t is None:
        return None
    assert t.is_contiguous() and t.dtype == torch.float64 and t.is_cuda, (t.dtype, t.device)
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
