"""Benchmark of the SD-ADER step on one MI355X.

Workload (the metric's configuration): 2-D Kelvin-Helmholtz on [0,1]^2,
periodic, 1024^2 elements, p = 3 (16.8 M solution points), update="FV" with
the a-posteriori fallback (NAD + SED + PAD, blending), HLLC on element and
fallback faces, minmod, CFL 0.8; rho = 2, vx = 0.5 for |y-0.5| < 0.25 else
rho = 1, vx = -0.5, P = 2.5, vy = 0.01 sin(4 pi x), plus seeded uniform noise
(amplitude 1e-3, numpy default_rng(0)) on the vx, vy subcell means.

A step is one pass of the hot path over the whole state: dt from the
maximum wave speed of the state and one spd_step (predictor + FV/FB
corrector).  The state is resident in HBM before the timed region (10
warm-up steps, then 200 timed steps from t = 0); the timed steps run the
device-side time loop of perform_time_evolution (dt, step and time on the
device) replayed from one HIP graph, without instrumentation; the same steps
with the host-driven loop (dt read back every step, as in the reference's
loop) are reported as host_loop_ms_per_step and must end in the same state
bit for bit; the per-kernel durations come from 20 further steps with HIP
events around every launch.  The JSON line carries:
  * value: DOF-updates/s of the FP64 form (DOF = solution points; one
    DOF-update = one point advanced one full step), with the roofline of the
    dominant kernel (HIP-event durations on the kernel's own stream; HBM
    traffic per launch from the committed counter summary), the step's HBM
    traffic and FP64 rate from the committed counter summaries
    (step_traffic, compute_roofline) and per-kernel times;
  * t1_window_recorded: the same workload timed from t = 1 (200 steps after
    the ~26k that reach it), as recorded by --t1-window in the committed
    summary (--t1-window measures it in the run, several minutes more);
  * fp32_bulk: the same workload in the FP32-bulk form, its speedup and its
    relative L1 difference from the FP64 form after the same steps;
  * ensemble / ensemble_small_members: 64 four-quadrant Riemann problems
    (N in {64,128,256,384} resp. {16,32,64}, p in {2..5}, outflow), a fixed
    number of steps per member, batched (members below 2^20 DOF share one
    launch per stage of the step per order, the others keep their own
    launches; one HIP graph) against the members run one after the other;
  * adjoint: the reverse-mode derivative on the 3-D Taylor-Green flow at
    128^3 elements (p = 3, FV+FB, HLLC, periodic [0, 2 pi]^3, seeded density
    perturbation of amplitude 1e-2), a short run (4 steps, a checkpoint every
    2): gradient of the final kinetic energy with respect to the initial
    density field through the checkpointed torch.autograd.Function, its time
    against the forward run, and central differences of the frozen-flag map
    on 16 random directions; adjoint_fp32_bulk: the FP32-bulk form's
    gradient at 64^3 (20 steps) against the FP64 form's
    (adjoint_fp64_reference);
  * cpu_baseline: the NumPy oracle on a bounded sample, one host core.

--precision-run fp32 (or perturbed) runs only the precision configuration's
Kelvin-Helmholtz leg instead: the FP64 form and the FP32-bulk form (or the
FP64 form from an initial state perturbed by 1e-14) to t = 2.0, with the
relative L1 difference per field at five times (about 15 minutes).

Usage: python bench.py [--gpus 1] [--steps 200] [--warmup 10]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBPS = 8000.0          # MI355X HBM3E, AMD published peak
FP64_PEAK_TFLOPS = 78.6         # MI355X FP64 vector / matrix dense peak


def kh_kwargs(N, p):
    return dict(p=p, N=(N, N), BC=(("periodic", "periodic"), ("periodic", "periodic")),
                update="FV", FB=True, riemann_solver_sd="hllc", riemann_solver_fv="hllc",
                slope_limiter="minmod", cfl_coeff=0.8)


def algorithmic_bytes(N, p, launches):
    """Compulsory HBM bytes per launch of each kernel class for one step of the
    2-D tiled FV+FB path in FP64 (each array element read or written once;
    flags 1 B).

    Per step (V = 4, T = n = p+1, m = p, E = N^2 elements, C = E n^2 cells,
    Fe = T V (N+1) N n element-face values per dimension, Fd = V (N n + 1) N n
    FV-face values per dimension, G = (N n + 4)^2 ghosted cells):
      k_face_sd  1 launch (k_face_sd2d, the final predictor state, for the
                 corrector): read T V C, write the element-face fluxes of both
                 dimensions (2 Fe)
      k_pred     m launches (k_pred2d, element faces solved inside): read the
                 state (T V C) and U_sp (V C), write T V C; the first iteration
                 reads U_sp and, when another follows, writes its increment
                 (V C), which the second reads with U_sp (2 V C) in place of
                 the state (the one-element halo reads are not counted)
      k_hoflux   T launches (k_hocand2d with the first half of the
                 detection): read the node's state slice (V C), its
                 element-face fluxes (2 V (N+1) N n) and U_cv (V C), write the
                 candidate (V C) and C flags (the FV-face arrays that the
                 column-strip fallback reads are written on the substeps that
                 take that path only, and are not counted)
      k_tensor   once: read + write V C
      k_detect   T launches (k_sed2d): read C flags (tiles with smooth-extrema
                 candidates also read the candidate's density and pressure;
                 not counted)
      k_theta1   T launches (k_theta2d, then the active-tile list): read C
                 flags, write G
      k_fbface   T launches (k_fbmarch2d or k_fbtile2d): read U_cv, theta, both
                 face arrays, write V C (at most: cells with no blended face
                 keep the candidate; rows and tiles without a blended face are
                 skipped, so the launch reads less where few cells are
                 troubled)
      k_cv2sp    once (k_tensor2d cv->sp with the wave-speed maximum of the
                 cells fused): read V C and write V C
    """
    V, n = 4, p + 1
    T = n
    m = p
    E = N * N
    C = E * n * n
    Fd = V * (N * n + 1) * N * n
    Fe = T * V * (N + 1) * N * n
    G = (N * n + 4) ** 2
    per_step = {
        "k_face_sd": T * V * C + 2 * Fe,
        "k_pred": ((2 * V * C) + (2 * V * C + T * V * C) + (m - 2) * (2 * T * V * C + V * C)
                   if m >= 2 else V * C + T * V * C),
        "k_hoflux": T * (3 * V * C + 2 * V * (N + 1) * N * n),
        "k_tensor": 2 * V * C,
        "k_detect": 0,
        "k_theta1": T * G,
        "k_fbface": T * (2 * V * C + G + 2 * Fd),
        "k_cv2sp": 2 * V * C,
    }
    extra_bytes = {"k_hoflux": T * C * 1.0, "k_detect": T * C * 1.0, "k_theta1": T * C * 1.0}
    out = {}
    steps = launches["_steps"]
    for k, val in launches.items():
        if k not in per_step:
            continue
        cnt = val[1]
        out[k] = (per_step[k] * 8.0 + extra_bytes.get(k, 0.0)) * steps / cnt
    return out


# rocprofv3 names of the kernels of each class in the FP64 2-D HLLC workload,
# with their launches per step
CLASS_KERNELS = {
    "k_fbface": [("k_fbmarch2d<1, double, false>", 4), ("k_fbtile2d<3, 1, double>", 4)],
    "k_face_sd": [("k_face_sd2d<3, 1, double, double>", 1)],
    "k_pred": [("k_pred2d<3, 1, double, double, true, false>", 1),
               ("k_pred2d<3, 1, double, double, false, true>", 1),
               ("k_pred2d<3, 1, double, double, false, false>", 1)],
    "k_detect": [("k_sed2d", 4)],
    "k_hoflux": [("k_hocand2d<3, double, true, true>", 4)],
    "k_theta1": [("k_theta2d", 4), ("k_fbmark_list", 4)],
    "k_cv2sp": [("k_tensor2d<3, 0>", 1)],
    "k_tensor": [("k_tensor2d<3, 1>", 1)],
}
TRAFFIC_SUMMARY = os.path.join(HERE, "profiles", "kh1024_p3_hbm_traffic.md")
FLOPS_SUMMARY = os.path.join(HERE, "profiles", "kh1024_p3_fp64_flops.md")
T1_SUMMARY = os.path.join(HERE, "profiles", "kh1024_p3_t1_window.md")


def recorded_t1_window():
    """The t = 1 window of the same workload as recorded by `--t1-window`
    (reaching t = 1 takes ~26k steps, several minutes, so the default run
    reports the committed record instead of measuring it); None when absent."""
    if not os.path.exists(T1_SUMMARY):
        return None
    rows = {}
    with open(T1_SUMMARY) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 2:
                rows[cells[0]] = cells[1]
    try:
        out = {"ms_per_step": float(rows["ms per step"]),
               "value": float(rows["DOF-updates/s"]),
               "steps_to_t1": int(rows["steps to t = 1.0"].split()[0]),
               "troubled_fraction": float(rows["troubled fraction (substep average of the next step)"]
                                          .rstrip("%")) / 100.0}
    except (KeyError, ValueError):
        return None
    out["source"] = os.path.relpath(T1_SUMMARY, HERE)
    return out


def measured_traffic(cls):
    """HBM bytes per launch of a kernel class from the committed counter
    summary (FETCH_SIZE / WRITE_SIZE passes scaled by the copy calibration),
    averaged over the class's kernels weighted by their launches per step;
    None when not recorded."""
    entries = CLASS_KERNELS.get(cls)
    if not entries or not os.path.exists(TRAFFIC_SUMMARY):
        return None
    names = dict(entries)
    vals = {}
    with open(TRAFFIC_SUMMARY) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 6 and cells[0].strip("`") in names:
                try:
                    vals[cells[0].strip("`")] = (float(cells[4]) + float(cells[5])) * 1024.0 * 1024.0
                except ValueError:
                    pass
    if len(vals) != len(names):
        return None
    return sum(vals[k] * w for k, w in names.items()) / sum(names.values())


def measured_step_traffic(prof, steps):
    """HBM bytes per step summed over the kernel classes of the event window
    (each class's counter bytes per launch times its launches per step);
    None when a class is missing from the counter summary."""
    total = 0.0
    for k, (_, cnt) in prof.items():
        t = measured_traffic(k)
        if t is None:
            return None
        total += t * cnt / steps
    return total


def measured_fp64_flops(cls):
    """FP64 flop per launch of a kernel class from the committed counter
    summary (64 x (ADD + MUL + 2 FMA + TRANS) FP64 instructions per dispatch),
    averaged over the class's kernels weighted by their launches per step;
    None when not recorded."""
    entries = CLASS_KERNELS.get(cls)
    if not entries or not os.path.exists(FLOPS_SUMMARY):
        return None
    names = dict(entries)
    vals = {}
    col = None  # the column headed "FP64 GFLOP per dispatch"
    with open(FLOPS_SUMMARY) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if "FP64 GFLOP per dispatch" in cells:
                col = cells.index("FP64 GFLOP per dispatch")
            elif col is not None and len(cells) > col and cells[0].strip("`") in names:
                try:  # the first table (the bench window) holds
                    vals.setdefault(cells[0].strip("`"), float(cells[col]) * 1e9)
                except ValueError:
                    pass
    if len(vals) != len(names):
        return None
    return sum(vals[k] * w for k, w in names.items()) / sum(names.values())


def step_traffic_summary(step_bytes, ms_step, alg_bytes):
    """Measured HBM bytes per step (counter summary) against the step time,
    beside the metric's algorithmic bytes per step (64 B per DOF)."""
    if step_bytes is None:
        return None
    gbps = step_bytes / (ms_step * 1e-3) / 1e9
    return {"bytes_per_step": step_bytes, "GBps": gbps, "frac": gbps / HBM_PEAK_GBPS,
            "algorithmic_bytes_per_step": alg_bytes,
            "traffic_over_algorithmic": step_bytes / alg_bytes, "source": os.path.relpath(TRAFFIC_SUMMARY, HERE)}


def copy_bandwidth(torch, nbytes=2 << 30, reps=10):
    """Measured device-to-device copy bandwidth (read + write bytes / time)."""
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del a, b
    return 2.0 * nbytes * reps / dt / 1e9


def cpu_baseline_child(N, steps, warmup):
    """Time the NumPy oracle on one core (run in a child with 1 thread)."""
    import numpy as np  # noqa: F401
    from spd_amd import initial_conditions as ic
    from oracle.spd_oracle import OracleSimulator
    p = 3
    noise = ic.subcell_noise((N, N), p, [1, 2], 1e-3, seed=0)
    o = OracleSimulator(ic.kelvin_helmholtz, noise=noise, **kh_kwargs(N, p))
    for _ in range(warmup):
        o.compute_dt()
        o.perform_update()
    t0 = time.perf_counter()
    for _ in range(steps):
        o.compute_dt()
        o.perform_update()
    dt = time.perf_counter() - t0
    dof = (N * (p + 1)) ** 2
    print(json.dumps({"value": dof * steps / dt, "seconds": dt, "steps": steps}))


def run_cpu_baseline(N=256, steps=1, warmup=1):
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               CUDA_VISIBLE_DEVICES="", HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child",
           "--cpu-N", str(N), "--cpu-steps", str(steps), "--cpu-warmup", str(warmup)]

    def pin():
        try:
            os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        except Exception:
            pass
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, preexec_fn=pin, timeout=900)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + r.stderr[-2000:])
    res = json.loads(r.stdout.strip().splitlines()[-1])
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": res["value"], "unit": "DOF-updates/s", "cores": 1, "kind": "port",
            "sample": (f"NumPy oracle (FP64 restatement of the reference), KH {N}^2 p=3 FV+FB, "
                       f"{warmup} warm-up + {steps} timed step(s) = {res['seconds']:.1f} s on one "
                       f"core of {os.cpu_count()} ({cpu})")}


def run_ensemble(torch, n=64, steps=10, sizes=(64, 128, 256, 384), seed=0):
    """The batched ensemble of 2-D four-quadrant Riemann problems (outflow,
    FV+FB, HLLC, p and N drawn from a seed) on a bounded sample of ``steps``
    steps per member.

    A probe run of the ``steps + 1`` steps finds the members the scheme drives
    to non-finite values (for the small-member draw the reference's census,
    tests/golden/ensemble_small_nonfinite.json, names the same members; it is
    reported beside the count); only the members that stay finite are timed.
    Both legs start from the same saved state (after one step that builds the
    batched graph): the batched leg (one HIP graph launch for the steps; the
    members of each order share one launch per stage of the step, the groups
    on separate capture branches) and the members one after the other with
    the host-driven loop (compute_dt + step, no snapshot writes) and the same
    kernels (substep form 2)."""
    import contextlib
    from spd_amd.ensemble import Ensemble, riemann_member_kwargs, riemann_members, timed
    specs = riemann_members(n, seed=seed, sizes=sizes)
    out = {"members": n, "sizes": list(sizes), "orders": [2, 3, 4, 5], "seed": seed,
           "steps_per_member": steps}
    with contextlib.redirect_stdout(sys.stderr):
        probe = Ensemble([riemann_member_kwargs(m) for m in specs], graph_steps=4, branches=8,
                         tune=False)
        probe.run(steps=steps + 1, finish=False)
        torch.cuda.synchronize()
        bad = [i for i, s_ in enumerate(probe.sims)
               if not bool(torch.isfinite(s_.dm.device_value("U_sp")).all())]
        del probe
        torch.cuda.empty_cache()
        out["members_nonfinite"] = len(bad)
        census = os.path.join(HERE, "tests", "golden", "ensemble_small_nonfinite.json")
        if os.path.exists(census):
            with open(census) as f:
                c = json.load(f)
            d = c["draw"]
            if (d["n"], d["seed"], tuple(d["sizes"])) == (n, seed, tuple(sizes)):
                ref_bad = [i for i, k in enumerate(c["first_nonfinite_step"])
                           if k is not None and k <= steps + 1]
                out["members_nonfinite_reference"] = len(ref_bad)
                out["nonfinite_same_members_as_reference"] = sorted(ref_bad) == sorted(bad)
        keep = [m for i, m in enumerate(specs) if i not in set(bad)]
        out["members_timed"] = len(keep)
        t0 = time.perf_counter()
        ens = Ensemble([riemann_member_kwargs(m) for m in keep], graph_steps=steps, branches=8,
                       tune=True)
        setup = time.perf_counter() - t0
        dof = sum(ens.dof)
        ens.run(steps=1, finish=False)                      # builds the graph once
        saved = [(s_.dm.device_value("U_sp").clone(), s_.time, s_.n_step, s_.dm.dt)
                 for s_ in ens.sims]

        def restore():
            for s_, (U, t, k, dt) in zip(ens.sims, saved):
                s_.dm.device_value("U_sp").copy_(U)
                s_.time, s_.n_step, s_.dm.dt = t, k, dt
                s_._dt_from_step = True

        def sequential():
            for s_ in ens.sims:
                s_.advance(steps)

        tb = timed(lambda: ens.run(steps=steps, finish=False))
        restore()
        # the recorded wave speeds belong to the batched leg's last states:
        # take each member's first dt from its saved state instead
        for s_ in ens.sims:
            s_._dt_from_step = False
            s_.dm.set_device_value("W_cv", s_._pointwise(
                "spd_cons_to_prims", s_._tensor_op("spd_sp_to_cv", s_.dm.device_value("U_sp"))))
        ts = timed(sequential)
    out.update({"dof_total": dof,
                "batched_DOF_updates_per_s": dof * steps / tb,
                "sequential_DOF_updates_per_s": dof * steps / ts,
                "speedup_vs_sequential": ts / tb, "graph_builds": ens.graph_builds,
                "graph_build_s": ens.build_seconds, "setup_and_tune_s": setup})
    del ens
    torch.cuda.empty_cache()
    return out


def smooth_direction(torch, s, rng, modes=6, kmax=2):
    """A random smooth field at the solution points of ``s`` (element layout):
    a sum of ``modes`` Fourier modes with integer wave numbers up to ``kmax``
    over the domain and random phases and amplitudes."""
    import numpy as np
    nd, n = s.ndim, s.p + 1
    X = []
    for a, d in enumerate(s.dims):
        lo = s.lim[d][0]
        e = np.arange(s.N[d])[:, None]
        x = lo + (e + np.asarray(s.x_sp)[None, :]) * s.h[d]          # (N_a, n)
        shape = [1] * (2 * nd)
        shape[nd - 1 - a] = s.N[d]
        shape[2 * nd - 1 - a] = n
        X.append(torch.as_tensor(x.reshape(shape), device="cuda"))
    out = 0.0
    for _ in range(modes):
        kv = rng.integers(-kmax, kmax + 1, size=nd)
        ph = rng.uniform(0, 2 * np.pi)
        arg = sum(float(kv[a]) * 2 * np.pi / s.len[d] * (X[a] - s.lim[d][0])
                  for a, d in enumerate(s.dims)) + ph
        out = out + rng.standard_normal() * torch.cos(arg)
    return out.contiguous()


def run_adjoint(torch, N=64, steps=20, k=5, n_fd=2, precision="fp64", ref_grad=None,
                fd_steps=(1e-7, 1e-8, 1e-9), offload=False):
    """Gradient of the final kinetic energy of the 3-D Taylor-Green flow with
    respect to the initial density field (``steps`` steps, a checkpoint every
    ``k``), timed against the forward run; central differences (steps 1e-7, 1e-8
    and 1e-9 of the density scale, the closest one kept) of the same frozen-flag
    map (the forward's dt and blending coefficients replayed) on ``n_fd``
    random smooth directions (white-noise directions put many limiter
    arguments within the difference step of a kink at this size).  With
    ``precision="fp32"`` the FP32-bulk form is differentiated and its gradient
    compared with ``ref_grad`` (the FP64 form's) instead.  With ``offload`` the
    checkpoints and recomputed states live in host memory (the configuration's
    200 steps at 128^3).  Returns (summary, gradient)."""
    import math
    import numpy as np
    from spd_amd import initial_conditions as ic
    from spd_amd.adjoint import Adjoint
    from spd_amd.simulator import SDADER_Simulator
    L = 2 * math.pi
    s = SDADER_Simulator(init_fct=ic.taylor_green, p=3, N=(N, N, N), xlim=(0, L), ylim=(0, L),
                         zlim=(0, L), update="FV", FB=True, riemann_solver_sd="hllc",
                         riemann_solver_fv="hllc", device="cuda:0", verbose=False,
                         precision=precision)
    adj = Adjoint(s, offload=offload, verbose=steps > 20)
    U0 = s.dm.device_value("U_sp").clone().contiguous()
    gen = torch.Generator(device="cuda").manual_seed(0)
    U0[0] *= 1.0 + 1e-2 * (2 * torch.rand(U0[0].shape, generator=gen, device="cuda",
                                          dtype=torch.float64) - 1)
    W0 = s._pointwise("spd_cons_to_prims", U0)
    del U0
    gam = s.gamma

    def cons(rho):
        vel = [W0[1 + a] for a in range(3)]
        E = W0[4] / (gam - 1) + 0.5 * rho * sum(v * v for v in vel)
        return torch.stack([rho] + [rho * v for v in vel] + [E]).contiguous()

    def ke_pts(U):
        return 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]

    def ke(U):
        return ke_pts(U).sum()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        UN, _, dts = adj._forward(cons(W0[0]), steps, k, None)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    del UN
    rho = W0[0].clone().requires_grad_(True)
    torch.cuda.reset_peak_memory_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = ke(adj.run(cons(rho), steps, checkpoint_every=k))
    J.backward()
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    progress(f"adjoint {N}^3 {precision}: forward {t_fwd:.2f} s, gradient {t_grad:.2f} s")
    peak = torch.cuda.max_memory_allocated()
    g = rho.grad
    errs = []
    with torch.no_grad():
        U = cons(W0[0])
        thetas = []
        for st in range(steps if n_fd > 0 else 0):
            U, th = adj.step(U, dts[st:st + 1], record_theta=True)
            thetas.append(th)

        def final(r):
            U = cons(r)
            for st in range(steps):
                U = adj.step(U, dts[st:st + 1], theta=thetas[st])
            return U

        def dJ(d, eps):
            # pointwise difference first: the sum of 10^7 kinetic energies
            # would bury the O(eps) difference in its own round-off
            return float((ke_pts(final(W0[0] + eps * d)) - ke_pts(final(W0[0] - eps * d))).sum())

        rng = np.random.default_rng(5)
        for i in range(n_fd):
            d = smooth_direction(torch, s, rng)
            ad = float((g * d).sum())
            e = [abs(dJ(d, eps) / (2 * eps) - ad) / abs(ad) for eps in fd_steps]
            errs.append(min(e))
            progress(f"adjoint {N}^3 {precision}: direction {i + 1} of {n_fd}, "
                     f"relative difference {errs[-1]:.2e}")
    out = {"workload": f"Taylor-Green 3-D {N}^3 p=3 FV+FB HLLC periodic, "
                       f"{precision.upper()}, density perturbation 1e-2, "
                       f"d(final kinetic energy)/d(initial density)",
           "steps": steps, "checkpoint_every": k, "checkpoints_in_host_memory": offload,
           "dof": (N * 4) ** 3,
           "forward_s": t_fwd, "gradient_s": t_grad, "gradient_over_forward": t_grad / t_fwd,
           "peak_torch_GiB": peak / 2 ** 30,
           "scratch_GiB": float(s._lib.spd_ctx_scratch_bytes(s._ctx)) / 2 ** 30,
           "fd_directions": n_fd, "fd_rel_err_max": max(errs) if errs else None}
    if ref_grad is not None:
        out["grad_rel_l2_vs_fp64"] = float((g - ref_grad).norm() / ref_grad.norm())
    g = g.detach().clone()
    del adj, s
    torch.cuda.empty_cache()
    return out, g


def progress(msg):
    """One progress line on stderr (the JSON result goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def l1_rel_fields(Ua, Ub):
    """Relative L1 difference of the 2-D states Ua and Ub per field (over all
    solution points, relative to Ub's L1 norm): rho, momenta (x and y
    together), E."""
    out = {name: float((Ua[v] - Ub[v]).abs().sum() / Ub[v].abs().sum())
           for name, v in (("rho", 0), ("E", 3))}
    out["momenta"] = float((Ua[1:3] - Ub[1:3]).abs().sum() / Ub[1:3].abs().sum())
    return out


def run_precision(torch, make_ref, make_other, other, marks, chunk=2000):
    """The Kelvin-Helmholtz run of the precision configuration: the FP64 form
    (make_ref) and a second simulation (make_other: the FP32-bulk form, or the
    FP64 form from a perturbed initial state) each advanced to every time in
    `marks` by the device-side loop with its clamp to that time (their dt
    sequences differ), and the relative L1 difference per field at each mark."""
    a, b = make_ref(), make_other()
    a.init_sim()
    b.init_sim()
    t0 = time.perf_counter()
    rows = []
    for tm in marks:
        for name, s in (("fp64", a), (other, b)):
            while s.time < tm:
                s._run_on_device(tm, chunk, 1)
                progress(f"precision run: {name} t = {s.time:.4f} after {s.n_step} steps, "
                         f"{time.perf_counter() - t0:.0f} s")
        torch.cuda.synchronize()
        Ua, Ub = a.dm.device_value("U_sp"), b.dm.device_value("U_sp")
        rows.append({"t": tm, "steps_fp64": a.n_step, f"steps_{other}": b.n_step,
                     "finite": bool(torch.isfinite(Ua).all() and torch.isfinite(Ub).all()),
                     "l1_rel": l1_rel_fields(Ub, Ua)})
        progress(f"precision run: t = {tm}: {rows[-1]['l1_rel']}")
    return {"other": other, "marks": rows, "seconds": time.perf_counter() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="steps of the HIP-event window that times each kernel")
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-child", action="store_true")
    ap.add_argument("--cpu-N", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-warmup", type=int, default=1)
    ap.add_argument("--no-fp32", action="store_true",
                    help="skip the FP32-bulk leg of the headline workload (counter passes "
                         "that need the FP64 kernels only)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg (for profiling) and print its JSON")
    ap.add_argument("--adjoint-N", type=int, default=128,
                    help="elements per side of the FP64 adjoint leg (with the FD check)")
    ap.add_argument("--adjoint-steps", type=int, default=4)
    ap.add_argument("--adjoint-k", type=int, default=2)
    ap.add_argument("--adjoint-fd", type=int, default=16)
    ap.add_argument("--adjoint-offload", action="store_true",
                    help="keep the adjoint's checkpoints in host memory (with --adjoint-only)")
    ap.add_argument("--adjoint-fp32-N", type=int, default=64,
                    help="elements per side of the FP32-bulk adjoint leg and of the FP64 "
                         "gradient it is compared with")
    ap.add_argument("--adjoint-fp32-steps", type=int, default=20)
    ap.add_argument("--adjoint-fp32-k", type=int, default=5)
    ap.add_argument("--precision-run", choices=["fp32", "perturbed"], default=None,
                    help="run only the precision configuration's Kelvin-Helmholtz leg: the FP64 "
                         "form and the FP32-bulk form (or the FP64 form from an initial state "
                         "perturbed by 1e-14 on vx) to --precision-t-end, with the relative L1 "
                         "difference per field at t_end/8, t_end/4, t_end/2, 3 t_end/4, t_end")
    ap.add_argument("--precision-t-end", type=float, default=2.0)
    ap.add_argument("--t1-window", action="store_true",
                    help="also run the FP64 form to t = 1.0 (about 23k steps, 5 minutes) and "
                         "time the same number of steps there")
    args = ap.parse_args()

    if args.cpu_baseline_child:
        cpu_baseline_child(args.cpu_N, args.cpu_steps, args.cpu_warmup)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py: this build runs the whole domain on one GPU; --gpus {args.gpus} "
              f"(WORLD_SIZE={world}) is not supported (multi-GPU halo exchange is design-only).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np  # noqa: F401
    import torch
    from spd_amd import initial_conditions as ic
    from spd_amd.simulator import SDADER_Simulator

    N, p = args.N, args.p
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.adjoint_only:
        print(json.dumps({"adjoint": run_adjoint(torch, args.adjoint_N, args.adjoint_steps,
                                                 args.adjoint_k, args.adjoint_fd,
                                                 offload=args.adjoint_offload)[0]}))
        return
    dof = (N * (p + 1)) ** 2
    noise = ic.subcell_noise((N, N), p, [1, 2], 1e-3, seed=0)

    def make_sim(precision):
        s = SDADER_Simulator(init_fct=ic.kelvin_helmholtz, device=dev, verbose=False,
                             precision=precision, **kh_kwargs(N, p))
        s.perturb_initial_condition(noise)
        return s

    if args.precision_run:
        te = args.precision_t_end
        marks = [te / 8, te / 4, te / 2, 3 * te / 4, te]
        if args.precision_run == "fp32":
            other = lambda: make_sim("fp32")  # noqa: E731
        else:
            def other():
                s = SDADER_Simulator(init_fct=ic.kelvin_helmholtz, device=dev, verbose=False,
                                     precision="fp64", **kh_kwargs(N, p))
                s.perturb_initial_condition(noise + ic.subcell_noise((N, N), p, [1], 1e-14, seed=7))
                return s
        res = run_precision(torch, lambda: make_sim("fp64"), other, args.precision_run, marks)
        res["workload"] = (f"Kelvin-Helmholtz {N}^2 p={p} FV+FB HLLC minmod CFL 0.8 periodic, "
                           f"seeded noise 1e-3 on vx, vy, to t = {te}")
        print(json.dumps({"precision_run": res}))
        return

    def run_form(precision):
        """Warm-up, then the timed steps: the device-side time loop of
        perform_time_evolution (dt, step and time on the device, replayed from
        one HIP graph captured beforehand); then the same steps from the same
        state with the host-driven loop (dt read back every step, kernels
        launched one by one) on a second simulator, which must end in the
        same state bit for bit; then as many steps with HIP events around
        every kernel (the per-kernel durations of the roofline), then one
        more step that measures the troubled fraction."""
        s = make_sim(precision)
        s.advance(args.warmup)
        progress(f"{precision}: {args.warmup} warm-up steps")
        early = s.dm.device_value("U_sp").clone()
        s.prepare_graph(args.steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.advance_device(args.steps)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        progress(f"{precision}: {args.steps} timed steps (device loop), "
                 f"{elapsed / args.steps * 1e3:.3f} ms per step")
        timed_end = s.dm.device_value("U_sp").clone()
        h = make_sim(precision)
        h.advance(args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.advance(args.steps)
        torch.cuda.synchronize()
        host_loop = time.perf_counter() - t0
        progress(f"{precision}: the same steps with the host loop, "
                 f"{host_loop / args.steps * 1e3:.3f} ms per step")
        same = bool(torch.equal(h.dm.device_value("U_sp"), timed_end))
        del h
        torch.cuda.empty_cache()
        s.profile_begin(max_launches=200 * max(args.profile_steps, 1))
        s.advance(args.profile_steps)
        prof = s.profile_end()
        frac = s.troubled_fraction_next_step()
        return s, elapsed, prof, frac, early, host_loop, timed_end, same

    s, elapsed, prof, frac, early64, hloop64, end64, same64 = run_form("fp64")
    fp32 = None
    if not args.no_fp32:
        s32, elapsed32, prof32, frac32, early32, hloop32, end32, same32 = run_form("fp32")

        # the two forms after the same number of steps from the same state
        l1_rel = l1_rel_fields
        torch.cuda.synchronize()
        l1 = l1_rel(s32.dm.device_value("U_sp"), s.dm.device_value("U_sp"))
        l1_early = l1_rel(early32, early64)
        l1_timed = l1_rel(end32, end64)
        del s32, early32, end32
        value32 = dof * args.steps / elapsed32
        fp32 = {
            "value": value32, "unit": "DOF-updates/s", "ms_per_step": elapsed32 / args.steps * 1e3,
            "host_loop_ms_per_step": hloop32 / args.steps * 1e3,
            "host_loop_same_state": same32,
            "speedup_vs_fp64": value32 / (dof * args.steps / elapsed),
            "step_hbm_roofline_frac": value32 * 64.0 / (HBM_PEAK_GBPS * 1e9),
            "l1_rel_vs_fp64": l1,
            "steps_compared": args.warmup + args.steps + args.profile_steps + 1,
            "l1_rel_vs_fp64_after_warmup": l1_early, "steps_compared_after_warmup": args.warmup,
            "l1_rel_vs_fp64_after_timed": l1_timed,
            "steps_compared_after_timed": args.warmup + args.steps,
            "troubled_fraction": frac32,
            "kernels_ms_per_step": {k: v[0] / args.profile_steps
                                    for k, v in sorted(prof32.items(), key=lambda kv: -kv[1][0])},
        }
    del early64, end64

    ms_step = elapsed / args.steps * 1e3
    value = dof * args.steps / elapsed
    launches = dict(prof)
    launches["_steps"] = args.profile_steps
    bpl = algorithmic_bytes(N, p, launches)
    dom = max(prof, key=lambda k: prof[k][0])
    dom_ms, dom_cnt = prof[dom]
    avg_ms = dom_ms / dom_cnt
    achieved = bpl[dom] / (avg_ms * 1e-3) / 1e9
    copy_gbps = copy_bandwidth(torch)
    kernels = {}
    step_flops = 0.0
    for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        fl = measured_fp64_flops(k)
        avg_s = v[0] / v[1] * 1e-3
        kernels[k] = {"ms_per_step": v[0] / args.profile_steps,
                      "launches_per_step": v[1] / args.profile_steps,
                      "GBps_alg": (bpl[k] / avg_s / 1e9) if k in bpl else None,
                      "TFLOPs_fp64": (fl / avg_s / 1e12) if fl is not None else None}
        if fl is not None:
            step_flops += fl * v[1] / args.profile_steps
    dom_fl = measured_fp64_flops(dom)

    out = {
        "metric": "DOF-updates/s, 2D p=3 1024\u00b2 KH: FP64 vs FP32-bulk/batched; %HBM roofline",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Kelvin-Helmholtz initial condition)",
        "config": {"workload": f"Kelvin-Helmholtz {N}^2 p={p} FV+FB HLLC/HLLC minmod "
                               f"CFL 0.8 periodic, FP64 (plain form)",
                   "N": N, "p": p, "dof_per_step": dof, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": measured_traffic(dom), "avg_launch_ms": avg_ms,
                     "alg_bytes_per_launch": bpl[dom],
                     "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, HERE)},
        "step_hbm_roofline_frac": value * 64.0 / (HBM_PEAK_GBPS * 1e9),
        "step_traffic": step_traffic_summary(measured_step_traffic(prof, args.profile_steps),
                                             ms_step, dof * 64.0),
        "compute_roofline": {
            "bound": "fp64 valu", "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "kernel": dom,
            "achieved": (dom_fl / (avg_ms * 1e-3) / 1e12) if dom_fl is not None else None,
            "frac": (dom_fl / (avg_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS)
            if dom_fl is not None else None,
            "step_achieved": step_flops / (ms_step * 1e-3) / 1e12 if step_flops else None,
            "step_flop_per_dof": step_flops / dof if step_flops else None,
            "flops_source": os.path.relpath(FLOPS_SUMMARY, HERE)},
        "copy_GBps": copy_gbps,
        "timed_loop": "device-side time loop (one HIP graph of the timed steps)",
        "host_loop_ms_per_step": hloop64 / args.steps * 1e3,
        "host_loop_same_state": same64,
        "troubled_fraction": frac,
        "kernels": kernels,
    }
    if fp32 is not None:
        out["fp32_bulk"] = fp32
    if not args.t1_window:
        rec = recorded_t1_window()
        if rec is not None:
            out["t1_window_recorded"] = rec
    if args.t1_window:
        # the device loop with the clamp to t_end, in chunks, with progress
        s.init_sim()
        t_start = time.perf_counter()
        while s.time < 1.0:
            s._run_on_device(1.0, 2000, 1)
            print(f"t1 window: t = {s.time:.4f} after {s.n_step} steps, "
                  f"{time.perf_counter() - t_start:.0f} s", file=sys.stderr, flush=True)
        t1, n1 = s.time, s.n_step
        s.prepare_graph(args.steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.advance_device(args.steps)
        torch.cuda.synchronize()
        el1 = time.perf_counter() - t0
        U = s.dm.device_value("U_sp")
        out["t1_window"] = {
            "t_start": t1, "step_start": n1,
            "value": dof * args.steps / el1, "ms_per_step": el1 / args.steps * 1e3,
            "troubled_fraction": s.troubled_fraction_next_step(),
            "finite": bool(torch.isfinite(U).all()),
            "rho_min": float(U[0].min()), "rho_max": float(U[0].max())}
    if not args.no_ensemble:
        del s
        torch.cuda.empty_cache()
        out["ensemble"] = run_ensemble(torch, steps=args.ensemble_steps)
        out["ensemble_small_members"] = run_ensemble(torch, steps=2 * args.ensemble_steps,
                                                     sizes=(16, 32, 64))
    if not args.no_adjoint:
        if "ensemble" not in out:
            del s
            torch.cuda.empty_cache()
        # the differentiation configuration's grid, a short run: the FP64 form
        # with central differences of its frozen-flag map on random directions
        out["adjoint"], g = run_adjoint(torch, args.adjoint_N, args.adjoint_steps,
                                        args.adjoint_k, args.adjoint_fd, fd_steps=(1e-7, 1e-8))
        del g
        torch.cuda.empty_cache()
        # the FP32-bulk form against the FP64 form's gradient on a grid where both
        # contexts' tapes and reverse-sweep buffers fit one after the other
        nf, sf, kf = args.adjoint_fp32_N, args.adjoint_fp32_steps, args.adjoint_fp32_k
        out["adjoint_fp64_reference"], g64 = run_adjoint(torch, nf, sf, kf, 0)
        out["adjoint_fp32_bulk"], _ = run_adjoint(torch, nf, sf, kf, 0, precision="fp32",
                                                  ref_grad=g64)
        del g64
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = run_cpu_baseline(args.cpu_N, args.cpu_steps, args.cpu_warmup)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
