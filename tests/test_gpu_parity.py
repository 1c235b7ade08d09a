"""GPU parity of the HIP step (through the C ABI, via the facade) with the
reference's golden vectors and with the CPU oracle run in lock-step.

Bar (FP64): every state array within 1e-12 norm-wise relative difference per
field (momenta share one scale), dt within 1e-12 relative, troubled-cell flags
identical to the oracle's except where the oracle's decision lies within 1e-10
(relative) of its threshold; theta and the face coefficients within 1e-12 when
the flags are identical (exact 0/0.25/.../1 values inside the domain; ghost
densities at outflow and reflective sides).
"""
import numpy as np
import pytest

from fixtures import (flag_mismatch, golden_names, load, make_ic, noise_of, relerr,
                      relerr_fields, sim_kwargs)

pytestmark = pytest.mark.gpu
TOL = 1e-12
NAMES = golden_names()


def gpu_sim(fx):
    from spd_amd.simulator import SDADER_Simulator
    cfg = fx["cfg"]
    s = SDADER_Simulator(init_fct=make_ic(cfg["ic"]), device="cuda:0", verbose=False,
                         **sim_kwargs(cfg))
    nz = noise_of(cfg)
    if nz is not None:
        s.perturb_initial_condition(nz)
    return s


def oracle_sim(fx):
    from oracle.spd_oracle import OracleSimulator
    cfg = fx["cfg"]
    return OracleSimulator(make_ic(cfg["ic"]), noise=noise_of(cfg), **sim_kwargs(cfg))


@pytest.mark.parametrize("name", NAMES)
def test_steps_vs_golden_and_oracle(name):
    print(f"[gpu parity] {name}", flush=True)
    fx = load(name)
    s = gpu_sim(fx)
    o = oracle_sim(fx)
    assert relerr_fields(s.W_init_cv, fx["W_init_cv"]) < TOL
    assert relerr_fields(s.dm.U_sp, fx["U_sp0"]) < TOL
    assert abs(s.dm.dt - float(fx["dt0"])) <= TOL * float(fx["dt0"])
    fb = s.update == "FV" and s.FB
    for k in range(1, int(fx["nsteps"]) + 1):
        s.compute_dt()
        o.compute_dt()
        assert abs(s.dm.dt - float(fx[f"dt_{k}"])) <= TOL * float(fx[f"dt_{k}"]), k
        o.captures = []
        o.perform_update()
        recs = s.perform_update_with_flags() if fb else None
        if not fb:
            s.perform_update()
        U = s.dm.U_sp
        assert relerr_fields(U, o.U_sp) < TOL, ("oracle", k, relerr_fields(U, o.U_sp))
        if f"U_sp_{k}" in fx:
            assert relerr_fields(U, fx[f"U_sp_{k}"]) < TOL, ("golden", k)
        if f"U_cv_{k}" in fx and s.update == "FV":
            assert relerr_fields(s.dm.U_cv, fx[f"U_cv_{k}"]) < TOL, k
            assert relerr_fields(s.dm.W_cv, fx[f"W_cv_{k}"]) < TOL, k
        if fb:
            for i, (rg, ro) in enumerate(zip(recs, o.captures)):
                bad, ndiff = flag_mismatch(rg["troubles"], ro["troubles"], ro["margin"])
                assert bad == 0, (k, i, bad)
                if ndiff:
                    continue
                # theta holds densities in outflow/reflective ghosts (the
                # reference's stale ghost fill), so it carries state round-off
                assert relerr(rg["theta"], ro["theta"]) < TOL, (k, i)
                for a in range(s.ndim):
                    d = "xyz"[a]
                    assert relerr(rg[f"aff_{d}"], ro[f"aff_{d}"]) < TOL, (k, i, d)


def _fb2d(name):
    kw = load(name)["cfg"]["kw"]
    return len(kw["N"]) == 2 and kw.get("update", "SD") == "FV" and kw.get("FB", False)


@pytest.mark.parametrize("form", [1, 2, 3])
@pytest.mark.parametrize("name", [n for n in NAMES if _fb2d(n)])
def test_fallback_forms_vs_oracle(name, form):
    """The 2-D FV substep has four kernel forms: 1 separate high-order flux
    and candidate kernels, separate detection and the fallback on cell tiles
    in LDS, all over the high-order face-flux arrays in HBM; 2 the fused
    high-order flux + candidate + first-half detection kernel writing the
    face-flux arrays and the fallback marching column strips over them; 3 the
    fused kernel without face-flux arrays and the fallback on the compacted
    list of active tiles; 4 (the default, which every other test runs) the
    path of 2 or of 3 chosen per substep on the device.  Forms 1 to 3 meet
    the same bar against the oracle."""
    from spd_amd import _lib
    print(f"[gpu parity fallback form {form}] {name}", flush=True)
    fx = load(name)
    s = gpu_sim(fx)
    _lib.check(s._lib.spd_set_launch(s._ctx, _lib.LAUNCH["fb_form"], form), "spd_set_launch")
    assert s._lib.spd_set_launch(s._ctx, _lib.LAUNCH["fb_form"], 5) < 0
    o = oracle_sim(fx)
    for k in range(1, int(fx["nsteps"]) + 1):
        s.compute_dt()
        o.compute_dt()
        o.perform_update()
        s.perform_update()
        assert relerr_fields(s.dm.U_sp, o.U_sp) < TOL, ("oracle", k, relerr_fields(s.dm.U_sp, o.U_sp))


@pytest.mark.parametrize("name", [n for n in NAMES if "final_U_sp" in load(n)])
def test_time_evolution_vs_golden(name):
    print(f"[gpu evolution] {name}", flush=True)
    fx = load(name)
    s = gpu_sim(fx)
    s.perform_time_evolution(fx["cfg"]["t_end"])
    assert s.n_step == int(fx["final_n_step"])
    assert abs(s.time - float(fx["final_time"])) < 1e-14
    assert relerr_fields(s.dm.U_sp, fx["final_U_sp"]) < TOL
    assert relerr_fields(s.dm.W_cv, fx["final_W_cv"]) < TOL


def test_reference_preserve_invariant():
    """The reference's own physics test (test/test_preserve.py:10-38): a density
    wave advected at v=1, P=1 to t=1 changes density but keeps velocity and
    pressure to 1e-14 (mean absolute), for SD, FV, FV+FB, FV+FB+godunov."""
    from spd_amd.simulator import SDADER_Simulator
    for ndim in (1, 2):
        for p in (1, 3):
            for kw in (dict(update="SD"), dict(update="FV"),
                       dict(update="FV", FB=True, godunov=True),
                       dict(update="FV", FB=True)):
                print(f"[preserve] ndim={ndim} p={p} {kw}", flush=True)
                s = SDADER_Simulator(p=p, N=(6,) * ndim, device="cuda:0", verbose=False, **kw)
                s.perform_time_evolution(1)
                assert np.mean(np.abs(s.dm.W_cv[0] - s.W_init_cv[0])) > 1e-14
                assert np.mean(np.abs(s.dm.W_cv[1:] - s.W_init_cv[1:])) < 1e-14


@pytest.mark.parametrize("bc", ["periodic", "gradfree"])
def test_theta_of_flags_on_large_grid(bc):
    """The 2-D blending kernel evaluates tiles whose staged cells (64 x 16
    outputs + 2 on every side) are all active cells on bytes, as the distance
    to the nearest troubled cell; that form needs grids of at least 68 cells
    per side, so the fixtures above never reach it.  On a Kelvin-Helmholtz grid
    of 256^2 cells, theta of every substep equals the oracle's blending rings
    (src/trouble_detection.py:139-169) applied to the step's own flags, bit for
    bit: everywhere when periodic, and on the cells whose 5 x 5 neighbourhood
    is active with outflow sides (their ghosts carry stale state values)."""
    from oracle.spd_oracle import OracleSimulator
    from spd_amd import initial_conditions as ic
    print(f"[gpu theta 256^2] {bc}", flush=True)
    kw = dict(p=3, N=(64, 64), BC=((bc, bc),) * 2, update="FV", FB=True,
              riemann_solver_sd="hllc", riemann_solver_fv="hllc")
    from spd_amd.simulator import SDADER_Simulator
    s = SDADER_Simulator(init_fct=ic.kelvin_helmholtz, device="cuda:0", verbose=False, **kw)
    # no velocity noise: the troubled cells lie along the two density jumps,
    # which leaves tiles with and without them
    blend = OracleSimulator(ic.kelvin_helmholtz, p=1, N=(2, 2))._blend
    n_trouble, n_empty_tiles = 0, 0
    for k in range(3):
        s.compute_dt()
        for rec in s.perform_update_with_flags():
            tr = rec["troubles"][0]
            n_trouble += int(tr.sum())
            pad = np.pad(tr, 2, mode="wrap" if bc == "periodic" else "constant")
            want = pad.copy()
            blend(pad, want)
            got = rec["theta"][0]
            assert got.shape == want.shape
            if bc == "periodic":
                assert np.array_equal(got, want), k
            else:
                assert np.array_equal(got[4:-4, 4:-4], want[4:-4, 4:-4]), k
            blocks = tr[:240, :192].reshape(15, 16, 3, 64)
            n_empty_tiles += int((blocks.sum(axis=(1, 3)) == 0).sum())
    assert n_trouble > 0 and n_empty_tiles > 0, (n_trouble, n_empty_tiles)


def _kh_sims(N, **extra):
    from oracle.spd_oracle import OracleSimulator
    from spd_amd import initial_conditions as ic
    from spd_amd.simulator import SDADER_Simulator
    kw = dict(p=3, N=(N, N), BC=(("periodic", "periodic"),) * 2, update="FV", FB=True,
              riemann_solver_sd="hllc", riemann_solver_fv="hllc", **extra)
    s = SDADER_Simulator(init_fct=ic.kelvin_helmholtz, device="cuda:0", verbose=False, **kw)
    return s, (lambda: OracleSimulator(ic.kelvin_helmholtz, **kw))


@pytest.mark.parametrize("limiter", ["minmod", "moncen"])
def test_fallback_quiet_rows_vs_oracle(limiter):
    """The column-strip fallback skips the rows whose 5-row neighbourhood has
    theta = 0 across the wave and the chunks that hold no such row.  Without
    velocity noise the Kelvin-Helmholtz layers on 128^2 cells trouble cells
    only along the two density jumps, so most chunks (4 rows at this size) and
    rows take those paths: states and flags against the oracle, 3 steps."""
    from spd_amd import _lib
    print(f"[gpu fallback quiet rows] {limiter}", flush=True)
    s, mk = _kh_sims(32, slope_limiter=limiter)
    _lib.check(s._lib.spd_set_launch(s._ctx, _lib.LAUNCH["fb_form"], 2), "spd_set_launch")
    o = mk()
    n_trouble = 0
    for k in range(3):
        s.compute_dt()
        o.compute_dt()
        o.captures = []
        o.perform_update()
        recs = s.perform_update_with_flags()
        for rg, ro in zip(recs, o.captures):
            bad, _ = flag_mismatch(rg["troubles"], ro["troubles"], ro["margin"])
            assert bad == 0, (k, bad)
            n_trouble += int(rg["troubles"].sum())
        assert relerr_fields(s.dm.U_sp, o.U_sp) < TOL, (k, relerr_fields(s.dm.U_sp, o.U_sp))
    assert n_trouble > 0


def test_xcd_band_tile_orders_vs_oracle():
    """At 128^2 elements (512^2 cells, p=3) the detection tiles (16 columns of
    32 cells) run in XCD column bands two columns wide; the final element
    faces have 8 tile columns of 16 elements, i.e. bands one column wide,
    which is launch order (bands of two columns, with gradfree and reflective
    sides too, are covered by test_gpu_parity_large.py): one step's flags and
    state against the oracle."""
    print("[gpu xcd band tile orders 512^2]", flush=True)
    s, mk = _kh_sims(128)
    o = mk()
    s.compute_dt()
    o.compute_dt()
    o.captures = []
    o.perform_update()
    recs = s.perform_update_with_flags()
    n_trouble = 0
    for rg, ro in zip(recs, o.captures):
        bad, _ = flag_mismatch(rg["troubles"], ro["troubles"], ro["margin"])
        assert bad == 0, bad
        n_trouble += int(rg["troubles"].sum())
    assert n_trouble > 0
    assert relerr_fields(s.dm.U_sp, o.U_sp) < TOL, relerr_fields(s.dm.U_sp, o.U_sp)


@pytest.mark.parametrize("forms", [(2, 1), (3, 1), (3, 2), (4, 2), (4, 3)])
def test_fallback_forms_agree_on_large_grid(forms):
    """The substep forms pairwise (column strips with quiet-row skipping,
    cell tiles, compacted tiles with the fused candidate, the adaptive
    choice) on a Kelvin-Helmholtz grid of 256^2 cells without noise, one step
    at a time from the same state: same flags, states within 1e-12 (forms
    2, 3 and 4: the same bits)."""
    from spd_amd import _lib
    print(f"[gpu fallback forms 256^2] {forms}", flush=True)
    a, _ = _kh_sims(64)
    b, _ = _kh_sims(64)
    _lib.check(a._lib.spd_set_launch(a._ctx, _lib.LAUNCH["fb_form"], forms[0]), "spd_set_launch")
    _lib.check(b._lib.spd_set_launch(b._ctx, _lib.LAUNCH["fb_form"], forms[1]), "spd_set_launch")
    for k in range(3):
        b.dm.U_sp = np.asarray(a.dm.U_sp)
        a.compute_dt()
        b.dm.dt = a.dm.dt
        ra = a.perform_update_with_flags()
        rb = b.perform_update_with_flags()
        for x, y in zip(ra, rb):
            assert np.array_equal(x["troubles"], y["troubles"]), k
        assert relerr_fields(a.dm.U_sp, b.dm.U_sp) < TOL, (k, relerr_fields(a.dm.U_sp, b.dm.U_sp))
        if 1 not in forms:
            assert np.array_equal(np.asarray(a.dm.U_sp), np.asarray(b.dm.U_sp)), k


def _forms_sim(kind, N, p, precision, limiter):
    from spd_amd import initial_conditions as ic
    from spd_amd.ensemble import riemann_member_kwargs, riemann_members
    from spd_amd.simulator import SDADER_Simulator
    if kind in ("kh", "kh-walls"):
        bc = (("periodic", "periodic"), ("reflective", "reflective")) if kind == "kh-walls" \
            else (("periodic", "periodic"),) * 2
        kw = dict(p=p, N=(N, N), BC=bc, update="FV", FB=True,
                  riemann_solver_sd="hllc", riemann_solver_fv="hllc", slope_limiter=limiter,
                  cfl_coeff=0.8, precision=precision)
        s = SDADER_Simulator(init_fct=ic.kelvin_helmholtz, device="cuda:0", verbose=False, **kw)
        s.perturb_initial_condition(ic.subcell_noise((N, N), p, [1, 2], 1e-3, seed=0))
        return s
    kw = riemann_member_kwargs(riemann_members(1, 7, (N,), (p,))[0], precision)
    kw["slope_limiter"] = limiter
    return SDADER_Simulator(device="cuda:0", **kw)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("kind,N,p,limiter", [("kh", 64, 3, "minmod"), ("kh", 128, 3, "moncen"),
                                              ("quadrants", 32, 5, "minmod"),
                                              ("kh-walls", 48, 2, "moncen")])
def test_fallback_forms_bit_identical(kind, N, p, limiter, precision):
    """Forms 2 (column strips over stored high-order face fluxes), 3
    (compacted tiles that form them again) and 4 (either, per substep, from
    the previous substep's count of active tiles, across steps too) give the
    same bits: the fallback's arithmetic is one set of uncontracted helpers.
    20 free-running steps of a noise-seeded Kelvin-Helmholtz grid (dense
    fallback first, sparse later; periodic, or reflective walls across y)
    and of a four-quadrant Riemann problem with gradient-free sides:
    identical states, times and step counts."""
    from spd_amd import _lib
    print(f"[gpu fallback forms bit-identical] {kind} N={N} p={p} {limiter} {precision}",
          flush=True)
    out = []
    for form in (2, 3, 4):
        s = _forms_sim(kind, N, p, precision, limiter)
        _lib.check(s._lib.spd_set_launch(s._ctx, _lib.LAUNCH["fb_form"], form), "spd_set_launch")
        s.perform_iterations(20)
        U = np.asarray(s.dm.U_sp)
        assert np.isfinite(U).all()
        out.append((U, s.time, s.n_step))
    for U, t, n in out[1:]:
        assert np.array_equal(U, out[0][0])
        assert t == out[0][1] and n == out[0][2]


@pytest.mark.parametrize("p,N", [(1, (64, 12)), (2, (56, 10)), (3, (32, 12)), (4, (20, 8)),
                                 (5, (14, 8))])
def test_candidate_kernel_block_runs_bit_identical(p, N):
    """The fused candidate kernel has a variant for rows of whole E-element
    blocks (E = 256 / (p+1)^2: 64, 28, 16, 10, 7; constant index arithmetic)
    and the general one (any elements per block, a partial last block).  The
    grids here are whole blocks wide, so the default takes the first; E - 1
    elements per block forces the second.  Both run the same arithmetic:
    3 steps of a noise-seeded Kelvin-Helmholtz grid give the same bits, and
    the default's first step stays within the bar of the oracle."""
    from oracle.spd_oracle import OracleSimulator
    from spd_amd import _lib
    from spd_amd import initial_conditions as ic
    from spd_amd.simulator import SDADER_Simulator
    print(f"[gpu candidate kernel block runs] p={p} N={N}", flush=True)
    E = 256 // (p + 1) ** 2
    assert N[0] % E == 0
    kw = dict(p=p, N=N, BC=(("periodic", "periodic"), ("periodic", "periodic")), update="FV",
              FB=True, riemann_solver_sd="hllc", riemann_solver_fv="hllc", cfl_coeff=0.8)
    noise = ic.subcell_noise(N, p, [1, 2], 1e-3, seed=0)
    out = []
    for epb in (0, E - 1):
        s = SDADER_Simulator(init_fct=ic.kelvin_helmholtz, device="cuda:0", verbose=False, **kw)
        s.perturb_initial_condition(noise)
        _lib.check(s._lib.spd_set_launch(s._ctx, _lib.LAUNCH["hoc_epb"], epb), "spd_set_launch")
        if epb == 0:
            o = OracleSimulator(ic.kelvin_helmholtz, noise=noise, **kw)
            o.perform_update()
            s.perform_update()
            assert relerr_fields(s.dm.U_sp, o.U_sp) < TOL, relerr_fields(s.dm.U_sp, o.U_sp)
            s.perform_iterations(2)
        else:
            s.perform_iterations(3)
        out.append(np.asarray(s.dm.U_sp))
    assert np.isfinite(out[0]).all()
    assert np.array_equal(out[0], out[1])
