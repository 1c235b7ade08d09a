"""GPU parity with the CPU oracle on grids large enough to reach the paths
that the golden fixtures (at most 16^2 elements) never take, with outflow
("gradfree") and reflective sides as well as periodic ones.

Paths and the sizes that reach them (2-D, p = 3 unless stated):
  * theta in bytes (k_theta2d): tiles whose staged 68 x 20 cells are all
    active cells, i.e. at least 68 cells per side;
  * XCD column bands: detection tiles when the grid has a multiple of 8
    columns of 32 cells (256 cells: 8 columns), the fused final element faces
    when it has a multiple of 8 columns of 16 elements; (256, 24) elements
    give 16 face-tile columns, so each XCD band is W = 2 columns wide and a
    tile's left neighbour is the other column of its band, or the last
    column of the band before it;
  * quiet-row skipping of the column-strip fallback: the troubled cells of a
    four-quadrant problem lie along the two centre lines, so whole rows of a
    strip have theta = 0.

Bar (FP64), every step: states within 1e-12 norm-wise relative per field,
flags identical to the oracle's except within 1e-10 of a threshold, theta
and the face coefficients within 1e-12 when the flags are identical.  The
four-quadrant states are drawn from seeded generators.  Seeds are not used
whose states the reference scheme drives to non-finite values within the 3
steps (see test_blowup.py for those), nor seeds whose candidate state grows
without bound at a wall (the blending coefficients of the ghost cells hold
the candidate's densities there, Q2 of the reference's stale ghost fill): on
such a state the scheme amplifies round-off by orders of magnitude per
substep (seed 12 at 64^2 reflective: ghost densities 1e8 within 3 steps, state
difference 4e-11), so each step also checks that the oracle's ghost
coefficients stay below three times the initial maximum density.
"""
import numpy as np
import pytest

from fixtures import flag_mismatch, relerr, relerr_fields

pytestmark = pytest.mark.gpu
TOL = 1e-12

CASES = [
    # (elements, p, side kind, seed of the quadrant states)
    ((64, 64), 3, "gradfree", 11),
    ((64, 64), 3, "reflective", 33),
    ((32, 32), 5, "gradfree", 37),
    ((32, 32), 5, "reflective", 34),
    ((256, 24), 3, "gradfree", 15),
    ((256, 24), 3, "reflective", 16),
    ((256, 24), 3, "periodic", 17),
]


def _pair(N, p, bc, seed, **extra):
    from oracle.spd_oracle import OracleSimulator
    from spd_amd import initial_conditions as ic
    from spd_amd.simulator import SDADER_Simulator
    states = ic.random_four_quadrant_states(np.random.default_rng(seed))
    kw = dict(p=p, N=N, BC=((bc, bc),) * 2, update="FV", FB=True,
              riemann_solver_sd="hllc", riemann_solver_fv="hllc", **extra)
    s = SDADER_Simulator(init_fct=ic.four_quadrant(states), device="cuda:0", verbose=False,
                         **kw)
    o = OracleSimulator(ic.four_quadrant(states), **kw)
    return s, o


def _lockstep(s, o, steps, check_theta=True):
    n_trouble = 0
    rho0 = float(np.max(o.W_cv[0]))
    for k in range(1, steps + 1):
        s.compute_dt()
        o.compute_dt()
        assert abs(s.dm.dt - o.dt) <= TOL * o.dt, (k, s.dm.dt, o.dt)
        o.captures = []
        o.perform_update()
        recs = s.perform_update_with_flags()
        assert len(recs) == len(o.captures)
        for i, (rg, ro) in enumerate(zip(recs, o.captures)):
            bad, ndiff = flag_mismatch(rg["troubles"], ro["troubles"], ro["margin"])
            assert bad == 0, (k, i, bad)
            n_trouble += int(np.sum(ro["troubles"]))
            assert float(np.nanmax(ro["theta"])) < 3.0 * rho0, ("unbounded candidate", k, i)
            if ndiff or not check_theta:
                continue
            assert relerr(rg["theta"], ro["theta"]) < TOL, (k, i)
            for a in range(s.ndim):
                d = "xyz"[a]
                assert relerr(rg[f"aff_{d}"], ro[f"aff_{d}"]) < TOL, (k, i, d)
        assert relerr_fields(s.dm.U_sp, o.U_sp) < TOL, (k, relerr_fields(s.dm.U_sp, o.U_sp))
    return n_trouble


@pytest.mark.parametrize("N,p,bc,seed", CASES,
                         ids=[f"{c[0][0]}x{c[0][1]}-p{c[1]}-{c[2]}" for c in CASES])
def test_quadrants_large_grid_vs_oracle(N, p, bc, seed):
    print(f"[gpu parity large] N={N} p={p} {bc} seed={seed}", flush=True)
    s, o = _pair(N, p, bc, seed)
    n_trouble = _lockstep(s, o, 3)
    ncell = N[0] * N[1] * (p + 1) ** 2
    # troubled cells exist, and most cells are not troubled (quiet rows/tiles)
    assert 0 < n_trouble < 0.5 * 4 * 3 * ncell, n_trouble


@pytest.mark.parametrize("form", [1, 2, 3])
@pytest.mark.parametrize("bc", ["gradfree", "reflective"])
def test_quadrants_other_fallback_forms_large_grid_vs_oracle(bc, form):
    """The fixed forms of the 2-D FV substep on the (256, 24)-element grid
    (the default adaptive form runs in the tests above): 1 cell tiles over the
    face-flux arrays in HBM, 2 column strips over them, 3 the fallback on the
    compacted list of active tiles (tiles of 8 x 2 elements: 32 columns, so
    the list holds tiles with and without active faces on this grid)."""
    from spd_amd import _lib
    print(f"[gpu parity large, fallback form {form}] {bc}", flush=True)
    s, o = _pair((256, 24), 3, bc, 15 if bc == "gradfree" else 16)
    _lib.check(s._lib.spd_set_launch(s._ctx, _lib.LAUNCH["fb_form"], form), "spd_set_launch")
    _lockstep(s, o, 2)


@pytest.mark.parametrize("bc", ["gradfree", "reflective"])
def test_quadrants_device_loop_large_grid(bc):
    """The device-side loop (graph replays) over the same grid gives the host
    loop's state bit for bit."""
    print(f"[gpu device loop large] {bc}", flush=True)
    a, _ = _pair((256, 24), 3, bc, 15)
    b, _ = _pair((256, 24), 3, bc, 15)
    b.device_loop, b.graph_steps = True, 2
    a.device_loop = False
    a.perform_iterations(4)
    b.perform_iterations(4)
    assert a.n_step == b.n_step == 4
    assert np.array_equal(a.dm.U_sp, b.dm.U_sp)


def test_taylor_green_3d_multiblock_vs_oracle():
    """3-D Taylor-Green at 8^3 elements, p = 3, with the density perturbation
    (amplitude 1e-2): several blocks of every 3-D kernel per dimension, two
    steps against the oracle (flags, theta and state)."""
    import math
    from oracle.spd_oracle import OracleSimulator
    from spd_amd import initial_conditions as ic
    from spd_amd.simulator import SDADER_Simulator
    print("[gpu parity 3-D Taylor-Green 8^3 p=3]", flush=True)
    L = 2 * math.pi
    kw = dict(p=3, N=(8, 8, 8), BC=(("periodic", "periodic"),) * 3, xlim=(0, L), ylim=(0, L),
              zlim=(0, L), update="FV", FB=True, riemann_solver_sd="hllc",
              riemann_solver_fv="hllc")
    noise = ic.subcell_noise((8, 8, 8), 3, [0], 1e-2, seed=0)
    s = SDADER_Simulator(init_fct=ic.taylor_green, device="cuda:0", verbose=False, **kw)
    s.perturb_initial_condition(noise)
    o = OracleSimulator(ic.taylor_green, noise=noise, **kw)
    assert relerr_fields(s.dm.U_sp, o.U_sp) < TOL
    n_trouble = _lockstep(s, o, 2)
    assert n_trouble > 0


@pytest.mark.parametrize("update", ["FV", "SD"])
@pytest.mark.parametrize("p", [4, 5])
def test_taylor_green_3d_high_orders_vs_oracle(p, update):
    """3-D Taylor-Green at p = 4 and 5 on a (4, 3, 2)-element grid (unequal
    sides, so a transposed index shows), density perturbation 1e-2: two steps
    against the oracle (flags, theta and state with the fallback; state with
    the SD update)."""
    import math
    from oracle.spd_oracle import OracleSimulator
    from spd_amd import initial_conditions as ic
    from spd_amd.simulator import SDADER_Simulator
    print(f"[gpu parity 3-D Taylor-Green (4,3,2) p={p} {update}]", flush=True)
    L = 2 * math.pi
    N = (4, 3, 2)
    kw = dict(p=p, N=N, BC=(("periodic", "periodic"),) * 3, xlim=(0, L), ylim=(0, L),
              zlim=(0, L), update=update, FB=(update == "FV"), riemann_solver_sd="hllc",
              riemann_solver_fv="hllc")
    noise = ic.subcell_noise(N, p, [0], 1e-2, seed=1)
    s = SDADER_Simulator(init_fct=ic.taylor_green, device="cuda:0", verbose=False, **kw)
    s.perturb_initial_condition(noise)
    o = OracleSimulator(ic.taylor_green, noise=noise, **kw)
    assert relerr_fields(s.dm.U_sp, o.U_sp) < TOL
    if update == "FV":
        assert _lockstep(s, o, 2) > 0
    else:
        for k in range(2):
            s.compute_dt()
            o.compute_dt()
            assert abs(s.dm.dt - o.dt) <= TOL * o.dt, (k, s.dm.dt, o.dt)
            s.perform_update()
            o.perform_update()
            assert relerr_fields(s.dm.U_sp, o.U_sp) < TOL, (k, relerr_fields(s.dm.U_sp, o.U_sp))
