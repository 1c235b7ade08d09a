"""The exact Riemann solver the Sod runs are measured against reproduces the
star states tabulated by Toro (Riemann Solvers and Numerical Methods for
Fluid Dynamics, 3rd ed., Table 4.2, tests 1-5) and the conservation of the
sampled solution."""
import numpy as np
import pytest

from exact_riemann import sample, star_state

TORO = [
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.30313, 0.92745),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.00189, 0.00000),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 460.894, 19.5975),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), 46.0950, -6.19633),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 1691.64, 8.68975),
]


@pytest.mark.parametrize("L,R,ps,us", TORO)
def test_star_state_matches_toro_table(L, R, ps, us):
    p, u = star_state(L, R)
    assert abs(p - ps) <= 5e-5 * max(1.0, abs(ps)) + 5e-6
    assert abs(u - us) <= 5e-5 * max(1.0, abs(us)) + 5e-6


def test_sod_sample_conserves_mass_and_states():
    L, R = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)
    x = (np.arange(200000) + 0.5) / 200000
    w = sample(x, 0.2, L, R)
    # mass in [0,1] is conserved while the waves stay inside the tube
    assert abs(w[0].mean() - 0.5 * (1.0 + 0.125)) < 1e-5
    assert w[0][0] == 1.0 and w[0][-1] == 0.125
    assert abs(w[2][100000 + 20000] - 0.30313) < 1e-4  # star pressure between the waves
