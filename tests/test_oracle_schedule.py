"""The oracle's runBand pass schedule against closed forms that tell its branches apart (CPU).

Each case is checked under the schedule it pins, and each closed form is shown to reject the other schedule:
the conserving schedule differs from the reference one only in the copy condition at s = 0, so it stands in
for the plausible mistake of getting that condition wrong."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from tests import schedule_cases as sc

TOL = 0.003  # relative; ray origins are not stratified: ~1e-3 Monte Carlo error at 10^6 rays per face


def _rel(a, b):
    return abs(a - b) / abs(b)


@pytest.fixture(scope="module")
def squares():
    T, ty, tw, _ = sc.squares_scene()
    return OracleScene(T, ty, tw)


@pytest.mark.parametrize("schedule", [0, 1])
def test_common_edge_squares_shortwave_with_sky(squares, schedule):
    R = squares.run_band(sc.squares_sw(schedule))
    e0, e1 = sc.squares_sw_exact(schedule)
    assert _rel(R[0, 0], e0) <= TOL, (R[0, 0], e0)
    assert _rel(R[1, 0], e1) <= TOL, (R[1, 0], e1)
    # the other schedule's closed form for the receiving patch is far outside the tolerance
    _, other = sc.squares_sw_exact(1 - schedule)
    assert _rel(R[1, 0], other) > 20 * TOL


def test_reference_schedule_gathers_reflected_sun_twice(squares):
    # depth 1, no sky: only the first scattering pass runs, and it gathers patch 0's reflected sun once;
    # adding the sky turns the primary pass on, which gathers it again (the reference's s = 0 branch)
    F = sc.squares_scene()[3]
    req = sc.squares_sw(0)
    req.dflux[:] = 0.0
    R = squares.run_band(req)
    assert _rel(R[1, 0], F * sc.RHO0 * sc.Q) <= TOL
    Rs = squares.run_band(sc.squares_sw(0))
    assert _rel(Rs[1, 0] - (1 - F) * sc.QD, 2 * F * sc.RHO0 * sc.Q) <= TOL


@pytest.mark.parametrize("schedule", [0, 1])
def test_common_edge_squares_longwave_and_shortwave(squares, schedule):
    R = squares.run_band(sc.squares_lw_sw(schedule))
    ex = sc.squares_lw_sw_exact()
    assert R[0, 0] == 0.0  # patch 1 is at 0 K and emits nothing
    for i, b in ((0, 1), (1, 0), (1, 1)):
        assert _rel(R[i, b], ex[i, b]) <= TOL, (i, b, R[i, b], ex[i, b])


@pytest.mark.parametrize("schedule", [0, 1])
def test_gray_parallel_rectangles(schedule):
    (T, ty, tw), req, exact = sc.gray_rectangles(schedule)
    R = OracleScene(T, ty, tw).run_band(req)
    # five scattering iterations truncate the geometric series at (rho F)^6 ~ 2e-6 of the answer
    assert np.all(np.abs(R[:, 0] - exact) / exact <= TOL), (R[:, 0], exact)
