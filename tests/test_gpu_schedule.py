"""The HIP runBand schedule (through the C ABI) against the closed forms that tell the pass schedule's branches
apart, under both schedules, on the whole-solve entry point and on the call-by-call backend driven in the
reference's own order.  The same requests run on the oracle (tests/test_oracle_schedule.py), so the GPU is
checked both against the closed form (Monte Carlo tolerance) and against the oracle at the parity bound."""
import numpy as np
import pytest

from helios_amd import native
from oracle.oracle import OracleScene
from tests import schedule_cases as sc
from tests.backend_schedule import run_band_via_backend
from tests.helpers import rel_err
from tests.test_oracle_schedule import TOL

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(a - b) / abs(b)


@pytest.fixture(scope="module")
def squares(ctx):
    T, ty, tw, _ = sc.squares_scene()
    return ctx.scene(T, ty, tw), OracleScene(T, ty, tw)


@pytest.mark.parametrize("schedule", [0, 1])
def test_common_edge_squares_shortwave_with_sky_gpu(squares, schedule):
    g, o = squares
    req = sc.squares_sw(schedule)
    R = native.run_band(g, req)
    e0, e1 = sc.squares_sw_exact(schedule)
    assert _rel(R[0, 0], e0) <= TOL, (R[0, 0], e0)
    assert _rel(R[1, 0], e1) <= TOL, (R[1, 0], e1)
    _, other = sc.squares_sw_exact(1 - schedule)
    assert _rel(R[1, 0], other) > 20 * TOL
    assert rel_err(R, o.run_band(req), floor=1.0).max() <= 1e-4


@pytest.mark.parametrize("schedule", [0, 1])
def test_common_edge_squares_longwave_and_shortwave_gpu(squares, schedule):
    g, o = squares
    req = sc.squares_lw_sw(schedule)
    R = native.run_band(g, req)
    ex = sc.squares_lw_sw_exact()
    assert R[0, 0] == 0.0
    for i, b in ((0, 1), (1, 0), (1, 1)):
        assert _rel(R[i, b], ex[i, b]) <= TOL, (i, b, R[i, b], ex[i, b])
    assert rel_err(R, o.run_band(req), floor=1.0).max() <= 1e-4


@pytest.mark.parametrize("schedule", [0, 1])
def test_gray_parallel_rectangles_gpu(ctx, schedule):
    (T, ty, tw), req, exact = sc.gray_rectangles(schedule)
    R = native.run_band(ctx.scene(T, ty, tw), req)
    assert np.all(np.abs(R[:, 0] - exact) / exact <= TOL), (R[:, 0], exact)
    assert rel_err(R, OracleScene(T, ty, tw).run_band(req), floor=1.0).max() <= 1e-4


@pytest.mark.parametrize("case", ["sw", "lw_sw"])
def test_common_edge_squares_call_by_call(ctx, case):
    # the reference's pass loop (reference schedule) over the RayTracingBackend-shaped entry points
    T, ty, tw, _ = sc.squares_scene()
    req = sc.squares_sw(0) if case == "sw" else sc.squares_lw_sw(0)
    be = native.Backend(ctx)
    be.update_geometry(T, ty, tw)
    be.build_acceleration_structure()
    be.update_materials(req.rho, req.tau)
    be.update_sources(req.sdir)
    R = run_band_via_backend(be, req, tw)
    if case == "sw":
        e0, e1 = sc.squares_sw_exact(0)
        assert _rel(R[0, 0], e0) <= TOL and _rel(R[1, 0], e1) <= TOL, (R[:, 0], e0, e1)
    else:
        ex = sc.squares_lw_sw_exact()
        for i, b in ((0, 1), (1, 0), (1, 1)):
            assert _rel(R[i, b], ex[i, b]) <= TOL, (i, b, R[i, b], ex[i, b])
    assert rel_err(R, native.run_band(ctx.scene(T, ty, tw), req), floor=1.0).max() <= 1e-5
