"""FP64 energy bookkeeping of the scattering iterations (rt_set_energy_accounting / rt_energy_residual_fp64 /
rt_energy_passes_fp64) through the C ABI.

Pinned by closed forms: on the 90-degree common-edge squares the conserving schedule's residual is zero and
the reference schedule's is -rho Q + rho (1 - F) Qd (patch 0's reflected sun handed on twice, its reflected sky
dropped: RadiationModel.cpp:4286-4289); in a closed enclosure nothing escapes and the emitted flux is absorbed.
The bookkeeping never changes the fluxes."""
import numpy as np
import pytest

from helios_amd import native, scenes
from tests import closed_forms as cf
from tests import schedule_cases as sc
from tests.helpers import config1_request

pytestmark = pytest.mark.gpu


@pytest.fixture()
def acct(ctx):
    native.set_energy_accounting(ctx, True)
    yield ctx
    native.set_energy_accounting(ctx, False)


def _inputs(e):
    return e["source"] + e["sky"] + e["emitted"]


@pytest.mark.parametrize("schedule", [0, 1])
def test_common_edge_squares_residual(acct, schedule):
    T, ty, tw, F = sc.squares_scene()
    g = acct.scene(T, ty, tw)
    req = sc.squares_sw(schedule)
    R = native.run_band(g, req)
    e = native.energy_fp64(acct, 1)
    Q, Qd, rho = sc.Q, sc.QD, sc.RHO0
    # inputs: the sun on patch 0, the sky on both patches through 1 - F
    assert abs(e["source"][0] - Q) <= 1e-3 * Q
    assert abs(e["sky"][0] - 2 * (1 - F) * Qd) <= 3e-3 * Qd
    assert e["emitted"][0] == 0.0
    assert abs(e["absorbed"][0] - float(R[:, 0].sum())) <= 1e-4 * e["absorbed"][0]  # unit areas
    assert e["passes"].shape == (2, 4, 1)
    exact = -rho * Q + rho * (1 - F) * Qd if schedule == 0 else 0.0
    assert abs(e["residual"][0] - exact) <= 3e-3 * (rho * Q), (e["residual"][0], exact)
    # each pass: what the launching faces hand on arrives at a surface or escapes
    for rad, arr, _, esc in e["passes"][:, :, 0]:
        assert abs(rad - (arr + esc)) <= 3e-3 * rad
    # the reference schedule hands patch 0's reflected sun on in both passes
    if schedule == 0:
        assert np.allclose(e["passes"][:, 0, 0], rho * Q, rtol=1e-6)


def test_enclosure_conserves_energy(acct):
    T, ty, tw = cf.closed_box()
    P = T.shape[0]
    g = acct.scene(T, ty, tw)
    req = native.BandRequest(P, [native.BandSpec("LW", 5, True, 0.0)], rho=np.full((P, 1), 0.2),
                             tau=np.zeros((P, 1)), direct_rays=0, diffuse_rays=10000, eps=np.full((P, 1), 0.8),
                             temperature=np.full(P, 300.0), diffuse_seeds=[3, 4, 5, 6, 7, 8])
    native.run_band(g, req)
    e = native.energy_fp64(acct, 1)
    A = 4.0  # each of the 6 x 25 patches is 2 x 2 m
    assert abs(e["emitted"][0] - P * A * float(np.float32(0.8)) * cf.SIGMA * 300.0 ** 4) <= 1e-9 * e["emitted"][0]
    assert e["source"][0] == 0.0 and e["sky"][0] == 0.0
    assert e["escaped"][0] <= 1e-4 * e["emitted"][0]  # only rays through the seams of the tiling
    assert abs(e["residual"][0]) <= 3e-3 * e["emitted"][0], e["residual"][0] / e["emitted"][0]
    assert e["passes"].shape[0] == 6


def test_one_sided_back_face_takes_nothing(acct):
    # an emitting one-sided rectangle under a one-sided rectangle whose back faces it: the upper one launches
    # no bottom-face rays (diffuse_raygen.comp:365-370), so what the lower one radiates at it is absorbed by
    # nothing and does not escape -- the residual is exactly that share, F12 of the emission
    T0, _, _, F12 = cf.parallel_rectangles()
    T = np.stack([T0[0], cf.patch_T((0, 0, 0.5), (1.0, 0, 0), (0, 2.0, 0))])  # both face +z
    g = acct.scene(T, [0, 0], [0, 0])
    req = native.BandRequest(2, [native.BandSpec("LW", 0, True, 0.0)], rho=np.zeros((2, 1)), tau=np.zeros((2, 1)),
                             direct_rays=0, diffuse_rays=250000, eps=np.ones((2, 1)), temperature=[300.0, 0.0],
                             diffuse_seeds=[41])
    R = native.run_band(g, req)
    e = native.energy_fp64(acct, 1)
    E = cf.SIGMA * 300.0 ** 4 * 2.0
    assert abs(e["emitted"][0] - E) <= 1e-6 * E
    assert R[1, 0] == 0.0
    assert abs(e["escaped"][0] - (1.0 - F12) * E) <= 3e-3 * E
    assert abs(e["residual"][0] - F12 * E) <= 3e-3 * E, (e["residual"][0] / E, F12)


def test_open_canopy_residual_both_schedules(acct):
    s = scenes.config1(3)
    g = acct.scene(s.transforms, s.types, s.twosided)
    res = {}
    for schedule in (0, 1):
        req = config1_request(s.P, schedule=schedule, nrays=400)
        native.run_band(g, req)
        res[schedule] = native.energy_fp64(acct, 1)
    e0, e1 = res[0], res[1]
    # the same direct and primary diffuse passes under both schedules
    for k in ("source", "sky"):
        assert abs(e0[k][0] - e1[k][0]) <= 1e-9 * e1[k][0]
    assert np.allclose(e0["passes"][0], e1["passes"][0], rtol=1e-12, atol=0)
    # conserving: zero up to the Monte Carlo error of the view factors
    assert abs(e1["residual"][0]) <= 0.01 * _inputs(e1)[0], e1["residual"][0] / _inputs(e1)[0]
    # reference: the first scattering pass hands on the direct scatter again instead of the primary pass's
    # scatter, so the residual is (that scatter) - (the direct scatter), up to the same error
    expect = e1["passes"][1, 0, 0] - e0["passes"][1, 0, 0]
    assert e0["passes"][1, 0, 0] == pytest.approx(e0["passes"][0, 0, 0], rel=1e-12)
    assert abs(e0["residual"][0] - expect) <= 0.01 * _inputs(e0)[0], (e0["residual"][0], expect)
    assert abs(e0["residual"][0]) > 5 * abs(e1["residual"][0])


def test_bookkeeping_leaves_fluxes_unchanged(ctx):
    s = scenes.config1(5)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = config1_request(s.P, nrays=64)
    R0 = native.run_band(g, req)
    native.set_energy_accounting(ctx, True)
    try:
        R1 = native.run_band(g, req)
    finally:
        native.set_energy_accounting(ctx, False)
    assert np.array_equal(R0, R1)
    native.run_band(g, req)
    with pytest.raises(native.RTError, match="rt_set_energy_accounting"):
        native.energy_fp64(ctx, 1)
