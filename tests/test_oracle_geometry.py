"""Oracle ray casting: reference golden cases, BVH == brute force, periodic wrap (CPU)."""
import numpy as np
import pytest

from helios_amd import scenes
from oracle.oracle import OracleScene
from tests import cd_cases


@pytest.mark.parametrize("case", cd_cases.load()["cases"], ids=lambda c: c["name"])
def test_cd_golden_case(case):
    T, ty, tw = cd_cases.case_scene(case)
    o = OracleScene(T, ty, tw)
    t, n, u = o.cast_rays(np.array([case["origin"]], np.float32), np.array([case["dir"]], np.float32),
                          case["max_dist"])
    cd_cases.check_case(case, float(t[0]), n[0].astype(np.float64), int(u[0]))


def test_cd_golden_grid():
    T, ty, tw, O, D, expect = cd_cases.grid_scene()
    o = OracleScene(T, ty, tw)
    t, n, u = o.cast_rays(O, D)
    assert np.array_equal(u != 0xFFFFFFFF, expect)


@pytest.mark.parametrize("brute_seed", [1, 2])
def test_castrays_bvh_equals_brute_force(brute_seed):
    s = scenes.random_canopy(brute_seed, 3000)
    o = OracleScene.from_arrays(s)
    O, D = scenes.rays(brute_seed, 20000, [-2, -2, 0], [2, 2, 2.2])
    tb, nb, ub = o.cast_rays(O, D, brute=False)
    tf, nf, uf = o.cast_rays(O, D, brute=True)
    assert np.array_equal(ub, uf)
    h = ub != 0xFFFFFFFF
    assert h.sum() > 500
    assert np.array_equal(tb[h], tf[h])


def test_radiation_trace_bvh_equals_brute_force():
    s = scenes.config1(3)
    o = OracleScene.from_arrays(s)
    rng = np.random.default_rng(0)
    bad = 0
    for k in range(3000):
        i = int(rng.integers(0, s.P))
        org, d = o.diffuse_ray(i, k % 10, (k // 10) % 10, 10, 1, 77)
        a = o.trace(org, d, 1e-5, i, brute=False)
        b = o.trace(org, d, 1e-5, i, brute=True)
        bad += a[0] != b[0]
    assert bad == 0


def test_periodic_faces_and_wrap():
    s = scenes.patch_canopy(4, 200, 0.1, 2.0, 1.0, 4, periodic=True)
    o = OracleScene.from_arrays(s)
    v, n, fr, dom, nb = o.geometry()
    assert nb == 4
    # domain = vertex bounds +- 1e-5 (R/src/RadiationModel.cpp:6046-6053)
    xs = v[:, :, 0][s.types == 0]
    assert abs(dom[0] - (xs.min() - 1e-5)) < 1e-6 and abs(dom[1] - (xs.max() + 1e-5)) < 1e-6
    # a horizontal ray above the canopy leaves through +x and re-enters at -x: after 32 wraps it
    # reports the last traversal result (a miss here)
    h = o.trace_periodic([0.0, 0.0, 1.9], [1.0, 0.0, 0.0], 1e-5, 0xFFFFFFFF)
    assert h == 0xFFFFFFFF


def test_diffuse_direction_is_cosine_distributed():
    # diffuse_raygen.comp:461-482: theta = asin(sqrt(Rt)) -> cos(theta) is distributed as sqrt(1-U):
    # E[cos] = 2/3, E[cos^2] = 1/2 on a horizontal patch
    T = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]], np.float32)
    o = OracleScene(T, [0], [0])
    cz = np.array([o.diffuse_ray(0, r % 40, r // 40, 40, 1, 5)[1][2] for r in range(1600)], np.float64)
    assert np.all(cz > 0)
    assert abs(cz.mean() - 2.0 / 3.0) < 2e-3
    assert abs((cz ** 2).mean() - 0.5) < 2e-3
