"""Watertight predicates in the oracle (CPU): closed edges where the reference's predicates leak, the same
decisions as the reference away from edges, BVH = brute force, and the closed-form radiation cases unchanged."""
import numpy as np
import pytest

from helios_amd import native, scenes
from oracle.oracle import OracleScene, predicates
from tests import closed_forms as cf
from tests.test_oracle_physics import req1


def _down_rays(xs, ys, z=1.0):
    O = np.stack([xs, ys, np.full_like(xs, z)], 1).astype(np.float32)
    D = np.tile(np.array([0.0, 0.0, -1.0], np.float32), (len(xs), 1))
    return O, D


def _trace_all(o, O, D, brute=False):
    hits = []
    for r in range(len(O)):
        h, t = o.trace(O[r], D[r], tmin=1e-5, brute=brute)
        hits.append((h, t))
    return hits


def test_shared_patch_edge_leaks_only_with_reference_predicates():
    # two unit patches sharing the edge x = 0; rays straight down exactly on that edge
    T = np.stack([cf.patch_T((-0.5, 0, 0), (1, 0, 0), (0, 1, 0)), cf.patch_T((0.5, 0, 0), (1, 0, 0), (0, 1, 0))])
    o = OracleScene(T, [0, 0], [1, 1])
    ys = np.linspace(-0.45, 0.45, 91).astype(np.float32)
    O, D = _down_rays(np.zeros_like(ys), ys)
    ref = _trace_all(o, O, D)
    assert all(h == 0xFFFFFFFF for h, _ in ref)  # the open interval 0 < d.a < |a|^2 excludes the edge on both
    with predicates("watertight"):
        wt = _trace_all(o, O, D)
        wt_brute = _trace_all(o, O, D, brute=True)
    assert all(h in (0, 1) and t == 1.0 for h, t in wt) and wt == wt_brute


def test_shared_triangle_diagonal_is_closed():
    # a unit square as two triangles sharing the diagonal y = x; rays down through points of the diagonal
    tris = [cf.tri_T((0, 0, 0), (1, 0, 0), (1, 1, 0)), cf.tri_T((0, 0, 0), (1, 1, 0), (0, 1, 0))]
    o = OracleScene(np.stack(tris), [1, 1], [1, 1])
    xs = np.random.default_rng(3).uniform(0.01, 0.99, 4000).astype(np.float32)
    O, D = _down_rays(xs, xs)
    with predicates("watertight"):
        assert all(h != 0xFFFFFFFF for h, _ in _trace_all(o, O, D))


def test_watertight_agrees_with_reference_away_from_edges():
    s = scenes.random_canopy(8, 3000)
    o = OracleScene.from_arrays(s)
    O, D = scenes.rays(4, 20000, (-0.6, -0.6, 0.0), (0.6, 0.6, 2.0))
    _, _, u_ref = o.cast_rays(O, D)
    with predicates("watertight"):
        _, _, u_wt = o.cast_rays(O, D)
        _, _, u_wt_b = o.cast_rays(O, D, brute=True)
    assert np.array_equal(u_wt, u_wt_b)
    assert np.mean(u_wt != u_ref) <= 1e-3
    assert np.mean(u_ref != 0xFFFFFFFF) > 0.2


def test_closed_forms_hold_with_watertight_predicates():
    T = cf.patch_T((0, 0, 0), (1, 0, 0), (0, 1, 0))[None]
    Tr, ty, tw, F12 = cf.parallel_rectangles()
    with predicates("watertight"):
        R = OracleScene(T, [0], [0]).run_band(req1(1, "SW", [[0.0]], [[0.0]], d=100, sun=[0, 0, 1], F=1000.0))
        assert abs(R[0, 0] - 1000.0) < 1e-3
        R = OracleScene(Tr, ty, tw).run_band(req1(2, "SW", np.zeros((2, 1)), np.zeros((2, 1)), n=40000, dflux=1.0))
        assert np.max(np.abs(R[:, 0] - (1.0 - F12))) / (1.0 - F12) <= 0.005
