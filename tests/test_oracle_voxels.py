"""Oracle castRays with voxels (type 4): the reference's voxel cases, an independent FP64 box test, face
normals, the traversal against brute force, and voxels left out of the radiation traversal (CPU)."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from tests import voxel_cases as vc
from tests.closed_forms import patch_T, voxel_box, voxel_T


@pytest.mark.parametrize("case", vc.KAT, ids=lambda c: c[0])
def test_reference_voxel_cases(case):
    name, boxes, o, d, maxd, hit, dist, normal = case
    T, ty, tw = vc.voxel_scene(boxes)
    t, n, u = OracleScene(T, ty, tw).cast_rays(np.array([o], np.float32), np.array([d], np.float32), maxd)
    if hit is None:
        assert u[0] == vc.MISS, name
        return
    assert u[0] == hit, name
    if dist is not None:
        assert dist[0] < t[0] < dist[1], (name, t[0])
    if normal is not None:
        assert np.array_equal(n[0], np.array(normal, np.float32)), (name, n[0])
    p = np.array(o, np.float64) + np.array(d, np.float64) * float(t[0])
    lo, hi = voxel_box(T[hit])
    # the hit point lies on the box's surface
    assert np.all(p >= lo - 1e-6) and np.all(p <= hi + 1e-6), (name, p)
    assert np.min(np.minimum(np.abs(p - lo), np.abs(p - hi))) < 1e-6, (name, p)


def test_reference_voxel_grid():
    boxes, O, D, expect = vc.grid_case()
    T, ty, tw = vc.voxel_scene(boxes)
    t, n, u = OracleScene(T, ty, tw).cast_rays(O, D)
    assert np.array_equal(u != vc.MISS, expect)


def test_reference_consistency_fan_against_fp64():
    boxes, O, D, maxd = vc.consistency_rays()
    T, ty, tw = vc.voxel_scene(boxes)
    t, n, u = OracleScene(T, ty, tw).cast_rays(O, D, maxd)
    lo, hi = zip(*[voxel_box(x) for x in T])
    nhit = 0
    for r in range(len(O)):
        k, tt, nn = vc.slab_fp64(lo, hi, O[r], D[r], maxd)
        assert (u[r] != vc.MISS) == (k is not None), r
        if k is not None:
            nhit += 1
            assert u[r] == k and abs(t[r] - tt) <= 1e-5 * tt, (r, t[r], tt)
            p = O[r].astype(np.float64) + D[r].astype(np.float64) * tt
            rel = np.sort(np.abs(p - (lo[k] + hi[k]) / 2) / ((hi[k] - lo[k]) / 2))
            if rel[2] - rel[1] > 1e-4:  # not on an edge, where FP32 rounding picks the face
                assert np.array_equal(n[r], nn.astype(np.float32)), (r, n[r], nn)
    assert nhit >= 2


@pytest.mark.parametrize("seed", [1, 2])
def test_random_voxels_against_fp64(seed):
    """Rotated and axis-aligned voxels, random rays, rays from inside: hit, index, distance and face normal
    equal the FP64 restatement wherever the FP64 answer is not within rounding of an edge or a tie."""
    rng = np.random.default_rng(seed)
    boxes = [(tuple(rng.uniform(-1, 1, 3)), tuple(rng.uniform(0.1, 0.6, 3)), float(rng.uniform(0, 3)) * (k % 2))
             for k in range(40)]
    T, ty, tw = vc.voxel_scene(boxes)
    O = rng.uniform(-1.5, 1.5, (3000, 3)).astype(np.float32)
    D = rng.normal(size=(3000, 3))
    D = (D / np.linalg.norm(D, axis=1, keepdims=True)).astype(np.float32)
    t, n, u = OracleScene(T, ty, tw).cast_rays(O, D)
    lo, hi = zip(*[voxel_box(x) for x in T])
    checked = 0
    for r in range(len(O)):
        k, tt, nn = vc.slab_fp64(lo, hi, O[r], D[r])
        if k is None:
            if u[r] != vc.MISS:
                # a grazing FP32 hit: the point is within rounding of the box
                p = O[r].astype(np.float64) + D[r].astype(np.float64) * float(t[r])
                L, H = voxel_box(T[u[r]])
                assert np.all(p >= L - 1e-5) and np.all(p <= H + 1e-5), r
            continue
        if u[r] != k:
            # FP64 tie or edge: the FP32 answer is at the same distance within rounding
            assert u[r] != vc.MISS and abs(float(t[r]) - tt) <= 1e-5 * max(tt, 1.0), (r, u[r], k)
            continue
        checked += 1
        assert abs(float(t[r]) - tt) <= 1e-5 * max(tt, 1.0), (r, t[r], tt)
        p = O[r].astype(np.float64) + D[r].astype(np.float64) * tt
        c, e = (lo[k] + hi[k]) / 2, (hi[k] - lo[k]) / 2
        rel = np.sort(np.abs(p - c) / e)
        if rel[2] - rel[1] > 1e-4:  # not at an edge of the box
            assert np.array_equal(n[r], nn.astype(np.float32)), (r, n[r], nn)
    assert checked > 500


def test_face_normals_are_outward_not_turned():
    """Entering through each face gives that face's outward normal (towards the origin); from inside, the exit
    face's outward normal (away from the origin) -- castRays does not turn voxel normals."""
    T, ty, tw = vc.voxel_scene([((0, 0, 0), (2, 2, 2))])
    o = OracleScene(T, ty, tw)
    for a in range(3):
        for s in (-1.0, 1.0):
            d = np.zeros(3, np.float32)
            d[a] = -s
            org = np.zeros(3, np.float32)
            org[a] = 3 * s
            org[(a + 1) % 3] = 0.3
            t, n, u = o.cast_rays(org[None], d[None])
            want = np.zeros(3, np.float32)
            want[a] = s
            assert u[0] == 0 and t[0] == np.float32(2.0) and np.array_equal(n[0], want)
            # from inside, moving towards the same face: its outward normal now points along the ray
            t, n, u = o.cast_rays(np.array([[0.2, -0.1, 0.3]], np.float32), -d[None])
            assert u[0] == 0 and np.array_equal(n[0], want)


def test_origin_on_a_face():
    """An origin on the lower face moving in hits the far face (the entry is below 1e-6); moving out, nothing."""
    T, ty, tw = vc.voxel_scene([((0, 0, 0), (2, 2, 2))])
    o = OracleScene(T, ty, tw)
    t, n, u = o.cast_rays(np.array([[0.1, 0.2, -1.0]], np.float32), np.array([[0, 0, 1]], np.float32))
    assert u[0] == 0 and t[0] == np.float32(2.0) and np.array_equal(n[0], np.array([0, 0, 1], np.float32))
    t, n, u = o.cast_rays(np.array([[0.1, 0.2, -1.0]], np.float32), np.array([[0, 0, -1]], np.float32))
    assert u[0] == vc.MISS


def test_max_distance_excludes_the_far_voxel():
    T, ty, tw = vc.voxel_scene([((0, 0, 0), (2, 2, 2))])
    o = OracleScene(T, ty, tw)
    O, D = np.array([[0, 0, -5]], np.float32), np.array([[0, 0, 1]], np.float32)
    assert o.cast_rays(O, D, 4.0)[2][0] == vc.MISS  # distance 4 is not below 4
    assert o.cast_rays(O, D, 4.001)[2][0] == 0


@pytest.mark.parametrize("seed", [3, 4])
def test_mixed_scene_bvh_equals_brute_force(seed):
    T, ty, tw = vc.mixed_scene(seed, 200, 2000)
    o = OracleScene(T, ty, tw)
    O, D = vc.mixed_rays(seed, T, ty, 12000)
    tb, nb, ub = o.cast_rays(O, D, brute=False)
    tf, nf, uf = o.cast_rays(O, D, brute=True)
    assert np.array_equal(ub, uf)
    h = ub != vc.MISS
    assert np.array_equal(tb[h], tf[h]) and np.array_equal(nb[h], nf[h])
    assert np.isin(ub[h], np.nonzero(ty == 4)[0]).sum() > 300  # voxels are hit (uuid = index here)


def test_radiation_trace_passes_through_voxels():
    """The radiation traversal never hits a voxel: a ray through one reaches the patch behind it."""
    T = np.stack([voxel_T((0, 0, 1), (1, 1, 1)), patch_T((0, 0, 0), (2, 0, 0), (0, 2, 0))])
    o = OracleScene(T, np.array([4, 0], np.uint32), np.zeros(2, np.uint32))
    h, t = o.trace(np.array([0.1, 0.1, 3.0], np.float32), np.array([0, 0, -1], np.float32), 1e-5)
    assert h == 1 and abs(t - 3.0) < 1e-6
    h, t = o.trace(np.array([0.1, 0.1, 3.0], np.float32), np.array([0, 0, -1], np.float32), 1e-5, brute=True)
    assert h == 1
    tc, nc, uc = o.cast_rays(np.array([[0.1, 0.1, 3.0]], np.float32), np.array([[0, 0, -1]], np.float32))
    assert uc[0] == 0 and tc[0] == np.float32(1.5)
