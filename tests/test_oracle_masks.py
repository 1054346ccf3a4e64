"""The oracle's texture masks (texture_mask.glsl, direct_raygen.comp:431-473, diffuse_raygen.comp:411-455)
against closed forms (CPU).

Texel lookup: a patch hit at parametric (u, v) reads column floor((W-1) u) and row floor((H-1)(1-v)) (row 0
is the top of the image); a triangle hit at barycentric (u, v) reads round((W-1)|s|), round((H-1)|1-t|) of
its interpolated UV (s, t).  A shadow ray from a small receiver straight below an occluder hits it at a
known (u, v), so whether the receiver is lit names exactly one texel.  Origin redraws: a direct ray whose
stratified origin falls on a transparent texel redraws a uniform point up to 10 times, so with half the
patch transparent a fraction 2^-10 of those rays keeps a transparent origin."""
import math

import numpy as np

from helios_amd.native import Camera
from oracle.oracle import OracleScene
from tests import closed_forms as cf

DOWN = (0.0, 0.0, 1.0)  # sun straight above


def _receiver_and_occluder(occ_T, occ_type):
    rec = cf.patch_T((0, 0, 0), (0.01, 0, 0), (0, 0.01, 0))
    return OracleScene(np.stack([rec, occ_T]), [0, occ_type], [0, 1])


def _lit(o):
    s = o.direct_sums(5, 3, [DOWN])
    return s[0, 0, 0] / (1.0 / 1.0)  # sum of strengths: 1 when every ray reaches the sun


MASK = np.array([[1, 0, 0],     # row 0 (top: v > 1/2)
                 [0, 1, 0],     # row 1 (v < 1/2)
                 [0, 0, 0]], np.uint8)


def test_patch_texel_lookup():
    for (u, v) in [(0.25, 0.25), (0.25, 0.75), (0.75, 0.25), (0.75, 0.75)]:
        # occluder 2 x 2 at z = 1 placed so the receiver (0, 0) projects to (u, v)
        cx, cy = 1.0 - 2.0 * u, 1.0 - 2.0 * v
        o = _receiver_and_occluder(cf.patch_T((cx, cy, 1.0), (2, 0, 0), (0, 2, 0)), 0)
        assert math.isclose(_lit(o), 0.0, abs_tol=1e-12)  # unmasked: shadowed
        o.set_masks([MASK], [-1, 0])
        opaque = MASK[int(math.floor(2 * (1 - v))), int(math.floor(2 * u))] != 0
        assert math.isclose(_lit(o), 0.0 if opaque else 1.0, rel_tol=1e-6, abs_tol=1e-12), (u, v)


def test_patch_custom_uv_lookup():
    """uv_id >= 0: u' = uv0.x + u (uv1.x - uv0.x), v' = uv0.y + v (uv2.y - uv1.y)."""
    uv = np.zeros((2, 4, 2), np.float32)
    uv[1] = [[0.5, 0.0], [1.0, 0.0], [1.0, 1.0], [0.5, 1.0]]  # the right half of the texture
    for (u, v) in [(0.25, 0.25), (0.25, 0.75), (0.75, 0.25), (0.75, 0.75)]:
        cx, cy = 1.0 - 2.0 * u, 1.0 - 2.0 * v
        o = _receiver_and_occluder(cf.patch_T((cx, cy, 1.0), (2, 0, 0), (0, 2, 0)), 0)
        mask = np.array([[0, 0, 1, 1, 0], [1, 1, 0, 0, 0]], np.uint8)  # W = 5, H = 2
        o.set_masks([mask], [-1, 0], uv=uv, uv_ids=[-1, 1])
        uu = 0.5 + u * 0.5
        col, row = int(math.floor(4 * uu)), int(math.floor(1 * (1 - 0.0 - v * 1.0)))
        opaque = mask[row, col] != 0
        assert math.isclose(_lit(o), 0.0 if opaque else 1.0, rel_tol=1e-6, abs_tol=1e-12), (u, v)


def test_triangle_texel_lookup():
    """Triangle v0 + u e1 + v e2 with UVs (0,0), (1,0), (0,1): texture (s, t) = (u, v), texel
    (round(2|u|), round(2|1-v|)) of a 3 x 3 mask."""
    mask = np.array([[0, 1, 0], [1, 0, 1], [0, 1, 1]], np.uint8)
    uv = np.zeros((2, 4, 2), np.float32)
    uv[1, :3] = [[0, 0], [1, 0], [0, 1]]
    for (u, v) in [(0.1, 0.1), (0.5, 0.1), (0.1, 0.5), (0.3, 0.3), (0.6, 0.3), (0.15, 0.8)]:
        x0, y0 = -2.0 * u, -2.0 * v
        T = cf.tri_T((x0, y0, 1.0), (x0 + 2, y0, 1.0), (x0, y0 + 2, 1.0))
        o = _receiver_and_occluder(T, 1)
        assert math.isclose(_lit(o), 0.0, abs_tol=1e-12)
        o.set_masks([mask], [-1, 0], uv=uv)
        opaque = mask[int(round(2 * abs(1 - v))), int(round(2 * abs(u)))] != 0
        assert math.isclose(_lit(o), 0.0 if opaque else 1.0, rel_tol=1e-6, abs_tol=1e-12), (u, v)


def test_all_opaque_patch_masks_change_nothing():
    """Every patch masked with an all-opaque texture: no redraw, no pass-through -> identical sums."""
    rng = np.random.default_rng(3)
    T = [cf.patch_T(rng.uniform(-1, 1, 3) * [1, 1, 0.5] + [0, 0, 0.6], (0.3, 0.05, 0.0), (0.0, 0.25, 0.1))
         for _ in range(60)]
    o = OracleScene(np.stack(T), [0] * 60, [1] * 60)
    a = o.direct_sums(6, 5, [[0.3, 0.2, 0.9]])
    h = o.diffuse_hits(4, 9)
    o.set_masks([np.ones((4, 5), np.uint8)], [0] * 60)
    assert np.array_equal(o.direct_sums(6, 5, [[0.3, 0.2, 0.9]]), a)
    assert np.array_equal(o.diffuse_hits(4, 9), h)


def test_transparent_patches_are_invisible_to_rays():
    """A fully transparent occluder passes every shadow ray; diffuse rays from it are all dropped while its
    direct rays keep their (transparent) origins."""
    o = _receiver_and_occluder(cf.patch_T((0, 0, 1.0), (2, 0, 0), (0, 2, 0)), 0)
    o.set_masks([np.zeros((2, 2), np.uint8)], [-1, 0])
    assert math.isclose(_lit(o), 1.0, rel_tol=1e-6)
    h = o.diffuse_hits(3, 1)
    assert np.all(h[1] == 0xFFFFFFFE)
    assert np.isfinite(o.direct_ray(1, 0, 0, 3, 1)).all()


def test_origin_redraws_on_patches():
    """Receiver 1 x 1 whose left half (u < 1/2) is transparent, the left half shadowed by an occluder: the
    rays that start on the left redraw until they land on the right (up to 10 times), so the lit fraction
    is 1 - 2^-11 instead of 1/2."""
    rec = cf.patch_T((0, 0, 0), (1, 0, 0), (0, 1, 0))
    occ = cf.patch_T((-0.25 - 0.01, 0, 1.0), (0.5, 0, 0), (0, 2, 0))  # covers x < -0.01
    o = OracleScene(np.stack([rec, occ]), [0, 0], [0, 1])
    d = 64
    base = o.direct_sums(d, 7, [DOWN])[0, 0, 0]
    assert abs(base - 0.5) < 0.02
    o.set_masks([np.array([[0, 1, 1]], np.uint8)], [0, -1])
    lit = o.direct_sums(d, 7, [DOWN])[0, 0, 0]
    assert abs(lit - (1 - 2.0 ** -11)) < 2e-3, lit


def test_triangle_origins_avoid_transparent_texels():
    """A masked triangle draws its origins by rejection: every direct origin lies on an opaque texel
    (floor lookup), and a fully transparent triangle drops every ray."""
    T = cf.tri_T((0, 0, 0), (1, 0, 0), (0, 1, 0))
    o = OracleScene(T[None], [1], [1])
    uv = np.zeros((1, 4, 2), np.float32)
    uv[0, :3] = [[0, 0], [1, 0], [0, 1]]
    mask = np.array([[1, 0, 0, 0, 0], [0, 0, 0, 0, 0], [0, 0, 1, 0, 0], [0, 0, 0, 0, 0], [1, 1, 0, 0, 0]], np.uint8)
    o.set_masks([mask], [0], uv=uv)
    n_ok = 0
    for r in range(49):
        p = o.direct_ray(0, r % 7, r // 7, 7, 11)
        if not np.isfinite(p).all():
            continue
        n_ok += 1
        # world = v0 + beta e1 + gamma e2 with e1 = (1,0,0), e2 = (0,1,0): (s, t) = (beta, gamma)
        s, t = float(p[0]), float(p[1])
        col = min(max(int(math.floor(4 * s)), 0), 4)
        row = min(max(int(math.floor(4 * (1 - t))), 0), 4)
        assert mask[row, col] != 0, (s, t)
    assert n_ok > 40
    o.set_masks([np.zeros((2, 2), np.uint8)], [0], uv=uv)
    assert all(not np.isfinite(o.direct_ray(0, r % 7, r // 7, 7, 11)).any() for r in range(49))


def test_camera_sees_through_transparent_texels():
    o = OracleScene(cf.patch_T((0, 0, 0), (4, 0, 0), (0, 4, 0))[None], [0], [0], uuids=[7])
    cam = Camera((0.0, 0.0, 2.0), (math.pi, 0.0), (8, 8), viewplane_length=1.0)
    lab, _ = o.pixel_labels(cam)
    assert np.all(lab == 8)
    o.set_masks([np.array([[1, 0], [0, 0]], np.uint8)], [0])  # only texel (0, 0) matters: col floor(u), row floor(1-v)
    lab, _ = o.pixel_labels(cam)
    assert np.all(lab == 8)
    o.set_masks([np.array([[0, 1], [1, 1]], np.uint8)], [0])
    lab, _ = o.pixel_labels(cam)
    assert np.all(lab == 0)
