"""Voxel scenes and rays for castRays tests: the reference's own voxel cases
(plugins/collisiondetection/tests/selfTest.cpp:4370-4515, 5239-5380), an independent FP64 restatement of
the box test for voxel-only scenes, and a mixed scene of voxels, patches and triangles."""
from __future__ import annotations

import numpy as np

from helios_amd import scenes
from tests.closed_forms import voxel_T, voxel_box

MISS = 0xFFFFFFFF


def voxel_scene(boxes):
    """boxes: [(center, size[, rotation])] -> transforms, types, twosided."""
    T = np.stack([voxel_T(b[0], b[1], b[2] if len(b) > 2 else 0.0) for b in boxes])
    n = len(boxes)
    return T, np.full(n, 4, np.uint32), np.zeros(n, np.uint32)


# (name, voxels, origin, direction, max_dist, expected hit index or None, distance (lo, hi) or None,
#  outward face normal of the hit or None); every one from selfTest.cpp
KAT = [
    ("basic_center", [((0, 0, 0), (2, 2, 2))], (0, 0, -5), (0, 0, 1), 10.0, 0, (3.9, 4.1), (0, 0, -1)),  # :4381
    ("basic_miss", [((0, 0, 0), (2, 2, 2))], (5, 0, -5), (0, 0, 1), 10.0, None, None, None),  # :4397
    ("basic_inside", [((0, 0, 0), (2, 2, 2))], (0, 0, 0), (0, 0, 1), 10.0, 0, (0.9, 1.1), (0, 0, 1)),  # :4410
    ("multi_first", [((-4, 0, 0), (2, 2, 2)), ((0, 0, 0), (2, 2, 2)), ((4, 0, 0), (2, 2, 2))], (-4, 0, -5),
     (0, 0, 1), 10.0, 0, None, (0, 0, -1)),  # :4438
    ("multi_closest", [((-4, 0, 0), (2, 2, 2)), ((0, 0, 0), (2, 2, 2)), ((4, 0, 0), (2, 2, 2))], (-8, 0, 0),
     (1, 0, 0), 20.0, 0, (2.9, 3.1), (-1, 0, 0)),  # :4451
    ("analytic_face", [((0.5, 0.5, 0.5), (1, 1, 1))], (0.5, 0.5, -1), (0, 0, 1), -1.0, 0, (1 - 1e-6, 1 + 1e-6),
     (0, 0, -1)),  # :5246
    ("edge", [((0, 0, 0), (2, 2, 2))], (1, 0, -2), (0, 0, 1), -1.0, 0, (1 - 1e-6, 1 + 1e-6), None),  # :5275
    ("corner", [((0, 0, 0), (2, 2, 2))], (-1, -1, -2), (0, 0, 1), -1.0, 0, (1 - 1e-6, 1 + 1e-6), None),  # :5289
]


def consistency_rays():
    """selfTest.cpp:4465-4486: three voxels and a 5 x 3 fan of rays from z = -8."""
    boxes = [((-3, -3, 0), (1, 1, 1)), ((0, 0, 0), (2, 2, 2)), ((4, 2, 1), (3, 3, 3))]
    O, D = [], []
    for i in range(5):
        for j in range(3):
            O.append((i * 2.0 - 4.0, j * 2.0 - 2.0, -8.0))
            d = np.array([0.1 * i, 0.1 * j, 1.0], np.float32)
            D.append(d / np.float32(np.sqrt(np.float32(d @ d))))
    return boxes, np.array(O, np.float32), np.array(D, np.float32), 20.0


def grid_case():
    """selfTest.cpp:5299-5380: a 3x3x3 grid of unit voxels 1.5 apart and a 10 x 10 grid of +z rays from
    z = -1; the expected hit flags are the test's own rule."""
    boxes = [((i * 1.5, j * 1.5, k * 1.5), (1, 1, 1)) for i in range(3) for j in range(3) for k in range(3)]
    O = np.array([(np.float32(i) * np.float32(0.4), np.float32(j) * np.float32(0.4), -1.0) for i in range(10)
                  for j in range(10)], np.float32)
    D = np.tile(np.array([0, 0, 1], np.float32), (100, 1))
    expect = np.zeros(100, bool)
    eps = np.float32(1e-6)
    for r, (x, y, _) in enumerate(O):
        for (cx, cy, cz), _s in boxes:
            cx, cy = np.float32(cx), np.float32(cy)
            if (x >= cx - np.float32(0.5) - eps and x <= cx + np.float32(0.5) + eps and
                    y >= cy - np.float32(0.5) - eps and y <= cy + np.float32(0.5) + eps):
                expect[r] = True
                break
    return boxes, O, D, expect


def slab_fp64(boxes_lo, boxes_hi, o, d, max_dist=-1.0):
    """Independent FP64 closest voxel hit: (index, distance, outward normal) or (None, inf, None).  Distances
    within 1e-9 of the closest count as ties and go to the lower index."""
    best, bt, bn = None, np.inf, None
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    for k, (lo, hi) in enumerate(zip(boxes_lo, boxes_hi)):
        tn, tf = -np.inf, np.inf
        ok = True
        for a in range(3):
            if abs(d[a]) < 1e-8:
                if o[a] < lo[a] or o[a] > hi[a]:
                    ok = False
                continue
            t1, t2 = sorted(((lo[a] - o[a]) / d[a], (hi[a] - o[a]) / d[a]))
            tn, tf = max(tn, t1), min(tf, t2)
        if not ok or tn > tf or tf < 0:
            continue
        t = tn if tn >= 1e-6 else tf
        if t < 1e-6 or (max_dist > 0 and t >= max_dist):
            continue
        if t < bt - 1e-9:
            best, bt = k, t
            p = o + d * t
            c, e = (lo + hi) / 2, (hi - lo) / 2
            rel = np.abs(p - c) / e
            ax = int(np.argmax(rel))
            bn = np.zeros(3)
            bn[ax] = 1.0 if p[ax] - c[ax] > 0 else -1.0
    return best, bt, bn


def mixed_scene(seed, n_vox=300, n_surf=3000):
    """Voxels (half of them rotated about z) shuffled among the patches and triangles of a random canopy."""
    rng = np.random.default_rng(seed)
    s = scenes.random_canopy(seed, n_surf)
    cen = np.c_[rng.uniform(-2, 2, n_vox), rng.uniform(-2, 2, n_vox), rng.uniform(0, 2, n_vox)]
    size = rng.uniform(0.05, 0.4, (n_vox, 3))
    rot = np.where(rng.random(n_vox) < 0.5, 0.0, rng.uniform(0, np.pi, n_vox))
    TV = np.stack([voxel_T(c, z, r) for c, z, r in zip(cen, size, rot)])
    T = np.concatenate([s.transforms.reshape(-1, 16), TV]).astype(np.float32)
    ty = np.concatenate([s.types, np.full(n_vox, 4, np.uint32)]).astype(np.uint32)
    tw = np.concatenate([s.twosided, np.zeros(n_vox, np.uint32)]).astype(np.uint32)
    perm = rng.permutation(T.shape[0])
    return T[perm], ty[perm], tw[perm]


def mixed_rays(seed, T, ty, n):
    """Random rays over the scene, plus rays from voxel centres (origins inside a box), axis-parallel rays
    (exact zero direction components) and rays starting on a voxel face."""
    rng = np.random.default_rng(seed + 1000)
    O, D = scenes.rays(seed, n, [-2.5, -2.5, -0.2], [2.5, 2.5, 2.4])
    vox = np.nonzero(ty == 4)[0]
    m = min(len(vox), n // 8)
    pick = rng.choice(vox, m, replace=False)
    Tm = T[pick].reshape(-1, 16)
    O[:m] = Tm[:, [3, 7, 11]]
    ax = rng.integers(0, 3, n)
    sg = np.where(rng.random(n) < 0.5, -1.0, 1.0).astype(np.float32)
    q = slice(m, 2 * m)
    Dq = np.zeros((m, 3), np.float32)
    Dq[np.arange(m), ax[q]] = sg[q]
    D[q] = Dq
    # origins on a face: the box's lower z face of an axis-aligned voxel, moving up (the exit is the hit)
    f = slice(2 * m, 3 * m)
    lo = np.array([voxel_box(T[k])[0] for k in pick], np.float32)
    ctr = Tm[:, [3, 7, 11]]
    O[f] = np.c_[ctr[:, 0], ctr[:, 1], lo[:, 2]]
    D[f] = np.array([0, 0, 1], np.float32)
    return O, D
