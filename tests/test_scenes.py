"""Seeded scene generators: deterministic, right sizes, golden checksums (CPU)."""
import hashlib
import json
import os

import numpy as np

from helios_amd import scenes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scene_sha256.json")


def digest(s):
    h = hashlib.sha256()
    for a in (s.transforms, s.types, s.twosided):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_sizes():
    assert scenes.config1(1).P == 1601
    s2 = scenes.config2(2)
    assert s2.P == 200_400 and s2.periodic == (True, True)
    s3 = scenes.orchard(3, nx=1, ny=1, leaves_per_tree=1000)
    assert s3.P == 1000 + 2000 + 64


def test_config2_leaf_areas_and_lai():
    s = scenes.config2(2)
    a = s.areas()
    leaf = s.twosided == 1
    assert a[leaf].min() >= 4e-4 * 0.999 and a[leaf].max() <= 16e-4 * 1.001
    lai = a[leaf].sum() / 100.0
    assert 1.6 < lai < 1.9  # log-uniform 4-16 cm^2 -> mean 8.66 cm^2 -> LAI ~ 1.73


def test_deterministic_and_golden():
    got = {"config1": digest(scenes.config1(1)), "config2_small": digest(scenes.config2(2, 5000)),
           "orchard_small": digest(scenes.orchard(3, 2, 1, 500, 320)),
           "mixed_small": digest(scenes.mixed_canopy(5, 4000))}
    assert got["config1"] == digest(scenes.config1(1))
    with open(GOLD) as f:
        want = json.load(f)
    assert got == want


def test_mixed_canopy_disks_are_triangle_fans():
    # every third leaf is a 6-triangle disk built as Context::addDisk builds one ring (core/src/Context.cpp:
    # 2503-2537): triangles (centre, p_t, p_t+1) with the p_t on a circle around the shared centre
    s = scenes.mixed_canopy(5, 4000)
    T = s.transforms.reshape(-1, 16).astype(np.float64)
    # per leaf triple: 1 patch, 1 triangle, 6 disk triangles
    assert np.array_equal(s.types[:8], [0, 1, 1, 1, 1, 1, 1, 1])
    v0 = T[2:8][:, [3, 7, 11]]                            # triangle vertex 0 = the disk centre
    v1 = v0 + T[2:8][:, [1, 5, 9]]                        # vertex 1 = v0 + column 1
    v2 = v1 + T[2:8][:, [0, 4, 8]]                        # vertex 2 = v1 + column 0
    assert np.allclose(v0, v0[0], atol=0)
    r1, r2 = np.linalg.norm(v1 - v0, axis=1), np.linalg.norm(v2 - v0, axis=1)
    rad = np.sqrt(1e-3 / np.pi)
    assert np.allclose(r1, rad, rtol=1e-5) and np.allclose(r2, rad, rtol=1e-5)
    assert np.allclose(v2[:-1], v1[1:], atol=1e-7) and np.allclose(v2[-1], v1[0], atol=1e-7)  # closed fan
    n = np.cross(v1 - v0, v2 - v0)
    assert np.allclose(n / np.linalg.norm(n, axis=1, keepdims=True), (n / np.linalg.norm(n, axis=1, keepdims=True))[0],
                       atol=1e-5)  # coplanar, one orientation
    assert np.all(s.twosided[:-100] == 1) and np.all(s.twosided[-100:] == 0)
