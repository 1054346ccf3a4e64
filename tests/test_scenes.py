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
           "orchard_small": digest(scenes.orchard(3, 2, 1, 500, 320))}
    assert got["config1"] == digest(scenes.config1(1))
    with open(GOLD) as f:
        want = json.load(f)
    assert got == want
