"""Loader of tests/golden/cd_castrays_cases.json into scenes and rays."""
from __future__ import annotations

import json
import os

import numpy as np

from tests.closed_forms import patch_T, tri_T

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    with open(os.path.join(HERE, "golden", "cd_castrays_cases.json")) as f:
        return json.load(f)


def case_scene(c):
    Ts, types = [], []
    for v0, v1, v2 in c["tris"]:
        Ts.append(tri_T(v0, v1, v2))
        types.append(1)
    for (ctr, size) in c["patches"]:
        Ts.append(patch_T(ctr, (size[0], 0, 0), (0, size[1], 0)))
        types.append(0)
    P = len(Ts)
    return np.stack(Ts), np.array(types, np.uint32), np.zeros(P, np.uint32)


def grid_scene():
    Ts = [patch_T((i * 0.8, j * 0.8, 0), (0.5, 0, 0), (0, 0.5, 0)) for i in range(5) for j in range(5)]
    P = len(Ts)
    O = np.array([[i * 0.4, j * 0.4, -1.0] for i in range(10) for j in range(10)], np.float32)
    D = np.tile(np.array([0, 0, 1], np.float32), (100, 1))
    expect = np.zeros(100, bool)
    for k, (x, y, _) in enumerate(O.astype(np.float64)):
        for i in range(5):
            for j in range(5):
                px, py = i * 0.8, j * 0.8
                if px - 0.25 - 1e-6 <= x <= px + 0.25 + 1e-6 and py - 0.25 - 1e-6 <= y <= py + 0.25 + 1e-6:
                    expect[k] = True
    return np.stack(Ts), np.zeros(P, np.uint32), np.zeros(P, np.uint32), O, D, expect


def check_case(c, t, nrm, uu):
    """Assert one case's outcome (t, normal, uuid of a single ray)."""
    if c["hit"] is None:
        assert uu == 0xFFFFFFFF, c["name"]
        return
    assert uu == c["hit"], (c["name"], uu)
    o = np.array(c["origin"], np.float64)
    d = np.array(c["dir"], np.float64)
    if "t" in c:
        assert abs(t - c["t"]) < c["t_tol"], (c["name"], t)
    if "t_range" in c:
        assert c["t_range"][0] < t < c["t_range"][1], (c["name"], t)
    p = o + d * t
    assert np.max(np.abs(p - np.array(c["point"]))) < c["point_tol"], (c["name"], p)
    # castRaysSoA contract: unit normal facing the ray origin
    assert abs(np.linalg.norm(nrm) - 1.0) < 1e-5
    assert np.dot(nrm, o - p) >= 0.0
    if "normal_axis" in c:
        assert abs(abs(np.dot(nrm, c["normal_axis"])) - 1.0) < 1e-6
