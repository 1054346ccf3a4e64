"""The homogeneous-canopy benchmark scene (benchmarks/radiation_homogeneous_canopy/main.cpp) (CPU).

Leaf count from the leaf area index and the mask's solid area, 2 x 2 sub-patches of w/2 per leaf sharing
one centre and normal, quarter-texture UVs, spherical leaf-angle distribution (mean |cos| of the normal's
zenith = 1/2), heights in [w/2, w/2 + h], one-sided ground tile over the periodic domain."""
import math

import numpy as np

from helios_amd import scenes


def test_homogeneous_canopy_geometry():
    s, masks, mid, uv, uvid = scenes.homogeneous_canopy(3, D=4.0, lai=1.0, h=1.0, w_leaf=0.1, mask_res=64)
    ng = 100 * 100
    nl = s.P - ng
    assert nl % 4 == 0
    solid = float(masks[0].mean())
    assert abs(solid - math.pi / 4) < 0.02
    assert nl // 4 == round(1.0 * 16.0 / (0.01 * solid))
    T = s.transforms.reshape(-1, 16).astype(np.float64)
    ex, ey, ez = T[:, [0, 4, 8]], T[:, [1, 5, 9]], T[:, [2, 6, 10]]
    c = T[:, [3, 7, 11]]
    # sub-patch sides are w/2, orthogonal, and the stored normal is their unit cross product
    assert np.allclose(np.linalg.norm(ex[:nl], axis=1), 0.05, rtol=1e-5)
    assert np.allclose(np.linalg.norm(ey[:nl], axis=1), 0.05, rtol=1e-5)
    assert np.allclose((ex[:nl] * ey[:nl]).sum(1), 0.0, atol=1e-7)
    n = np.cross(ex, ey)
    assert np.allclose(n[:nl] / np.linalg.norm(n[:nl], axis=1, keepdims=True), ez[:nl], atol=1e-5)
    # the four sub-patches of a leaf surround its centre
    leaf_c = c[:nl].reshape(-1, 4, 3).mean(1)
    assert ((leaf_c[:, 2] >= 0.05 - 1e-6) & (leaf_c[:, 2] <= 1.05 + 1e-6)).all()
    assert (np.abs(leaf_c[:, :2]) <= 2.0 + 1e-6).all()
    # spherical leaf angles: |cos zenith| of the normals is uniform on [0, 1]
    cz = np.abs(ez[:nl:4, 2])
    assert abs(cz.mean() - 0.5) < 0.02 and abs(np.median(cz) - 0.5) < 0.03
    # quarter UVs per sub-patch, masks on leaves only, ground one-sided, periodic domain
    q = uv[:nl].reshape(-1, 4, 4, 2)[0]
    assert np.allclose(q[:, 2] - q[:, 0], 0.5)
    assert (mid[:nl] == 0).all() and (mid[nl:] == -1).all() and (uvid[:nl] >= 0).all()
    assert (s.twosided[:nl] == 1).all() and (s.twosided[nl:] == 0).all()
    assert s.periodic == (True, True)
    assert np.isclose(np.abs(ex[nl:, 0]).sum() * np.abs(ey[nl:, 1]).mean(), 16.0)
