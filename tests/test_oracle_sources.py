"""The oracle's non-collimated sources (direct_raygen.comp:482-628) against closed forms (CPU).

A disk or rectangle source is a Lambertian emitter: the mean of cos(theta_1) cos(theta_2) / s^2 over
uniform source points, times the source area / pi, is the view factor from a small receiver to the
source; a sphere source seen along the receiver's normal from distance r gives (R/r)^2 (times the mean
of r^2/s^2 over the sphere).  Each test uses a 1 cm receiver patch, so the closed forms for a
differential area apply, and enough stratified rays for a 1-2 % Monte Carlo error."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleScene, lib
from tests import closed_forms as cf

DOWN = (math.pi, 0.0, 0.0)  # turns a planar source's normal R*(0,0,1) to -z: it lights what lies below it


def _receiver(tilt_deg=0.0):
    a = math.radians(tilt_deg)
    ex = (0.01, 0.0, 0.0)
    ey = (0.0, 0.01 * math.cos(a), 0.01 * math.sin(a))
    T = cf.patch_T((0, 0, 0), ex, ey)[None]
    return OracleScene(T, [0], [0])


def _sums(o, pos, ty, width, rot=(0.0, 0.0, 0.0), d=200, seed=5):
    out = o.direct_sums_src(d, seed, [pos], types=[ty], widths=[width], rotations=[rot])
    return out[0, 0, 0], out[0, 0, 1]


def disk_view_factor(R, h):
    return R * R / (h * h + R * R)


def rect_view_factor(w, l, h):
    """Receiver below the centre of a parallel w x l rectangle at height h: four corner rectangles."""
    def corner(a, b):
        A, B = a / h, b / h
        return (A / math.sqrt(1 + A * A) * math.atan(B / math.sqrt(1 + A * A)) +
                B / math.sqrt(1 + B * B) * math.atan(A / math.sqrt(1 + B * B))) / (2 * math.pi)
    return 4 * corner(w / 2, l / 2)


def test_rotation_matrix_closed_form():
    m = np.zeros(9, np.float32)
    for rot in [(0.0, 0.0, 0.0), (math.pi, 0.0, 0.0), (0.3, -0.7, 1.1)]:
        r = np.asarray(rot, np.float32)
        lib().orc_rotation_matrix(r.ctypes.data_as(lib().orc_rotation_matrix.argtypes[0]),
                                  m.ctypes.data_as(lib().orc_rotation_matrix.argtypes[1]))
        rx, ry, rz = (float(v) for v in r)
        Rx = np.array([[1, 0, 0], [0, math.cos(rx), -math.sin(rx)], [0, math.sin(rx), math.cos(rx)]])
        Ry = np.array([[math.cos(ry), 0, math.sin(ry)], [0, 1, 0], [-math.sin(ry), 0, math.cos(ry)]])
        Rz = np.array([[math.cos(rz), -math.sin(rz), 0], [math.sin(rz), math.cos(rz), 0], [0, 0, 1]])
        # transforms.glsl:49-64 lists the columns of Rz*Ry*Rx (right-handed rotations); m holds its rows
        M = m.reshape(3, 3).astype(np.float64)
        assert abs(np.linalg.det(M) - 1.0) < 1e-6
        assert np.allclose(M, Rz @ Ry @ Rx, atol=1e-6)
    # (pi, 0, 0) flips the source normal R*(0,0,1) to -z
    r = np.asarray(DOWN, np.float32)
    lib().orc_rotation_matrix(r.ctypes.data_as(lib().orc_rotation_matrix.argtypes[0]),
                              m.ctypes.data_as(lib().orc_rotation_matrix.argtypes[1]))
    assert np.allclose(m[[2, 5, 8]], [0, 0, -1], atol=1e-6)


@pytest.mark.parametrize("R,h", [(0.5, 1.0), (1.0, 0.5), (0.2, 2.0)])
def test_disk_source_view_factor(R, h):
    o = _receiver()
    top, bot = _sums(o, (0.0, 0.0, h), 4, (R, R), DOWN)
    exact = disk_view_factor(R, h)
    assert bot == 0.0
    assert abs(top - exact) / exact <= 0.02, (top, exact)


@pytest.mark.parametrize("w,l,h", [(1.0, 2.0, 1.0), (0.5, 0.5, 2.0), (3.0, 1.0, 0.5)])
def test_rectangle_source_view_factor(w, l, h):
    o = _receiver()
    top, bot = _sums(o, (0.0, 0.0, h), 3, (w, l), DOWN)
    exact = rect_view_factor(w, l, h)
    assert bot == 0.0
    assert abs(top - exact) / exact <= 0.02, (top, exact)


def test_back_face_of_planar_sources_is_dark():
    o = _receiver()
    # unrotated sources face +z: a receiver below sees their back side
    for ty, wd in ((3, (1.0, 1.0)), (4, (0.5, 0.5))):
        top, bot = _sums(o, (0.0, 0.0, 1.0), ty, wd, (0.0, 0.0, 0.0))
        assert top == 0.0 and bot == 0.0


@pytest.mark.parametrize("r,zen", [(10.0, 0.0), (20.0, 40.0)])
def test_sphere_source_inverse_square(r, zen):
    o = _receiver()
    R = 1.0
    z = math.radians(zen)
    pos = (r * math.sin(z), 0.0, r * math.cos(z))
    top, bot = _sums(o, pos, 1, (2 * R, 2 * R))
    # mean of 1/s^2 over the sphere surface seen from distance r
    inv_s2 = math.log((r + R) / (r - R)) / (2 * r * R)
    exact = math.cos(z) * R * R * inv_s2
    assert bot == 0.0
    assert abs(top - exact) / exact <= 0.02, (top, exact)


def test_sun_sphere_equals_sphere():
    o = _receiver()
    a = _sums(o, (3.0, 1.0, 12.0), 1, (1.5, 1.5))
    b = _sums(o, (3.0, 1.0, 12.0), 2, (1.5, 1.5))
    assert a == b


def test_source_beyond_occluder_is_blocked_and_beyond_source_is_not():
    # receiver at z=0 facing +z, an opaque 1 m plate at z=1, sphere source at z=2 (blocked) or z=0.5 (the
    # plate lies beyond the source: not blocking)
    T = np.stack([cf.patch_T((0, 0, 0), (0.01, 0, 0), (0, 0.01, 0)), cf.patch_T((0, 0, 1.0), (1, 0, 0), (0, 1, 0))])
    o = OracleScene(T, [0, 0], [0, 1])
    blocked = o.direct_sums_src(20, 3, [(0, 0, 2.0)], types=[1], widths=[(0.2, 0.2)])
    assert blocked[0, 0, 0] == 0.0
    lit = o.direct_sums_src(20, 3, [(0, 0, 0.5)], types=[1], widths=[(0.2, 0.2)])
    assert lit[0, 0, 0] > 0.0


def test_mixed_sources_draw_in_order():
    # a collimated source draws nothing; each other source draws two numbers from the ray's stream, so a
    # disk's result depends on the sources listed before it
    o = _receiver()
    disk = dict(pos=(0.0, 0.0, 1.0), w=(0.5, 0.5), rot=DOWN)
    alone = o.direct_sums_src(20, 7, [disk["pos"]], types=[4], widths=[disk["w"]], rotations=[disk["rot"]])
    after_col = o.direct_sums_src(20, 7, [(0, 0, 1), disk["pos"]], types=[0, 4], widths=[(0, 0), disk["w"]],
                                  rotations=[(0, 0, 0), disk["rot"]])
    after_sph = o.direct_sums_src(20, 7, [(0, 0, 5), disk["pos"]], types=[1, 4], widths=[(0.1, 0.1), disk["w"]],
                                  rotations=[(0, 0, 0), disk["rot"]])
    assert after_col[0, 1, 0] == alone[0, 0, 0]
    assert after_sph[0, 1, 0] != alone[0, 0, 0]
    assert abs(after_sph[0, 1, 0] - alone[0, 0, 0]) / alone[0, 0, 0] < 0.05
