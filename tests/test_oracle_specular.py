"""The oracle's Blinn-Phong camera term (camera_raygen.comp:362-409) against closed forms (CPU).

A horizontal plane of patches (normal +z) with no outgoing radiation is seen by a narrow pinhole camera
(viewplane length 40, so every ray lies within 0.72 degrees of the view axis); a hit pixel then holds
only the specular term divided by pi.  For one source whose radiation_specular on the hit primitive is E,
camera weight w, exponent e and scale k (mode 2, else 1), a pixel of any antialiasing count reads

    E * w * 0.25 * k * cos(g)^e * (e + 2) / (2 pi) / pi,

with g the angle between the surface normal and the half vector of the light direction l and the
reversed view direction -d, i.e. half the angle between l and -d.  Each case sets g through the
geometry: a camera looking straight down under a zenith sun (g = 0), a camera 60 degrees off nadir
(g = 30 degrees), and a point-like sphere source at 45 degrees elevation seen by a nadir camera
(g = 22.5 degrees, l taken towards the source's centre from the hit point).  The narrow view bounds each
pixel's g within 0.4 degrees of the nominal value, and the expected value is checked between the
bounds.  The radiation a fully lit horizontal face receives from a zenith sun is the source flux (the
face's strength sum is 1), the E the direct launch stores for it."""
import math

import numpy as np

from helios_amd.native import Camera
from oracle.oracle import OracleScene
from tests import closed_forms as cf

DOWN = (math.pi, 0.0)
SPREAD = math.degrees(math.atan(math.hypot(0.5, 0.5) / 40.0))  # widest ray angle from the view axis


def _plane():
    T = np.stack([cf.patch_T((-0.5, 0, 0), (1.0, 0, 0), (0, 2.0, 0)), cf.patch_T((0.5, 0, 0), (1.0, 0, 0), (0, 2.0, 0))])
    return OracleScene(T, [0, 0], [0, 0])


def _sky(B):
    return dict(params=np.tile(np.asarray([0.0, 1.0, 1.0, 0.0], np.float32), (B, 1)), radiance=np.zeros(B, np.float32),
                sun=(0.0, 0.0, 1.0), solar=np.zeros(B, np.float32), solar_cos=2.0,
                diffuse_flux=np.zeros(B, np.float32), emission=np.zeros(B, np.uint32))


def _render(o, cam, E, w, e, k, mode, pos, types=None, widths=None, B=1):
    P = o.P
    S = len(pos)
    spec = dict(spec=np.full((S, P, B), E, np.float32), weight=np.full((S, B), w, np.float32),
                exponent=np.full(P, e, np.float32), scale=np.full(P, k, np.float32), mode=mode)
    zero = np.zeros((P, B), np.float32)
    flux = np.full((S, B), 500.0, np.float32)
    return o.camera_pixels(cam, zero, zero, flux, np.asarray(pos, np.float32), _sky(B), seed=4, specular=spec,
                           source_types=types, source_widths=widths)[0]


def _expect(E, w, e, k, g_deg):
    return E * w * 0.25 * k * math.cos(math.radians(g_deg)) ** e * (e + 2.0) / (2.0 * math.pi) / math.pi


def _check_band(px, E, w, e, k, g_deg, spread):
    lo = _expect(E, w, e, k, g_deg + spread)
    hi = _expect(E, w, e, k, max(g_deg - spread, 0.0))
    assert np.all(px >= lo * (1 - 1e-5)) and np.all(px <= hi * (1 + 1e-5)), (px.min(), px.max(), lo, hi)


def test_zenith_sun_nadir_camera():
    o = _plane()
    for spp in (1, 3):
        cam = Camera((0.0, 0.0, 2.0), DOWN, (12, 10), viewplane_length=40.0, antialiasing_samples=spp)
        px = _render(o, cam, E=800.0, w=0.8, e=50.0, k=1.0, mode=1, pos=[(0.0, 0.0, 1.0)])
        # g is half the angle between -d and l: at most half the ray spread
        _check_band(px, 800.0, 0.8, 50.0, 1.0, 0.0, SPREAD / 2)


def test_off_nadir_camera_and_scale_mode():
    o = _plane()
    r = 3.0
    # view direction (sin 120, 0, cos 120): 60 degrees from nadir, towards +x; -d is 60 degrees from the zenith
    cam = Camera((-r * math.sin(math.radians(120)), 0.0, -r * math.cos(math.radians(120))),
                 (math.radians(120), math.pi / 2), (10, 8), viewplane_length=40.0)
    px1 = _render(o, cam, E=600.0, w=1.0, e=10.0, k=0.3, mode=1, pos=[(0.0, 0.0, 1.0)])
    _check_band(px1, 600.0, 1.0, 10.0, 1.0, 30.0, SPREAD / 2)
    # mode 2 multiplies by the primitive's scale, mode 1 ignores it
    px2 = _render(o, cam, E=600.0, w=1.0, e=10.0, k=0.3, mode=2, pos=[(0.0, 0.0, 1.0)])
    assert np.allclose(px2, px1 * 0.3, rtol=1e-6)


def test_sphere_source_light_points_at_its_centre():
    o = _plane()
    cam = Camera((0.0, 0.0, 2.0), DOWN, (10, 10), viewplane_length=40.0)
    # a 1 mm sphere 424 m away at 45 degrees elevation: l within 0.01 degrees of 45 degrees over the view
    px = _render(o, cam, E=400.0, w=1.0, e=20.0, k=1.0, mode=1, pos=[(300.0, 0.0, 300.0)], types=[1],
                 widths=[(0.001, 0.0)])
    _check_band(px, 400.0, 1.0, 20.0, 1.0, 22.5, SPREAD / 2 + 0.01)
    # the same position read as a collimated direction gives the same l
    pc = _render(o, cam, E=400.0, w=1.0, e=20.0, k=1.0, mode=1, pos=[(1.0, 0.0, 1.0)])
    assert np.allclose(px, pc, rtol=1e-3)  # hit points differ by a few cm, 424 m from the source


def test_no_term_without_positive_exponent_weight_or_radiation():
    o = _plane()
    cam = Camera((0.0, 0.0, 2.0), DOWN, (6, 6), viewplane_length=40.0)
    for E, w, e in ((800.0, 1.0, -1.0), (800.0, 1.0, 0.0), (800.0, 0.0, 50.0), (0.0, 1.0, 50.0)):
        assert np.all(_render(o, cam, E=E, w=w, e=e, k=1.0, mode=1, pos=[(0.0, 0.0, 1.0)]) == 0.0)
    # two sources add
    one = _render(o, cam, E=800.0, w=1.0, e=5.0, k=1.0, mode=1, pos=[(0.0, 0.0, 1.0)])
    two = _render(o, cam, E=800.0, w=1.0, e=5.0, k=1.0, mode=1, pos=[(0.0, 0.0, 1.0), (0.0, 0.0, 1.0)])
    assert np.allclose(two, 2 * one, rtol=1e-6)


def test_direct_launch_radiation_of_a_lit_horizontal_face_is_the_flux():
    o = _plane()
    sums = o.direct_sums_src(8, 5, [(0.0, 0.0, 1.0)])  # [P][S][2]
    assert np.allclose(sums[:, 0, 0] + sums[:, 0, 1], 1.0, rtol=1e-6)
