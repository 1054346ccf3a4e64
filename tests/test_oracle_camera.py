"""The oracle's camera rays (camera_raygen.comp, pixel_label_raygen.comp) against closed forms (CPU).

Geometry that pins the pixel -> ray map: a camera looking straight down from height h over the plane
z = 0 sees, through the centre of pixel (i, j), the point (-h*a, h*b) with a = spy/viewplane and
b = spz/viewplane (the image's horizontal axis runs along -x and its vertical axis along +y at azimuth
0), and every pixel-centre depth along the view axis is h.  Radiance: a surface of uniform outgoing
flux L appears as L/pi whatever the antialiasing; a miss shows the sky radiance times the Prague
pattern (or 1 for an isotropic sky); a band that emits adds camera_diffuse_flux/pi; sources in view
add flux / area / pi (times the cosine to a planar source's normal)."""
import ctypes as C
import math

import numpy as np

from helios_amd.native import Camera
from oracle.oracle import OracleScene, lib
from tests import closed_forms as cf

PI_F = float(np.float32(3.14159265359))
DOWN = (math.pi, 0.0)          # direction[0] = pi looks along -z
UP = (0.0, 0.0)                # direction[0] = 0 looks along +z
EAST = (math.pi / 2, math.pi / 2)  # horizontal, along +x


def _plane(uuids=(10, 20), half=1.0):
    """Two patches tiling [-half, half]^2 at z = 0: x < 0 (uuids[0]) and x > 0 (uuids[1]), normals +z."""
    T = np.stack([cf.patch_T((-half / 2, 0, 0), (half, 0, 0), (0, 2 * half, 0)),
                  cf.patch_T((half / 2, 0, 0), (half, 0, 0), (0, 2 * half, 0))])
    return OracleScene(T, [0, 0], [0, 0], uuids=list(uuids))


def _sky(B=1, radiance=0.0, params=(0.0, 1.0, 1.0, 0.0), sun=(0.0, 0.0, 1.0), solar=0.0, solar_cos=2.0,
         diffuse_flux=0.0, emission=0):
    return dict(params=np.tile(np.asarray(params, np.float32), (B, 1)), radiance=np.full(B, radiance, np.float32),
                sun=sun, solar=np.full(B, solar, np.float32), solar_cos=solar_cos,
                diffuse_flux=np.full(B, diffuse_flux, np.float32), emission=np.full(B, emission, np.uint32))


def _centre_dirs(cam):
    """Pixel-centre unit directions of a downward camera at azimuth 0 in double: (-a, b, -1)/norm."""
    W, H = cam.resolution
    ii, jj = np.meshgrid(np.arange(W), np.arange(H))
    spy = -0.5 + (ii + 0.5) / W
    spz = (0.5 - (jj + 0.5) / H) / cam.fov_aspect
    a, b = spy / cam.viewplane_length, spz / cam.viewplane_length
    return a.reshape(-1), b.reshape(-1)


def test_labels_and_depths_of_a_plane_seen_from_above():
    o = _plane(half=1.0)
    h = 3.0
    cam = Camera((0.0, 0.0, h), DOWN, (32, 26), viewplane_length=1.0, fov_aspect=1.25)
    lab, dep = o.pixel_labels(cam)
    a, b = _centre_dirs(cam)
    x, y = -h * a, h * b
    inside = (np.abs(x) < 1.0) & (np.abs(y) < 1.0)
    expect = np.where(inside, np.where(x < 0, 11, 21), 0)
    assert np.array_equal(lab, expect)
    assert inside.sum() > 20 and (~inside).sum() > 20
    assert np.allclose(dep[inside], h, rtol=1e-5)
    assert np.all(dep[~inside] == -1.0)


def test_label_tiles_compose_the_full_image():
    o = _plane()
    cam = Camera((0.2, -0.1, 3.0), (math.pi - 0.2, 0.4), (20, 16), viewplane_length=0.3)
    full_l, full_d = o.pixel_labels(cam)
    lab = np.zeros(20 * 16, np.uint32)
    dep = np.zeros(20 * 16, np.float32)
    for tile in [(0, 0, 12, 10), (12, 0, 8, 10), (0, 10, 12, 6), (12, 10, 8, 6)]:
        o.pixel_labels(cam, tile=tile, labels=lab, depths=dep)
    assert np.array_equal(lab, full_l) and np.array_equal(dep, full_d)


def _render(o, cam, top, bottom, sky, seed=3, **kw):
    P = 2
    ot = np.full((P, 1), top, np.float32)
    ob = np.full((P, 1), bottom, np.float32)
    S = len(kw.get("source_pos", []))
    return o.camera_pixels(cam, ot, ob, kw.pop("source_flux", np.zeros((S, 1))), kw.pop("source_pos", np.zeros((0, 3))),
                           sky, seed=seed, **kw)[0]


def test_uniform_surface_radiance_and_face():
    o = _plane(half=50.0)
    L_top, L_bot = 300.0, 40.0
    for spp in (1, 4):
        cam = Camera((0.0, 0.0, 2.0), DOWN, (16, 16), viewplane_length=1.0, antialiasing_samples=spp)
        px = _render(o, cam, L_top, L_bot, _sky())
        assert np.allclose(px, L_top / math.pi, rtol=1e-5), spp
    cam = Camera((0.0, 0.0, -2.0), UP, (16, 16), viewplane_length=1.0)
    px = _render(o, cam, L_top, L_bot, _sky())
    assert np.allclose(px, L_bot / math.pi, rtol=1e-5)


def test_isotropic_sky_and_emission():
    o = _plane(half=1.0)
    cam = Camera((0.0, 0.0, 1.0), UP, (8, 8), viewplane_length=1.0, antialiasing_samples=2)
    px = _render(o, cam, 0.0, 0.0, _sky(radiance=25.0))
    assert np.allclose(px, 25.0, rtol=1e-6)
    px = _render(o, cam, 0.0, 0.0, _sky(radiance=25.0, diffuse_flux=60.0, emission=1))
    assert np.allclose(px, 25.0 + 60.0 / math.pi, rtol=1e-6)
    px = _render(o, cam, 0.0, 0.0, _sky(radiance=25.0, diffuse_flux=60.0, emission=0))
    assert np.allclose(px, 25.0, rtol=1e-6)


def _jitter(n, seed):
    """Sub-pixel offsets (Rx, Ry) of sample 0 of pixels 0..n-1: the first two draws of the ray's stream."""
    L = lib()
    out = np.zeros((n, 2))
    st = C.c_uint32()
    for k in range(n):
        st.value = L.orc_tea16(k, seed)
        out[k, 0] = L.orc_rnd(C.byref(st))
        out[k, 1] = L.orc_rnd(C.byref(st))
    return out


def test_prague_sky_pattern():
    o = _plane(half=1.0)
    W = H = 24
    cam = Camera((0.0, 0.0, 1.0), EAST, (W, H), viewplane_length=0.5)
    sun = np.array([0.6, 0.0, 0.8])
    prm = (2.5, 40.0, 1.7, 0.3)
    px = _render(o, cam, 0.0, 0.0, _sky(radiance=10.0, params=prm, sun=tuple(sun)), seed=9)
    # sample directions of a camera looking along +x: image x-axis -> -y, image y-axis -> +z
    R = _jitter(W * H, 9)
    ii, jj = np.meshgrid(np.arange(W), np.arange(H))
    a = (-0.5 + (ii.reshape(-1) + R[:, 0]) / W) / cam.viewplane_length
    b = (0.5 - (jj.reshape(-1) + R[:, 1]) / H) / cam.viewplane_length
    d = np.stack([np.ones_like(a), -a, b], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    gamma = np.degrees(np.arccos(np.clip(d @ sun, -1, 1)))
    w = (1 + prm[0] * np.exp(-gamma / prm[1])) * (1 + (prm[2] - 1) * (1 - np.maximum(d[:, 2], 0))) * prm[3] * math.pi
    sky = 10.0 * w
    e = np.abs(px - sky) / sky
    assert e.max() < 2e-5, e.max()
    assert px.max() / px.min() > 1.5


def test_solar_disk_replaces_the_sky():
    o = _plane(half=1.0)
    sun = np.array([0.0, 0.0, 1.0])
    cam = Camera((0.0, 0.0, 1.0), UP, (33, 33), viewplane_length=1.0)
    cos_disk = math.cos(math.radians(5.0))
    sky = _sky(radiance=20.0, sun=tuple(sun), solar=5000.0, solar_cos=cos_disk)
    px = _render(o, cam, 0.0, 0.0, sky, source_pos=[[0, 0, 1]], source_flux=[[800.0]])
    c = 16 * 33 + 16
    assert math.isclose(px[c], 5000.0, rel_tol=1e-6)
    assert math.isclose(px[0], 20.0, rel_tol=1e-6)
    # a zero-flux collimated source shows no disk
    px0 = _render(o, cam, 0.0, 0.0, sky, source_pos=[[0, 0, 1]], source_flux=[[0.0]])
    assert math.isclose(px0[c], 20.0, rel_tol=1e-6)


def test_sources_in_view():
    o = _plane(half=1.0)
    cam = Camera((0.0, 0.0, 1.0), UP, (33, 33), viewplane_length=1.0)
    c = 16 * 33 + 16
    # sphere of diameter 0.4 at height 3: radiance flux / (4 pi r^2) / pi at the centre pixel
    flux, r = 500.0, 0.2
    px = _render(o, cam, 0.0, 0.0, _sky(radiance=7.0), source_pos=[[0, 0, 3.0]], source_flux=[[flux]],
                 source_types=[1], source_widths=[[2 * r, 0.0]], source_rotations=[[0, 0, 0]])
    assert math.isclose(px[c], flux / (4 * math.pi * r * r) / math.pi, rel_tol=1e-5)
    assert math.isclose(px[0], 7.0, rel_tol=1e-6)
    # rectangle 0.5 x 0.3 at height 3 facing down (rotation (pi, 0, 0)): flux/area * cos / pi, cos ~ 1
    px = _render(o, cam, 0.0, 0.0, _sky(), source_pos=[[0, 0, 3.0]], source_flux=[[flux]], source_types=[3],
                 source_widths=[[0.5, 0.3]], source_rotations=[[math.pi, 0, 0]])
    assert math.isclose(px[c], flux / 0.15 / math.pi, rel_tol=1e-4)
    # facing up (no rotation) it is seen from behind: nothing
    px = _render(o, cam, 0.0, 0.0, _sky(), source_pos=[[0, 0, 3.0]], source_flux=[[flux]], source_types=[3],
                 source_widths=[[0.5, 0.3]], source_rotations=[[0, 0, 0]])
    assert px[c] == 0.0
    # disk of radius 0.25 facing down
    px = _render(o, cam, 0.0, 0.0, _sky(), source_pos=[[0, 0, 3.0]], source_flux=[[flux]], source_types=[4],
                 source_widths=[[0.25, 0.0]], source_rotations=[[math.pi, 0, 0]])
    assert math.isclose(px[c], flux / (math.pi * 0.0625) / math.pi, rel_tol=1e-4)
    # a source behind the surface hit is hidden: camera looks down at the plane, sphere below it
    cam_d = Camera((0.0, 0.0, 1.0), DOWN, (33, 33), viewplane_length=1.0)
    px = _render(o, cam_d, 90.0, 0.0, _sky(), source_pos=[[0, 0, -1.0]], source_flux=[[flux]], source_types=[1],
                 source_widths=[[0.4, 0.0]], source_rotations=[[0, 0, 0]])
    assert math.isclose(px[c], 90.0 / math.pi, rel_tol=1e-5)


def test_periodic_wrap_depth():
    """Periodic domain [-1, 1]^2 (ground patch), a wall at x = 0.5 facing -x; the camera at x = 0.8
    looking along +x reaches x = 1 after 0.2, wraps to x = -1 and meets the wall after 1.5 more."""
    T = np.stack([cf.patch_T((0, 0, 0), (2, 0, 0), (0, 2, 0)),
                  cf.patch_T((0.5, 0, 0.5), (0, 2, 0), (0, 0, 1))])
    o = OracleScene(T, [0, 0], [0, 0], uuids=[1, 2], periodic=(True, True))
    cam = Camera((0.8, 0.0, 0.5), EAST, (5, 5), viewplane_length=4.0)
    lab, dep = o.pixel_labels(cam)
    assert lab[12] == 3
    assert abs(dep[12] - 1.7) < 1e-4
