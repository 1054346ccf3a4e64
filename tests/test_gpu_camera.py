"""Camera and pixel-label rays through the C ABI against the oracle (GPU).

Same rays bit for bit: labels must match exactly; depths to 1e-6 relative (both FP32 with the same
operation order, the wrap distances summed alike); pixel radiance to 1e-5 relative (the device sums each
pixel's FP32 samples in order in FP32, the oracle in FP64; the Prague pattern's acos/exp come from
different libraries)."""
import math

import numpy as np
import pytest

from helios_amd import native
from helios_amd.native import Camera
from oracle.oracle import OracleScene
from tests.helpers import rel_err, small_canopy

pytestmark = pytest.mark.gpu

TILES = [(0, 0, 40, 20), (40, 0, 24, 20), (0, 20, 40, 28), (40, 20, 24, 28)]


def _setup(ctx, s, B=2, S_pos=None, types=None, widths=None, rotations=None, seed=3):
    be = native.Backend(ctx)
    be.update_geometry(s.transforms, s.types, s.twosided, periodic=s.periodic)
    be.build_acceleration_structure()
    be.update_materials(np.full((s.P, B), 0.1), np.full((s.P, B), 0.05))
    pos = np.asarray([[0.3, 0.2, 0.9]] if S_pos is None else S_pos, np.float32)
    be.update_sources(pos, types=types, widths=widths, rotations=rotations)
    be.zero_radiation_buffers(B)
    rng = np.random.default_rng(seed)
    top = rng.uniform(0, 400, (s.P, B)).astype(np.float32)
    bot = rng.uniform(0, 100, (s.P, B)).astype(np.float32)
    be.upload_radiation_out(top, bot)
    flux = rng.uniform(100, 900, (pos.shape[0], B)).astype(np.float32)
    be.upload_source_fluxes(flux)
    return be, top, bot, flux, pos


def _sky(B, sun):
    sun = np.asarray(sun, np.float32) / np.linalg.norm(sun)
    return dict(params=np.tile(np.asarray([2.0, 30.0, 1.5, 0.25], np.float32), (B, 1)),
                radiance=np.linspace(5, 15, B).astype(np.float32), sun=tuple(float(v) for v in sun),
                solar=np.linspace(3000, 4000, B).astype(np.float32), solar_cos=math.cos(math.radians(4.0)),
                diffuse_flux=np.linspace(50, 80, B).astype(np.float32), emission=np.arange(B, dtype=np.uint32) % 2)


def _upload_sky(be, sky):
    be.update_sky_model(sky["params"], sky["radiance"], sky["sun"], sky["solar"], sky["solar_cos"],
                        sky["diffuse_flux"], sky["emission"])


def _oracle_pixels(s, cam, top, bot, flux, pos, sky, seed, tiles=None, **src):
    o = OracleScene.from_arrays(s)
    fx, fy = cam.resolution
    out = np.zeros((top.shape[1], fx * fy))
    for t in tiles or [None]:
        o.camera_pixels(cam, top, bot, flux, pos, sky, seed=seed, tile=t, pixels=out, **src)
    return out.T  # [pixel][band]


CAMS = {
    "oblique_pinhole": Camera((0.5, -4.0, 2.5), (math.radians(115), math.radians(20)), (64, 48),
                              viewplane_length=0.8, fov_aspect=64 / 48, antialiasing_samples=3),
    "lens_top_down": Camera((0.0, 0.0, 4.0), (math.pi, 0.3), (64, 48), focal_length=3.0, lens_diameter=0.2,
                            viewplane_length=0.6, fov_aspect=64 / 48, antialiasing_samples=2),
    "horizon_sun": Camera((-5.0, 0.0, 0.8), (math.radians(80), math.radians(90)), (64, 48),
                          viewplane_length=0.5, fov_aspect=64 / 48),
}


@pytest.mark.parametrize("name", sorted(CAMS))
@pytest.mark.parametrize("periodic", [False, True])
def test_camera_pixels_match_oracle(ctx, name, periodic):
    cam = CAMS[name]
    s = small_canopy(11, 1500, periodic=periodic)
    sun = (0.3, 0.2, 0.9) if name != "horizon_sun" else (1.0, 0.0, 0.15)
    be, top, bot, flux, pos = _setup(ctx, s, S_pos=[sun])
    sky = _sky(2, sun)
    _upload_sky(be, sky)
    be.launch_camera_rays(cam, seed=17)
    R = be.get_camera_results(cam.resolution)
    ref = _oracle_pixels(s, cam, top, bot, flux, pos, sky, 17)
    assert R["pixel_data"].shape == ref.shape
    e = rel_err(R["pixel_data"], ref, floor=1e-3)
    assert e.max() <= 1e-5, f"{name}: max rel err {e.max():.3e} at {np.unravel_index(e.argmax(), e.shape)}"
    assert (ref > 0).mean() > 0.5


def test_camera_tiles_and_sources_match_oracle(ctx):
    """Four tiles of one image with a sphere, a rectangle and a disk source in view."""
    cam = Camera((0.0, 0.0, 0.2), (0.35, 0.0), (64, 48), viewplane_length=0.9, fov_aspect=64 / 48,
                 antialiasing_samples=2)
    s = small_canopy(12, 1200)
    S_pos = [[0.3, 0.2, 0.9], [0.0, 0.6, 3.0], [0.8, -0.5, 3.5], [-0.9, 0.4, 2.5]]
    types = [0, 1, 3, 4]
    widths = [[0, 0], [0.5, 0.0], [0.8, 0.6], [0.4, 0.0]]
    rots = [[0, 0, 0], [0, 0, 0], [math.pi, 0.2, 0.1], [math.pi - 0.3, 0, 0.5]]
    be, top, bot, flux, pos = _setup(ctx, s, B=3, S_pos=S_pos, types=types, widths=widths, rotations=rots)
    sky = _sky(3, S_pos[0])
    _upload_sky(be, sky)
    for t in TILES:
        be.launch_camera_rays(cam, seed=5, tile=t)
    R = be.get_camera_results(cam.resolution)
    ref = _oracle_pixels(s, cam, top, bot, flux, pos, sky, 5, tiles=TILES, source_types=types,
                         source_widths=widths, source_rotations=rots)
    e = rel_err(R["pixel_data"], ref, floor=1e-3)
    assert e.max() <= 1e-5, f"max rel err {e.max():.3e} at {np.unravel_index(e.argmax(), e.shape)}"
    # the planar and sphere sources each fill some pixels on their own
    assert (ref.max(1) > 100).sum() > 3


@pytest.mark.parametrize("periodic", [False, True])
def test_pixel_labels_match_oracle(ctx, periodic):
    s = small_canopy(13, 1500, periodic=periodic)
    be = native.Backend(ctx)
    uu = (np.arange(s.P, dtype=np.uint32) * 7 + 100)
    be.update_geometry(s.transforms, s.types, s.twosided, uuids=uu, periodic=s.periodic)
    be.build_acceleration_structure()
    o = OracleScene(s.transforms, s.types, s.twosided, uuids=uu, periodic=s.periodic)
    for cam in CAMS.values():
        be.zero_camera_pixel_buffers(cam.resolution)
        for t in TILES:
            be.launch_pixel_label_rays(cam, tile=t)
        R = be.get_camera_results(cam.resolution)
        lab, dep = o.pixel_labels(cam)
        assert np.array_equal(R["pixel_labels"], lab)
        hit = lab > 0
        assert hit.mean() > 0.05 and (~hit).any() or periodic
        assert np.all(R["pixel_depths"][~hit] == -1.0)
        assert np.allclose(R["pixel_depths"][hit], dep[hit], rtol=1e-6, atol=0)


def test_camera_errors():
    fresh = native.Context(0)
    try:
        _check_camera_errors(fresh)
    finally:
        fresh.close()


def _check_camera_errors(ctx):
    s = small_canopy(3, 300)
    be, *_ = _setup(ctx, s, B=1)
    cam = CAMS["oblique_pinhole"]
    with pytest.raises(native.RTError, match="rt_update_specular_materials"):
        be.launch_camera_rays(cam, specular=1)
    with pytest.raises(native.RTError, match="exceeds the full resolution"):
        be.launch_camera_rays(cam, tile=(40, 0, 30, 10))
    with pytest.raises(native.RTError, match="sky model has 2 bands"):
        be.update_sky_model(np.zeros((2, 4)), [1, 1], (0, 0, 1), [0, 0], 2.0)
        be.launch_camera_rays(cam)
    with pytest.raises(native.RTError, match="label buffers hold"):
        be.launch_pixel_label_rays(cam)
    # results of an image the buffers do not hold read as zeros
    R = be.get_camera_results((7, 5))
    assert not R["pixel_data"].any() and not R["pixel_labels"].any()
