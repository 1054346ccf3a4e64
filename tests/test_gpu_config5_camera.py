"""The config-5 camera at its benchmarked size, through the C ABI: the bench's 1024 x 1024 x 64-sample view of
the 5x tiled orchard (5,008,000 primitives, 3 bands, a collimated sun with its solar disk, a Prague sky).
Three 32 x 32 tiles (canopy, ground, the upper edge of the view) are rendered and compared with the oracle:
pixel radiance within 1e-5 relative (the device sums each pixel's FP32 samples in order, the oracle in FP64);
pixel labels of the whole image bit-exact and depths within 1e-6."""
import math

import numpy as np
import pytest

import bench
from helios_amd import native, scenes
from oracle.oracle import OracleScene
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

B = 3
TILES = [(496, 600, 32, 32), (200, 960, 32, 32), (800, 40, 32, 32)]


@pytest.fixture(scope="module")
def c5cam(ctx):
    s = scenes.orchard(3, tiles=5, jitter=0.05)
    be = native.Backend(ctx)
    be.update_geometry(s.transforms, s.types, s.twosided)
    be.build_acceleration_structure()
    be.update_materials(np.full((s.P, B), 0.1, np.float32), np.full((s.P, B), 0.05, np.float32))
    sun = native.sun_direction(40.0, 150.0)
    be.update_sources([sun])
    be.zero_radiation_buffers(B)
    rng = np.random.default_rng(7)
    top = rng.uniform(0, 300, (s.P, B)).astype(np.float32)
    bot = rng.uniform(0, 100, (s.P, B)).astype(np.float32)
    be.upload_radiation_out(top, bot)
    flux = np.array([[400.0, 300.0, 500.0]], np.float32)
    be.upload_source_fluxes(flux)
    sky = dict(params=np.tile(np.array([2.0, 30.0, 1.5, 0.25], np.float32), (B, 1)),
               radiance=np.array([20.0, 15.0, 30.0], np.float32), sun=tuple(float(v) for v in sun),
               solar=np.array([4e3, 3e3, 5e3], np.float32), solar_cos=math.cos(math.radians(0.27)),
               diffuse_flux=np.zeros(B, np.float32), emission=np.zeros(B, np.uint32))
    be.update_sky_model(sky["params"], sky["radiance"], sky["sun"], sky["solar"], sky["solar_cos"],
                        sky["diffuse_flux"], sky["emission"])
    cam = bench.orchard_camera(s, 1024, 64)
    return s, be, cam, top, bot, flux, np.asarray([sun], np.float32), sky, OracleScene.from_arrays(s)


def test_config5_camera_tiles_match_oracle(c5cam):
    s, be, cam, top, bot, flux, pos, sky, o = c5cam
    be.zero_camera_pixel_buffers(cam.resolution)
    for t in TILES:
        be.launch_camera_rays(cam, seed=2, tile=t)
    R = be.get_camera_results(cam.resolution)["pixel_data"]
    fx, fy = cam.resolution
    ref = np.zeros((B, fx * fy))
    for t in TILES:
        o.camera_pixels(cam, top, bot, flux, pos, sky, seed=2, tile=t, pixels=ref)
    ref = ref.T
    px = np.zeros(fx * fy, bool)
    for (x, y, w, h) in TILES:
        for j in range(y, y + h):
            px[j * fx + x:j * fx + x + w] = True
    e = rel_err(R[px], ref[px], floor=1e-3)
    assert e.max() <= 1e-5, f"max rel err {e.max():.3e}"
    assert (ref[px] > 0).mean() > 0.9
    assert np.all(R[~px] == 0.0)  # nothing outside the tiles


def test_config5_pixel_labels_match_oracle(c5cam):
    s, be, cam, top, bot, flux, pos, sky, o = c5cam
    be.zero_camera_pixel_buffers(cam.resolution)
    be.launch_pixel_label_rays(cam)
    R = be.get_camera_results(cam.resolution)
    lab, dep = o.pixel_labels(cam)
    assert np.array_equal(R["pixel_labels"], lab), f"{np.count_nonzero(R['pixel_labels'] != lab)} labels differ"
    hit = lab > 0
    assert 0.3 < hit.mean() < 1.0
    assert np.allclose(R["pixel_depths"][hit], dep[hit], rtol=1e-6, atol=0)
