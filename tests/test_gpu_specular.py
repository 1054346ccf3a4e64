"""Specular reflection through the C ABI against the oracle (GPU).

A direct launch with specular on stores radiation_specular [S][P][B]: each source's flux times the
strength sums of its unoccluded rays on both faces of each primitive (R/shaders/direct_raygen.comp:
717-726).  The rays are the oracle's bit for bit, so the buffer matches the oracle's FP64 sums times the
flux within 1e-5 relative (the device sums strengths in FP32).  Camera launches with the Blinn-Phong term
(camera_raygen.comp:362-409) then match the oracle's pixels, given the same radiation_specular, within
1e-5 relative in both modes, for a collimated and a sphere source (the FP32 operations are the same; powf
comes from different libraries)."""
import math

import numpy as np
import pytest

from helios_amd import native
from helios_amd.native import Camera
from oracle.oracle import OracleScene
from tests.helpers import rel_err, small_canopy

pytestmark = pytest.mark.gpu

B = 2
POS = np.asarray([[0.3, 0.2, 0.9], [1.5, -1.0, 6.0]], np.float32)
TYPES = np.asarray([0, 1], np.uint32)
WIDTHS = np.asarray([[0.0, 0.0], [0.4, 0.0]], np.float32)
ROTS = np.zeros((2, 3), np.float32)
CAM = Camera((0.5, -4.0, 2.5), (math.radians(115), math.radians(20)), (48, 36), viewplane_length=0.8,
             fov_aspect=48 / 36, antialiasing_samples=3)


def _sky():
    return dict(params=np.tile(np.asarray([0.0, 1.0, 1.0, 0.0], np.float32), (B, 1)),
                radiance=np.asarray([5.0, 9.0], np.float32), sun=(0.0, 0.0, 1.0), solar=np.zeros(B, np.float32),
                solar_cos=2.0, diffuse_flux=np.zeros(B, np.float32), emission=np.zeros(B, np.uint32))


@pytest.fixture(scope="module")
def lit(ctx):
    s = small_canopy(3, 300)
    be = native.Backend(ctx)
    be.update_geometry(s.transforms, s.types, s.twosided)
    be.build_acceleration_structure()
    be.update_materials(np.full((s.P, B), 0.1), np.full((s.P, B), 0.05))
    be.update_sources(POS, types=TYPES, widths=WIDTHS, rotations=ROTS)
    be.zero_radiation_buffers(B)
    rng = np.random.default_rng(11)
    flux = rng.uniform(100, 900, (2, B)).astype(np.float32)
    be.upload_source_fluxes(flux)
    be.launch_direct_rays(64, 9, B, specular=1)
    spec = be.get_radiation_specular()
    top = rng.uniform(0, 400, (s.P, B)).astype(np.float32)
    bot = rng.uniform(0, 100, (s.P, B)).astype(np.float32)
    be.upload_radiation_out(top, bot)
    sky = _sky()
    be.update_sky_model(sky["params"], sky["radiance"], sky["sun"], sky["solar"], sky["solar_cos"],
                        sky["diffuse_flux"], sky["emission"])
    return s, be, flux, spec, top, bot, sky, OracleScene.from_arrays(s), rng


def test_direct_launch_stores_delivered_radiation(lit):
    s, be, flux, spec, *_, o, rng = lit
    sums = o.direct_sums_src(8, 9, POS, TYPES, WIDTHS, ROTS)  # [P][S][2]
    ref = flux[:, None, :].astype(np.float64) * (sums[:, :, 0] + sums[:, :, 1]).T[:, :, None]
    assert spec.shape == (2, s.P, B)
    assert (ref > 0).mean() > 0.3 and np.all((ref == 0) == (spec == 0))
    e = rel_err(spec, ref, floor=1e-3 * ref.max())
    assert e.max() <= 1e-5, f"max rel err {e.max():.3e}"


def test_camera_specular_matches_oracle(lit):
    s, be, flux, spec, top, bot, sky, o, rng = lit
    exp = np.where(rng.random(s.P) < 0.8, rng.uniform(2, 60, s.P), -1.0).astype(np.float32)
    scale = rng.uniform(0.1, 2.0, s.P).astype(np.float32)
    w = rng.uniform(0.3, 1.2, (2, B)).astype(np.float32)
    be.update_specular_materials(exp, scale)
    be.upload_source_fluxes_cam(w)
    fx, fy = CAM.resolution
    plain = None
    for mode in (0, 1, 2):
        be.zero_camera_pixel_buffers(CAM.resolution)
        be.launch_camera_rays(CAM, seed=6, specular=mode)
        R = be.get_camera_results(CAM.resolution)["pixel_data"]
        sp = None if mode == 0 else dict(spec=spec, weight=w, exponent=exp, scale=scale, mode=mode)
        ref = o.camera_pixels(CAM, top, bot, flux, POS, sky, seed=6, source_types=TYPES, source_widths=WIDTHS,
                              source_rotations=ROTS, specular=sp).T
        e = rel_err(R, ref, floor=1e-3)
        assert e.max() <= 1e-5, f"mode {mode}: max rel err {e.max():.3e}"
        if mode == 0:
            plain = R
        else:
            # the term only adds: a visible share of the hit pixels brightens
            assert np.all(R >= plain - 1e-6 * np.abs(plain)) and (R > plain * 1.01).mean() > 0.05


def test_specular_errors_and_inert_cases(lit):
    s, be, flux, spec, top, bot, sky, o, rng = lit
    be.zero_camera_pixel_buffers(CAM.resolution)
    be.launch_camera_rays(CAM, seed=6)
    plain = be.get_camera_results(CAM.resolution)["pixel_data"]
    # no positive exponent: the specular launch reproduces the plain image bit for bit
    be.update_specular_materials(np.full(s.P, -1.0, np.float32))
    be.upload_source_fluxes_cam(np.ones((2, B), np.float32))
    be.zero_camera_pixel_buffers(CAM.resolution)
    be.launch_camera_rays(CAM, seed=6, specular=2)
    assert np.array_equal(be.get_camera_results(CAM.resolution)["pixel_data"], plain)
    with pytest.raises(native.RTError, match="must be 0, 1"):
        be.launch_camera_rays(CAM, specular=3)
    with pytest.raises(native.RTError, match="S\\*B_launch"):
        be.upload_source_fluxes_cam(np.ones(3, np.float32))
    with pytest.raises(native.RTError, match="primitives"):
        be.update_specular_materials(np.ones(s.P + 1, np.float32))
    # a new launch set clears radiation_specular, and the camera weights follow the band count
    be.zero_radiation_buffers(B)
    assert np.all(be.get_radiation_specular() == 0.0)
    be.zero_radiation_buffers(1)
    be.upload_source_fluxes(np.ones((2, 1), np.float32))
    be.update_sky_model(np.zeros((1, 4), np.float32), [1.0], (0.0, 0.0, 1.0), [0.0], 2.0)
    with pytest.raises(native.RTError, match="rt_upload_source_fluxes_cam"):
        be.launch_camera_rays(CAM, specular=1)


def test_specular_needs_materials(ctx):
    s = small_canopy(3, 200)
    be = native.Backend(ctx)
    be.update_geometry(s.transforms, s.types, s.twosided)
    be.build_acceleration_structure()
    be.update_materials(np.full((s.P, 1), 0.1), np.full((s.P, 1), 0.05))
    be.update_sources(POS[:1])
    be.zero_radiation_buffers(1)
    be.upload_source_fluxes(np.ones((1, 1), np.float32))
    with pytest.raises(native.RTError, match="rt_update_specular_materials"):
        be.launch_camera_rays(CAM, specular=1)
