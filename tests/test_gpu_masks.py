"""Texture-masked primitives through the C ABI against the oracle (GPU).

Masks change which hits count (a transparent texel lets the ray pass) and where direct and diffuse rays
start (redraws off transparent texels, dropped rays), on the BVH, BVH4-packet and occlusion-grid paths
alike.  Bars: diffuse hit words bit-exact; direct sums within 1e-5 of the FP64 oracle; runBand within
1e-4; camera labels bit-exact; grid and BVH direct passes equal."""
import math

import numpy as np
import pytest

from helios_amd import native, scenes
from helios_amd.native import Camera
from oracle.oracle import OracleScene
from tests.backend_schedule import run_band_via_backend
from tests.helpers import rel_err, small_canopy, two_band_request

pytestmark = pytest.mark.gpu


def _masks(s, seed):
    """Random masks (6 of 3x3 .. 9x7 texels, ~55 % opaque) on ~70 % of the primitives; random UVs; a third of
    the masked patches use custom UVs."""
    rng = np.random.default_rng(seed)
    masks = [(rng.uniform(size=(int(rng.integers(3, 8)), int(rng.integers(3, 10)))) < 0.55).astype(np.uint8)
             for _ in range(6)]
    P = s.P
    mid = np.where(rng.uniform(size=P) < 0.7, rng.integers(0, 6, P), -1).astype(np.int32)
    uv = rng.uniform(0, 1, (P, 4, 2)).astype(np.float32)
    uv[:, 1, 0] = np.maximum(uv[:, 1, 0], uv[:, 0, 0])
    uvid = np.where(rng.uniform(size=P) < 0.33, np.arange(P), -1).astype(np.int32)
    return masks, mid, uv, uvid


def _both(ctx, s, seed=1):
    masks, mid, uv, uvid = _masks(s, seed)
    g = ctx.scene(s.transforms, s.types, s.twosided, periodic=s.periodic)
    g.set_masks(masks, mid, uv, uvid)
    o = OracleScene.from_arrays(s)
    o.set_masks(masks, mid, uv, uvid)
    return g, o, (masks, mid, uv, uvid)


SCENES = {
    "patches": lambda: small_canopy(5, 2500),
    "patches_periodic": lambda: small_canopy(6, 2500, periodic=True),
    "mixed": lambda: scenes.random_canopy(7, 6000),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_diffuse_hits_bit_exact_with_masks(ctx, name):
    s = SCENES[name]()
    g, o, _ = _both(ctx, s)
    hg = g.diffuse_hits(6, 77)
    ho = o.diffuse_hits(6, 77)
    assert np.array_equal(hg, ho), f"{np.count_nonzero(hg != ho)} hit words differ"
    assert (ho == 0xFFFFFFFE).sum() > 0  # some rays dropped by their masks


@pytest.mark.parametrize("name", sorted(SCENES))
def test_direct_sums_with_masks(ctx, name):
    s = SCENES[name]()
    g, o, _ = _both(ctx, s)
    src = [native.sun_direction(25, 40), native.sun_direction(65, 220)]
    sg = g.direct_sums(7, 13, src)
    so = o.direct_sums(7, 13, src)
    assert np.max(np.abs(sg - so)) <= 1e-5 * max(1.0, np.abs(so).max())


def test_direct_grid_equals_bvh_with_masks(ctx):
    s = small_canopy(23, 4000)
    g, _, _ = _both(ctx, s, seed=4)
    src = [native.sun_direction(z, a) for z, a in ((5, 0), (35, 120), (70, 250))]
    try:
        ctx.set_direct_grids(True)
        sg = g.direct_sums(9, 5, src)
        ctx.set_direct_grids(False)
        sb = g.direct_sums(9, 5, src)
    finally:
        ctx.set_direct_grids(True)
    assert np.max(np.abs(sg - sb) / np.maximum(np.abs(sb), 1e-30)) <= 1e-6


@pytest.mark.parametrize("name", ["patches", "mixed"])
def test_run_band_with_masks(ctx, name):
    s = SCENES[name]()
    g, o, _ = _both(ctx, s, seed=2)
    req = two_band_request(s, depth=2, nrays=25)
    Rg = native.run_band(g, req)
    Ro = o.run_band(req)
    err = rel_err(Rg, Ro, floor=1.0)
    assert err.max() <= 1e-4, f"max rel err {err.max()}"


def test_backend_masks_equal_scene_masks(ctx):
    """rt_update_texture_masks + build (call by call) == rt_scene_set_masks + rt_run_band."""
    s = small_canopy(8, 1500)
    g, _, (masks, mid, uv, uvid) = _both(ctx, s, seed=3)
    req = two_band_request(s, depth=1, nrays=25)
    R_dev = native.run_band(g, req)
    be = native.Backend(ctx)
    be.update_geometry(s.transforms, s.types, s.twosided, periodic=s.periodic)
    be.update_texture_masks(masks, mid, uv, uvid)
    be.build_acceleration_structure()
    be.update_materials(req.rho, req.tau)
    be.update_sources(req.sdir)
    R_be = run_band_via_backend(be, req, s.twosided)
    assert rel_err(R_be, R_dev, floor=1e-2).max() <= 1e-5
    # masks off again: a plain build differs
    be.update_texture_masks([], np.full(s.P, -1))
    be.build_acceleration_structure()
    be.update_materials(req.rho, req.tau)
    be.update_sources(req.sdir)
    R_plain = run_band_via_backend(be, req, s.twosided)
    assert rel_err(R_plain, R_dev, floor=1e-2).max() > 1e-3


def test_camera_labels_with_masks(ctx):
    s = scenes.random_canopy(9, 4000)
    masks, mid, uv, uvid = _masks(s, 5)
    uu = np.arange(s.P, dtype=np.uint32) + 10
    be = native.Backend(ctx)
    be.update_geometry(s.transforms, s.types, s.twosided, uuids=uu)
    be.update_texture_masks(masks, mid, uv, uvid)
    be.build_acceleration_structure()
    o = OracleScene(s.transforms, s.types, s.twosided, uuids=uu)
    o.set_masks(masks, mid, uv, uvid)
    # looking down on the canopy (about 1.4 m wide, 2 m tall) from 4 m
    cam = Camera((0.0, 0.0, 4.0), (math.pi, 0.0), (48, 32), viewplane_length=2.0, fov_aspect=1.5)
    be.zero_camera_pixel_buffers(cam.resolution)
    be.launch_pixel_label_rays(cam)
    R = be.get_camera_results(cam.resolution)
    lab, dep = o.pixel_labels(cam)
    assert np.array_equal(R["pixel_labels"], lab)
    assert (lab > 0).mean() > 0.5


def test_mask_errors(ctx):
    s = small_canopy(2, 200)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    with pytest.raises(native.RTError, match="names mask 3"):
        g.set_masks([np.ones((2, 2))], np.full(s.P, 3))
    with pytest.raises(native.RTError, match="size"):
        g.set_masks([np.ones((0, 2))], np.zeros(s.P))


def test_direct_sums_all_source_types_with_masks(ctx):
    """Masked scene lit by collimated, sphere, sun-sphere, rectangle and disk sources: origin redraws
    continue the RNG stream before the per-source draws; occlusion honours transparent texels."""
    from tests.test_gpu_sources import POS, ROTS, TYPES, WIDTHS
    s = scenes.random_canopy(12, 5000)
    g, o, _ = _both(ctx, s, seed=6)
    sg = g.direct_sums_src(6, 41, POS, TYPES, WIDTHS, ROTS)
    so = o.direct_sums_src(6, 41, POS, TYPES, WIDTHS, ROTS)
    for k in range(len(TYPES)):
        scale = max(np.abs(so[:, k]).max(), 1e-30)
        assert np.abs(so[:, k]).max() > 0.0, f"source {k} lights nothing"
        assert np.max(np.abs(sg[:, k] - so[:, k])) <= 1e-5 * scale, f"source type {TYPES[k]}"


def test_camera_pixels_with_masks(ctx):
    """Camera radiance of a masked canopy: hits behind transparent texels show the primitives behind."""
    s = scenes.random_canopy(13, 4000)
    masks, mid, uv, uvid = _masks(s, 8)
    B = 2
    be = native.Backend(ctx)
    be.update_geometry(s.transforms, s.types, s.twosided)
    be.update_texture_masks(masks, mid, uv, uvid)
    be.build_acceleration_structure()
    be.update_materials(np.full((s.P, B), 0.1), np.full((s.P, B), 0.05))
    be.update_sources([[0.3, 0.2, 0.9]])
    be.zero_radiation_buffers(B)
    rng = np.random.default_rng(2)
    top = rng.uniform(0, 300, (s.P, B)).astype(np.float32)
    bot = rng.uniform(0, 80, (s.P, B)).astype(np.float32)
    be.upload_radiation_out(top, bot)
    be.upload_source_fluxes([[500.0, 700.0]])
    sky = dict(params=np.tile(np.asarray([0, 1, 1, 0], np.float32), (B, 1)), radiance=np.array([10.0, 20.0], np.float32),
               sun=(0.0, 0.0, 1.0), solar=np.zeros(B, np.float32), solar_cos=2.0, diffuse_flux=np.zeros(B, np.float32),
               emission=np.zeros(B, np.uint32))
    be.update_sky_model(sky["params"], sky["radiance"], sky["sun"], sky["solar"], sky["solar_cos"])
    cam = Camera((0.0, 0.0, 4.0), (math.pi, 0.0), (48, 32), viewplane_length=2.0, fov_aspect=1.5,
                 antialiasing_samples=2)
    be.launch_camera_rays(cam, seed=9)
    R = be.get_camera_results(cam.resolution)
    o = OracleScene.from_arrays(s)
    o.set_masks(masks, mid, uv, uvid)
    ref = o.camera_pixels(cam, top, bot, [[500.0, 700.0]], [[0.3, 0.2, 0.9]], sky, seed=9).T
    assert rel_err(R["pixel_data"], ref, floor=1e-3).max() <= 1e-5
