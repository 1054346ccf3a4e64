"""Translucent (glass) covers on the HIP path (through the C ABI): the reference's glass test scenes against
their closed forms and the oracle, and a canopy under a glass roof with two bands and two suns against the
oracle (1e-4) and against its single-sun solves."""
import math

import numpy as np
import pytest

from helios_amd import native, scenes
from helios_amd import workloads as W
from oracle.oracle import OracleScene, glass_tra
from tests import glass_cases as gc
from tests.backend_schedule import run_band_via_backend
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

F = gc.F_SUN


def _both(ctx, sc, req):
    T, ty, tw = sc
    Rg = native.run_band(ctx.scene(T, ty, tw), req)
    Ro = OracleScene(T, ty, tw).run_band(req)
    assert rel_err(Rg, Ro, floor=1.0).max() <= 1e-4
    return Rg


def test_normal_and_oblique_transmittance_gpu(ctx):
    R = _both(ctx, gc.cover_over_receiver(), gc.request(2, [1, 0]))
    assert abs(R[1, 0] - F * glass_tra(1.0, 1.5)[0]) <= 1e-4 * F
    th = math.radians(30.0)
    R = _both(ctx, gc.cover_over_receiver(8.0), gc.request(2, [1, 0], sun=(math.sin(th), 0.0, math.cos(th)),
                                                             direct_rays=20000))
    assert abs(R[1, 0] - F * math.cos(th) * gc.fresnel_slab(math.cos(th), 1.5)[0]) <= 1e-4 * F


def test_double_layer_window_and_opaque_cases_gpu(ctx):
    R = _both(ctx, gc.double_layer(), gc.request(3, [1, 1, 0]))
    assert abs(R[2, 0] - F * glass_tra(1.0, 1.5)[0] ** 2) <= 1e-4 * F
    th = math.radians(45.0)
    R = _both(ctx, gc.window(th), gc.request(5, [0, 0, 1, 0, 0], sun=(math.sin(th), 0.0, math.cos(th)),
                                             direct_rays=40000))
    assert abs(R[3, 0] - F * math.cos(th) * gc.fresnel_slab(math.cos(th), 1.5)[0]) <= 1e-4 * F
    assert R[4, 0] == 0.0
    R = _both(ctx, gc.cover_over_receiver(), gc.request(2, [[1, 0], [0, 0]], bands=("LW", "SW")))
    assert R[1, 0] == 0.0 and R[1, 1] == 0.0


def test_cover_shares_and_sky_gpu(ctx):
    tau = glass_tra(1.0, 1.5, 0.05)[0]
    R = _both(ctx, gc.cover_over_receiver(), gc.request(2, [1, 0], KL=0.05, depth=0))
    assert abs(R[0, 0] - F * (2.0 - tau)) <= 1e-4 * F
    sc = gc.scene([gc.horiz((0, 0, 1), 4000, 4000), gc.horiz((0, 0, 0), 2, 2, tw=0)])
    R = _both(ctx, sc, gc.request(2, [1, 0], F=0.0, sky=100.0, diffuse_rays=250000, depth=0))
    assert abs(R[1, 0] - 100.0 * gc.sky_through_cover()) <= 0.2


def _greenhouse():
    """A 3,000-leaf canopy (4 x 4 m) under a two-panel glass roof at 2 m and a glass wall on one side."""
    s = scenes.patch_canopy(9, 3000, 0.06, 4.0, 1.5, 6)
    roof = gc.scene([gc.horiz((-1.5, 0, 2.0), 3.0, 7.0), gc.horiz((1.5, 0, 2.0), 3.0, 7.0),
                     ((3.0, 0, 1.0), (0, 7.0, 0), (0, 0, 2.0), 1)])
    T = np.concatenate([s.transforms.reshape(-1, 16), roof[0]])
    ty = np.concatenate([s.types, roof[1]])
    tw = np.concatenate([s.twosided, roof[2]])
    return T.astype(np.float32), ty, tw, s.P


def _greenhouse_request(tw, P0, P, schedule=0):
    leaf = np.asarray(tw) == 1  # two-sided leaves; the one-sided ground and the (two-sided) glass get own values
    rho = np.stack([np.where(leaf, 0.45, 0.2), np.where(leaf, 0.10, 0.15)], 1)
    tau = np.stack([np.where(leaf, 0.40, 0.0), np.where(leaf, 0.05, 0.0)], 1)
    rho[P0:] = 0.0
    tau[P0:] = 0.0
    glass = np.zeros((P, 2), np.uint8)
    glass[P0:] = 1
    gn = np.full((P, 2), 1.5, np.float32)
    gkl = np.zeros((P, 2), np.float32)
    gkl[P0:, 1] = 0.08  # more absorbing in PAR
    bands = [native.BandSpec("NIR", 2, False, 60.0), native.BandSpec("PAR", 2, False, 120.0)]
    return native.BandRequest(P, bands, rho=rho, tau=tau, direct_rays=36, diffuse_rays=36,
                              source_dirs=[native.sun_direction(35.0, 120.0)], source_flux=[[350.0, 500.0]],
                              direct_seed=21, diffuse_seeds=[31, 32, 33], schedule=schedule, is_glass=glass,
                              glass_n=gn, glass_KL=gkl)


def test_greenhouse_call_by_call(ctx):
    """The reference's pass loop driven call by call (tests/backend_schedule.py): covers pass the direct and
    diffuse rays of the call-by-call launches as they do on the whole-solve path (same rays; per-face launches
    and deposits change only the FP32 summation order: 1e-5), and the result matches the oracle (1e-4)."""
    T, ty, tw, P0 = _greenhouse()
    P = T.shape[0]
    req = _greenhouse_request(tw, P0, P)
    Rg = native.run_band(ctx.scene(T, ty, tw), req)
    be = native.Backend(ctx)
    be.update_geometry(T, ty, tw)
    be.build_acceleration_structure()
    be.update_materials(req.rho, req.tau)
    be.update_glass_materials(req.is_glass, req.glass_n, req.glass_KL)
    be.update_sources(req.sdir)
    Rb = run_band_via_backend(be, req, tw)
    e = rel_err(Rb, Rg, floor=1.0)
    assert e.max() <= 1e-5, f"max rel err {e.max():.3e}"
    Ro = OracleScene(T, ty, tw).run_band(req)
    assert rel_err(Rb, Ro, floor=1.0).max() <= 1e-4
    # the covers take a share (their own radiation_flux), and without them the same launches differ
    assert np.all(Rb[P0:].sum(0) > 0)
    be.update_glass_materials(None)
    R0 = run_band_via_backend(be, req, tw)
    assert np.abs(R0[:P0] - Rb[:P0]).max() > 1.0


@pytest.mark.parametrize("schedule", [0, 1])
def test_greenhouse_against_oracle(ctx, schedule):
    T, ty, tw, P0 = _greenhouse()
    P = T.shape[0]
    req = _greenhouse_request(tw, P0, P, schedule)
    Rg = native.run_band(ctx.scene(T, ty, tw), req)
    Ro = OracleScene(T, ty, tw).run_band(req)
    assert rel_err(Rg, Ro, floor=1.0).max() <= 1e-4
    # the roof shades the canopy: it absorbs less than the same canopy in the open, but far from nothing
    Ropen = native.run_band(ctx.scene(T[:P0], ty[:P0], tw[:P0]), _open(req, P0))
    frac = Rg[:P0].sum(0) / Ropen.sum(0)
    assert np.all((0.5 < frac) & (frac < 0.99)), frac


def _open(req, P0):
    return native.BandRequest(P0, req.bands, rho=req.rho[:P0], tau=req.tau[:P0], direct_rays=36, diffuse_rays=36,
                              source_dirs=req.sdir, source_flux=req.sflux, direct_seed=21, diffuse_seeds=[31, 32, 33],
                              schedule=req.schedule)


def test_greenhouse_batch_equals_single_suns(ctx):
    T, ty, tw, P0 = _greenhouse()
    P = T.shape[0]
    g = ctx.scene(T, ty, tw)
    base = _greenhouse_request(tw, P0, P)
    dirs = np.stack([[native.sun_direction(30.0, 100.0)], [native.sun_direction(55.0, 220.0)]]).astype(np.float32)
    flux = np.array([[[350.0, 500.0]], [[200.0, 300.0]]], np.float32)
    sky = np.array([[60.0, 120.0], [80.0, 90.0]], np.float32)
    seeds = [7, 8]
    Rb = native.run_band_batch(g, base, dirs, flux, sky, seeds)
    for k in range(2):
        single = W.single_sun_request(base, dirs[k], flux[k], sky[k], seeds[k])
        single.is_glass, single.glass_n, single.glass_KL = base.is_glass, base.glass_n, base.glass_KL
        Rp = native.run_band(g, single)
        assert rel_err(Rb[k], Rp, floor=1.0).max() <= 1e-5


def test_greenhouse_served(ctx):
    """The request server with a glass model: requests batched at most 4 suns (8 channels) at a time, each
    result equal to its single-sun solve within 1e-5 (the covers' shares are FP64 atomic sums)."""
    T, ty, tw, P0 = _greenhouse()
    P = T.shape[0]
    g = ctx.scene(T, ty, tw)
    base = _greenhouse_request(tw, P0, P)
    srv = native.Server(ctx, slots=2, max_batch_suns=16)
    try:
        mid = srv.add_model(g, base)
        rng = np.random.default_rng(3)
        jobs = []
        for k in range(10):
            d = np.asarray([native.sun_direction(float(rng.uniform(15, 70)), float(rng.uniform(60, 300)))],
                           np.float32)
            f = rng.uniform(200, 600, (1, 2)).astype(np.float32)
            q = rng.uniform(40, 120, 2).astype(np.float32)
            seed = int(rng.integers(1, 2**31))
            jobs.append((d, f, q, seed, srv.submit(mid, d, f, q, seed)))
        for d, f, q, seed, t in jobs:
            R, _ = srv.wait(t)
            single = W.single_sun_request(base, d, f, q, seed)
            single.is_glass, single.glass_n, single.glass_KL = base.is_glass, base.glass_n, base.glass_KL
            assert rel_err(R, native.run_band(g, single), floor=1.0).max() <= 1e-5
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == 10 and st["failed"] == 0 and 1 <= st["max_batch"] <= 4, st


def test_glass_limits(ctx):
    T, ty, tw, P0 = _greenhouse()
    P = T.shape[0]
    g = ctx.scene(T, ty, tw)
    base = _greenhouse_request(tw, P0, P)
    K = 5  # 5 suns x 2 bands = 10 channels > 8
    dirs = np.tile(np.array([[native.sun_direction(30.0, 100.0)]], np.float32), (K, 1, 1))
    with pytest.raises(native.RTError, match="at most 8 channels"):
        native.run_band_batch(g, base, dirs, np.ones((K, 1, 2), np.float32), np.ones((K, 2), np.float32),
                              list(range(K)))
    # the FP64 energy bookkeeping follows the hit-table passes, which a solve with covers does not run
    native.set_energy_accounting(ctx, True)
    try:
        native.run_band(g, base)
        with pytest.raises(native.RTError, match="translucent covers"):
            native.energy_fp64(ctx, 2)
    finally:
        native.set_energy_accounting(ctx, False)
    bad = _greenhouse_request(tw, P0, P)
    bad.glass_n[P0, 0] = 0.5
    with pytest.raises(native.RTError, match="refractive index"):
        native.run_band(g, bad)
