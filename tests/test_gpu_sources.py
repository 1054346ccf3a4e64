"""Sphere, sun-sphere, rectangle and disk sources (direct_raygen.comp:482-628) on the HIP path.

The direct strength sums of k_direct_src are compared with the oracle's (FP64) within 1e-5 relative
(the rays, their strengths and occlusion decisions are bit-identical; only the summation differs), on a
plain and on a periodic canopy; whole solves within the 1e-4 parity bound; the batched, call-by-call
and serving paths must agree with the plain device solve."""
import math

import numpy as np
import pytest

from helios_amd import native
from oracle.oracle import OracleScene
from tests.backend_schedule import run_band_via_backend
from tests.helpers import rel_err, small_canopy

pytestmark = pytest.mark.gpu

SUN = native.sun_direction(35.0, 120.0)
# collimated, sphere, sun sphere (at 150e9 m), rectangle and disk facing down over the canopy
POS = np.array([SUN, [1.0, 0.5, 4.0], [150e9 * c for c in native.sun_direction(50.0, 250.0)], [0.0, 0.0, 3.0],
                [-1.0, 1.0, 2.5]], np.float32)
TYPES = np.array([0, 1, 2, 3, 4], np.uint32)
WIDTHS = np.array([[0, 0], [0.5, 0.5], [1.391e9, 1.391e9], [2.0, 1.0], [0.6, 0.6]], np.float32)
ROTS = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0], [math.pi, 0.2, 0.3], [math.pi - 0.3, 0.1, 0.0]], np.float32)


def _both(ctx, s):
    g = ctx.scene(s.transforms, s.types, s.twosided, periodic=s.periodic)
    return g, OracleScene.from_arrays(s)


def _request(s, order=None, depth=2, nrays=36):
    k = np.arange(len(TYPES)) if order is None else np.asarray(order)
    P = s.P
    leaf = s.twosided == 1
    rho = np.stack([np.where(leaf, 0.45, 0.30), np.where(leaf, 0.10, 0.15)], 1)
    tau = np.stack([np.where(leaf, 0.40, 0.0), np.where(leaf, 0.05, 0.0)], 1)
    bands = [native.BandSpec("NIR", depth, False, 60.0), native.BandSpec("PAR", depth, False, 120.0)]
    flux = np.array([[350, 500], [2e3, 3e3], [300, 450], [800, 900], [1.5e3, 1e3]], np.float32)
    return native.BandRequest(P, bands, rho=rho, tau=tau, direct_rays=nrays, diffuse_rays=nrays,
                              source_dirs=POS[k], source_flux=flux[k], direct_seed=21,
                              diffuse_seeds=[31 + q for q in range(depth + 1)], source_types=TYPES[k],
                              source_widths=WIDTHS[k], source_rotations=ROTS[k])


@pytest.mark.parametrize("periodic", [False, True])
def test_direct_sums_all_source_types(ctx, periodic):
    s = small_canopy(13, 2500, periodic=periodic)
    g, o = _both(ctx, s)
    sg = g.direct_sums_src(8, 77, POS, TYPES, WIDTHS, ROTS)
    so = o.direct_sums_src(8, 77, POS, TYPES, WIDTHS, ROTS)
    for k in range(len(TYPES)):
        scale = max(np.abs(so[:, k]).max(), 1e-30)
        assert np.abs(so[:, k]).max() > 0.0, f"source {k} lights nothing"
        assert np.max(np.abs(sg[:, k] - so[:, k])) <= 1e-5 * scale, f"source type {TYPES[k]}"


def test_source_order_changes_the_draws(ctx):
    s = small_canopy(13, 1500)
    g, o = _both(ctx, s)
    order = [4, 0, 3, 1]
    sg = g.direct_sums_src(6, 5, POS[order], TYPES[order], WIDTHS[order], ROTS[order])
    so = o.direct_sums_src(6, 5, POS[order], TYPES[order], WIDTHS[order], ROTS[order])
    assert np.max(np.abs(sg - so)) <= 1e-5 * np.abs(so).max()


def test_run_band_all_source_types(ctx):
    s = small_canopy(17, 2000)
    g, o = _both(ctx, s)
    req = _request(s)
    Rg = native.run_band(g, req)
    Ro = o.run_band(req)
    err = rel_err(Rg, Ro, floor=1.0)
    assert err.max() <= 1e-4, f"max rel err {err.max()}"


def test_batch_equals_plain_with_sources(ctx):
    s = small_canopy(17, 2000)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    base = _request(s)
    pos2 = POS.copy()
    pos2[1] = [-0.5, -1.0, 5.0]
    pos2[4] = [0.5, 0.0, 2.0]
    dirs = np.stack([POS, pos2])
    flux = np.stack([base.sflux, base.sflux * 0.5])
    sky = np.stack([base.dflux, base.dflux * 2.0])
    seeds = np.array([21, 22], np.uint32)
    Rb = native.run_band_batch(g, base, dirs, flux, sky, seeds)
    for k in range(2):
        spec = [native.BandSpec(b.label, b.scattering_depth, b.emission, float(q)) for b, q in zip(base.bands, sky[k])]
        rq = native.BandRequest(s.P, spec, rho=base.rho, tau=base.tau, direct_rays=base.direct_rays,
                                diffuse_rays=base.diffuse_rays, source_dirs=dirs[k], source_flux=flux[k],
                                direct_seed=int(seeds[k]), diffuse_seeds=list(base.dseeds), source_types=TYPES,
                                source_widths=WIDTHS, source_rotations=ROTS)
        assert np.array_equal(Rb[k], native.run_band(g, rq))


def test_backend_call_by_call_with_sources(ctx):
    s = small_canopy(19, 1500)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = _request(s)
    be = native.Backend(ctx)
    be.update_geometry(s.transforms, s.types, s.twosided)
    be.build_acceleration_structure()
    be.update_materials(req.rho, req.tau)
    be.update_sources(POS, TYPES, WIDTHS, ROTS)
    Rb = run_band_via_backend(be, req, s.twosided)
    Rd = native.run_band(g, req)
    assert rel_err(Rb, Rd, floor=1.0).max() <= 1e-5


def test_invalid_source_shapes_are_rejected(ctx):
    s = small_canopy(19, 500)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = _request(s)
    bad = native.BandRequest(s.P, req.bands, rho=req.rho, tau=req.tau, direct_rays=16, diffuse_rays=16,
                             source_dirs=POS[:2], source_flux=req.sflux[:2], source_types=[0, 1],
                             source_widths=[[0, 0], [-1.0, 0]])
    with pytest.raises(native.RTError, match="positive finite width"):
        native.run_band(g, bad)
    bad2 = native.BandRequest(s.P, req.bands, rho=req.rho, tau=req.tau, direct_rays=16, diffuse_rays=16,
                              source_dirs=POS[:1], source_flux=req.sflux[:1], source_types=[7])
    with pytest.raises(native.RTError, match="unknown type 7"):
        native.run_band(g, bad2)


def test_server_with_sources(ctx):
    s = small_canopy(19, 1500)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = _request(s)
    srv = native.Server(ctx, slots=2, max_batch_suns=4)
    try:
        m = srv.add_model(g, req)
        t = srv.submit(m, POS, req.sflux, req.dflux, req.direct_seed)
        R, _ = srv.wait(t)
    finally:
        srv.close()
    assert np.array_equal(R, native.run_band(g, req))
