"""Anisotropic diffuse skies (power law, Prague) through the C ABI against the oracle (GPU).

The gather regenerates each sky-bound ray's direction from the pass seed and weighs the sky flux by the
distribution there, so the device and the oracle evaluate fd on the same FP32 directions; acos/pow/exp
come from different libraries, so results agree to the 1e-4 runBand bar rather than bit for bit."""
import math

import numpy as np
import pytest

from helios_amd import native
from oracle.oracle import OracleScene
from tests.backend_schedule import run_band_via_backend
from tests.helpers import rel_err, small_canopy, two_band_request

pytestmark = pytest.mark.gpu

PEAK = [0.42, 0.25, 0.87]


def _unit(p):
    p = np.asarray(p, np.float64)
    return (p / np.linalg.norm(p)).astype(np.float32)


def _sky_request(s, depth=2, peak=PEAK):
    base = two_band_request(s, depth=depth, nrays=36)
    pk = _unit(peak)
    return native.BandRequest(s.P, base.bands, rho=base.rho, tau=base.tau, direct_rays=base.direct_rays,
                              diffuse_rays=base.diffuse_rays, source_dirs=base.sdir, source_flux=base.sflux,
                              direct_seed=base.direct_seed, diffuse_seeds=base.dseeds,
                              sky_extinction=[0.8, 0.0], sky_peak_dir=[pk, pk], sky_norm=[0.35, 0.0],
                              sky_prague=[[0, 1, 1, 0], [2.5, 30.0, 1.6, 0.21]])


@pytest.mark.parametrize("periodic", [False, True])
def test_run_band_anisotropic_sky_vs_oracle(ctx, periodic):
    s = small_canopy(14, 2500, periodic=periodic)
    g = ctx.scene(s.transforms, s.types, s.twosided, periodic=s.periodic)
    o = OracleScene.from_arrays(s)
    req = _sky_request(s)
    Rg = native.run_band(g, req)
    Ro = o.run_band(req)
    assert rel_err(Rg, Ro, floor=1.0).max() <= 1e-4
    # the sky distribution matters: the isotropic solve differs
    iso = _sky_request(s)
    iso.sky_ext = iso.sky_peak = iso.sky_norm = iso.sky_prague = None
    assert rel_err(native.run_band(g, iso), Rg, floor=1.0).max() > 1e-2


def test_batch_sky_peaks_equal_single_solves(ctx):
    s = small_canopy(15, 2000)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    peaks = [[0.42, 0.25, 0.87], [-0.6, 0.3, 0.74], [0.1, -0.8, 0.59]]
    req = _sky_request(s)
    K, B = len(peaks), req.B
    dirs = np.stack([req.sdir] * K)
    flux = np.stack([req.sflux] * K)
    sky = np.stack([req.dflux] * K)
    seeds = [req.direct_seed] * K
    pk = np.stack([_unit(p) for p in peaks])
    Rb = native.run_band_batch(g, req, dirs, flux, sky, seeds, sky_peak_dirs=np.repeat(pk[:, None, :], B, 1))
    for k in range(K):
        one = _sky_request(s, peak=peaks[k])
        Rk = native.run_band(g, one)
        assert rel_err(Rb[k], Rk, floor=1.0).max() <= 1e-6, k


def test_backend_diffuse_launch_with_sky():
    """Call-by-call launches with the power-law fields and the Prague parameters of rt_update_sky_model
    reproduce rt_run_band (own context: the sky model is context state)."""
    ctx = native.Context(0)
    try:
        _backend_sky(ctx)
    finally:
        ctx.close()


def _backend_sky(ctx):
    s = small_canopy(16, 1500)
    req = _sky_request(s, depth=1)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    R_dev = native.run_band(g, req)
    be = native.Backend(ctx)
    be.update_geometry(s.transforms, s.types, s.twosided)
    be.build_acceleration_structure()
    be.update_materials(req.rho, req.tau)
    be.update_sources(req.sdir)
    be.update_sky_model(req.sky_prague, [0.0, 0.0], PEAK, [0.0, 0.0], 2.0)
    R_be = run_band_via_backend(be, req, s.twosided)
    assert rel_err(R_be, R_dev, floor=1e-2).max() <= 1e-5
    g.close()


def test_sky_errors(ctx):
    s = small_canopy(3, 300)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = _sky_request(s)
    req.sky_peak = None
    with pytest.raises(native.RTError, match="diffuse_peak_dir"):
        native.run_band(g, req)
    req = _sky_request(s)
    req.sky_norm = None
    with pytest.raises(native.RTError, match="diffuse_dist_norm"):
        native.run_band(g, req)
