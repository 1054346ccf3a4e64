"""Request server (rt_server_*) through the C ABI: continuous batching with memory-footprint admission.

The contract: every request's result equals a solve of that sun alone (rt_run_band with the model's
settings) within 1e-5 relative (floor 1 W/m^2; batched and plain runs trace identical rays and differ
only in summation order), whatever mix it ran in."""
import time

import numpy as np
import pytest

from helios_amd import native, scenes
from helios_amd import workloads as W
from tests.helpers import config1_request, rel_err, small_canopy, two_band_request

pytestmark = pytest.mark.gpu


def _emission_model(s):
    P = s.P
    leaf = s.twosided == 1
    rho = np.stack([np.where(leaf, 0.05, 0.1), np.where(leaf, 0.10, 0.15)], 1)
    tau = np.stack([np.where(leaf, 0.0, 0.0), np.where(leaf, 0.05, 0.0)], 1)
    eps = np.stack([np.where(leaf, 0.95, 0.9), np.ones(P)], 1)
    temp = np.where(leaf, 298.0, 293.0).astype(np.float32)
    bands = [native.BandSpec("LW", 1, True, 350.0), native.BandSpec("PAR", 2, False, 100.0)]
    return native.BandRequest(P, bands, rho=rho, tau=tau, eps=eps, temperature=temp, direct_rays=36,
                              diffuse_rays=36, source_dirs=[[0, 0, 1]], source_flux=[[0.0, 500.0]], direct_seed=1,
                              diffuse_seeds=[91, 92, 93])


def _setup(ctx):
    s1 = small_canopy(5, 2500)
    s2 = scenes.config1(1)
    s3 = small_canopy(7, 1500, periodic=True)
    mods = [(s1, two_band_request(s1, depth=2, nrays=36)),
            (s2, config1_request(s2.P, depth=3, nrays=49)),
            (s3, _emission_model(s3))]
    out = []
    for s, req in mods:
        g = ctx.scene(s.transforms, s.types, s.twosided, periodic=s.periodic)
        out.append((s, g, req))
    return out


def _random_sun(rng, req):
    z = np.radians(rng.uniform(10, 80))
    a = np.radians(rng.uniform(0, 360))
    d = np.array([[np.sin(z) * np.sin(a), np.sin(z) * np.cos(a), np.cos(z)]], np.float32)
    flux = (req.sflux * rng.uniform(0.5, 1.5)).astype(np.float32)
    sky = (req.dflux * rng.uniform(0.5, 1.5)).astype(np.float32)
    return d, flux, sky, int(rng.integers(1, 2**31))


def test_server_results_equal_alone_in_a_mix(ctx):
    models = _setup(ctx)
    srv = native.Server(ctx, slots=3, max_batch_suns=8)
    ids = [srv.add_model(g, req) for (_, g, req) in models]
    rng = np.random.default_rng(11)
    jobs = []
    for i in range(40):
        m = int(rng.integers(len(models)))
        sun = _random_sun(rng, models[m][2])
        jobs.append((m, sun, srv.submit(ids[m], *sun)))
        if rng.random() < 0.25:
            time.sleep(0.003)
    worst = 0.0
    for m, (d, f, q, seed), t in jobs:
        R, lat = srv.wait(t)
        assert lat > 0
        s, g, req = models[m]
        alone = native.run_band(g, W.single_sun_request(req, d, f, q, seed))
        worst = max(worst, float(rel_err(R, alone, floor=1.0).max()))
    st = srv.stats()
    srv.close()
    assert worst <= 1e-5, worst
    assert st["completed"] == 40 and st["failed"] == 0 and st["queued"] == 0
    assert st["batches"] < 40 and st["max_batch"] > 1, st  # requests were batched
    assert 0 < st["p50_ms"] <= st["p99_ms"]


def test_server_memory_budget_limits_batches(ctx):
    s = small_canopy(5, 20000)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = two_band_request(s, depth=2, nrays=36)
    # room for one batch of three suns: admission must shrink batches to fit
    budget = native.solve_footprint(s.P, req, 3) + 4096
    assert native.solve_footprint(s.P, req, 4) > budget
    srv = native.Server(ctx, slots=2, max_batch_suns=16, memory_budget=budget)
    mid = srv.add_model(g, req)
    rng = np.random.default_rng(3)
    suns = [_random_sun(rng, req) for _ in range(12)]
    tickets = [srv.submit(mid, *sun) for sun in suns]
    for sun, t in zip(suns, tickets):
        R, _ = srv.wait(t)
        alone = native.run_band(g, W.single_sun_request(req, *sun))
        assert rel_err(R, alone, floor=1.0).max() <= 1e-5
    st = srv.stats()
    srv.close()
    assert st["completed"] == 12
    assert st["peak_workspace_bytes"] <= budget, st
    assert 1 < st["max_batch"] <= 3, st


def test_server_rejects_what_cannot_fit(ctx):
    s = scenes.config1(1)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = config1_request(s.P, depth=3, nrays=49)
    srv = native.Server(ctx, slots=1, memory_budget=1 << 16)
    mid = srv.add_model(g, req)
    t = srv.submit(mid, *_random_sun(np.random.default_rng(0), req))
    with pytest.raises(native.RTError, match="memory budget"):
        srv.wait(t)
    with pytest.raises(native.RTError, match="unknown model"):
        srv.submit(7, [[0, 0, 1]], [[1.0]], [1.0], 1)
    st = srv.stats()
    srv.close()
    assert st["failed"] == 1
