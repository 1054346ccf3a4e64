"""HIP path vs the CPU oracle, through the C ABI (libhelios_rt.so).

Bars: hit words and castRays IDs bit-exact (ties resolve to the lower primitive index on both
sides); FP32 compensated per-primitive sums within 1e-5 relative of the FP64 oracle; radiation_flux
within 1e-4 relative (the parity bound of the flux)."""
import numpy as np
import pytest

from helios_amd import native, scenes
from oracle.oracle import OracleScene
from tests.helpers import config1_request, rel_err, small_canopy, two_band_request

pytestmark = pytest.mark.gpu


def _both(ctx, s):
    g = ctx.scene(s.transforms, s.types, s.twosided, periodic=s.periodic)
    o = OracleScene.from_arrays(s)
    return g, o


def test_scene_info(ctx):
    s = scenes.config1(1)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    info = g.info()
    assert info["nprims"] == 1601 and info["nleaves"] == 1601
    assert 0 < info["depth"] < 64
    assert info["builder"] == "ploc"
    assert 0 < info["depth4"] < info["depth"]  # the direct pass traverses the 4-wide tree


def _diagnose(g, o, hg, ho, n, seed):
    """Describe each differing hit word: ray bits, GPU BVH / GPU brute / oracle BVH / oracle brute."""
    lines = []
    rays = g.diffuse_rays(n, seed)
    for (i, slot, r) in list(zip(*np.nonzero(hg != ho)))[:10]:
        face = 1 - slot
        og, dg = rays[i, slot, r, :3], rays[i, slot, r, 3:]
        oo, do = o.diffuse_ray(int(i), int(r % n), int(r // n), n, face, seed)
        ray = np.r_[og, dg][None, :]
        gb = g.trace(ray, [i], 1e-5, brute=False)
        gf = g.trace(ray, [i], 1e-5, brute=True)
        ob = o.trace(og, dg, 1e-5, int(i), brute=False)
        of = o.trace(og, dg, 1e-5, int(i), brute=True)
        lines.append(f"prim {i} slot {slot} ray {r}: gpu {hg[i, slot, r]:#x} oracle {ho[i, slot, r]:#x}; "
                     f"ray bits equal {np.array_equal(og, oo) and np.array_equal(dg, do)} "
                     f"(gpu {og.tolist()} {dg.tolist()} / cpu {oo.tolist()} {do.tolist()}); "
                     f"gpu bvh {gb[0][0]} t={gb[1][0]!r} gpu brute {gf[0][0]} t={gf[1][0]!r} "
                     f"cpu bvh {ob} cpu brute {of}")
    return "\n".join(lines[:10])


@pytest.mark.parametrize("seed", [3, 77])
def test_diffuse_hits_bit_exact_config1(ctx, seed):
    s = scenes.config1(1)
    g, o = _both(ctx, s)
    hg = g.diffuse_hits(10, seed)
    ho = o.diffuse_hits(10, seed)
    assert np.array_equal(hg, ho), f"{np.count_nonzero(hg != ho)} hit words differ\n" + _diagnose(g, o, hg, ho, 10,
                                                                                               seed)


def test_diffuse_rays_bit_exact(ctx):
    s = small_canopy(3, 500)
    g, o = _both(ctx, s)
    n, seed = 6, 4242
    rg = g.diffuse_rays(n, seed)
    bad = 0
    for i in range(0, s.P, 7):
        for slot in range(2):
            if slot == 1 and s.twosided[i] == 0:
                continue
            for r in range(n * n):
                oo, do = o.diffuse_ray(i, r % n, r // n, n, 1 - slot, seed)
                bad += not (np.array_equal(rg[i, slot, r, :3], oo) and np.array_equal(rg[i, slot, r, 3:], do))
    assert bad == 0


def test_diffuse_hits_bit_exact_periodic(ctx):
    s = small_canopy(9, 4000, periodic=True)
    g, o = _both(ctx, s)
    assert g.info()["nbbox"] == 4
    hg = g.diffuse_hits(8, 1234)
    ho = o.diffuse_hits(8, 1234)
    assert np.array_equal(hg, ho), f"{np.count_nonzero(hg != ho)} hit words differ"


def test_direct_sums_config1(ctx):
    s = scenes.config1(1)
    g, o = _both(ctx, s)
    src = [native.sun_direction(30, 0), native.sun_direction(60, 200)]
    sg = g.direct_sums(10, 99, src)
    so = o.direct_sums(10, 99, src)
    assert np.max(np.abs(sg - so)) <= 1e-5 * max(1.0, np.abs(so).max())


def test_direct_grid_equals_bvh(ctx):
    """Occlusion grids and BVH traversal make the same decision for every shadow ray, on a scene with many
    overlapping primitives and several sun directions.  The grid path hands rays to lanes dynamically, so
    its compensated sums group the (equal) per-ray strengths differently: they may differ in the last bit,
    while a single differing decision would move a sum by 1/144 relative."""
    s = small_canopy(23, 4000)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    src = [native.sun_direction(z, a) for z, a in ((5, 0), (35, 120), (70, 250), (85, 10))]
    req = two_band_request(s, depth=2, nrays=36)
    try:
        ctx.set_direct_grids(True)
        sg = g.direct_sums(12, 5, src)
        Rg = native.run_band(g, req)
        assert ctx.timing()["direct_grid"] == 1
        ctx.set_direct_grids(False)
        sb = g.direct_sums(12, 5, src)
        Rb = native.run_band(g, req)
        assert ctx.timing()["direct_grid"] == 0
    finally:
        ctx.set_direct_grids(True)
    assert np.max(np.abs(sg - sb) / np.maximum(np.abs(sb), 1e-30)) <= 1e-6
    assert rel_err(Rg, Rb, floor=1.0).max() <= 1e-6


@pytest.mark.parametrize("schedule", [0, 1])
def test_run_band_config1(ctx, schedule):
    s = scenes.config1(1)
    g, o = _both(ctx, s)
    req = config1_request(s.P, depth=3, schedule=schedule)
    Rg = native.run_band(g, req)
    Ro = o.run_band(req)
    err = rel_err(Rg, Ro, floor=1.0)
    assert err.max() <= 1e-4, f"max rel err {err.max()}"


def test_run_band_two_bands_periodic(ctx):
    s = small_canopy(11, 3000, periodic=True)
    g, o = _both(ctx, s)
    req = two_band_request(s, depth=2, nrays=36)
    Rg = native.run_band(g, req)
    Ro = o.run_band(req)
    err = rel_err(Rg, Ro, floor=1.0)
    assert err.max() <= 1e-4, f"max rel err {err.max()}"


def test_cast_rays_vs_oracle(ctx):
    s = scenes.random_canopy(4, 20000)
    g, o = _both(ctx, s)
    O, D = scenes.rays(8, 200000, [-3, -3, 0], [3, 3, 2.2])
    tg, ng, ug = g.cast_rays(O, D)
    to, no, uo = o.cast_rays(O, D)
    assert np.array_equal(ug, uo), f"{np.count_nonzero(ug != uo)} ids differ"
    hit = uo != 0xFFFFFFFF
    assert hit.sum() > 1000
    assert np.array_equal(tg[hit], to[hit])
    assert np.array_equal(ng[hit], no[hit])


def test_batch_equals_plain(ctx):
    s = scenes.config1(2)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = config1_request(s.P, depth=2, dseeds=(5, 6, 7))
    K = 5
    dirs = np.stack([native.sun_direction(10 + 15 * k, 60 + 40 * k) for k in range(K)])[:, None, :]
    flux = np.array([[[1000.0 - 100 * k]] for k in range(K)], np.float32)
    sky = np.array([[80.0 + 5 * k] for k in range(K)], np.float32)
    seeds = np.arange(K) + 100
    Rb = native.run_band_batch(g, req, dirs, flux, sky, seeds)
    for k in range(K):
        rk = native.BandRequest(s.P, [native.BandSpec("PAR", 2, False, float(sky[k, 0]))], rho=req.rho, tau=req.tau,
                                direct_rays=100, diffuse_rays=100, source_dirs=dirs[k], source_flux=flux[k],
                                direct_seed=int(seeds[k]), diffuse_seeds=[5, 6, 7])
        Rp = native.run_band(g, rk)
        err = rel_err(Rb[k], Rp, floor=1.0)
        assert err.max() <= 1e-5, f"sun {k}: {err.max()}"
