"""The homogeneous-canopy benchmark scene (benchmarks/radiation_homogeneous_canopy/main.cpp) at a small
domain through the C ABI against the oracle (GPU).

Disk-masked 2 x 2 sub-patch leaves with custom quarter UVs, periodic in x and y, a one-sided ground tile;
the benchmark's two runBand calls (band "direct": sun at 36 degrees zenith, unit horizontal flux; band
"diffuse": unit isotropic sky), each at depth 0, with fewer rays.  Bars: runBand within 1e-4 of the FP64
oracle; diffuse hit words bit-exact."""
import math

import numpy as np
import pytest

from helios_amd import native, scenes
from oracle.oracle import OracleScene
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def canopy():
    return scenes.homogeneous_canopy(4, D=2.0, lai=1.5, h=1.0, w_leaf=0.1, mask_res=64)


def _pair(ctx, canopy):
    s, masks, mid, uv, uvid = canopy
    g = ctx.scene(s.transforms, s.types, s.twosided, periodic=s.periodic)
    g.set_masks(masks, mid, uv, uvid)
    o = OracleScene.from_arrays(s)
    o.set_masks(masks, mid, uv, uvid)
    return s, g, o


def _requests(P):
    zero = np.zeros((P, 1), np.float32)
    th = 0.2 * math.pi
    sun = native.sun_direction(math.degrees(th), 0.0)
    direct = native.BandRequest(P, [native.BandSpec("direct", 0, False, 0.0)], rho=zero, tau=zero, direct_rays=64,
                                diffuse_rays=16, source_dirs=[sun], source_flux=[[1.0 / math.cos(th)]],
                                direct_seed=3, diffuse_seeds=[5])
    diffuse = native.BandRequest(P, [native.BandSpec("diffuse", 0, False, 1.0)], rho=zero, tau=zero,
                                 direct_rays=16, diffuse_rays=64, source_dirs=[sun], source_flux=[[0.0]],
                                 direct_seed=7, diffuse_seeds=[9])
    return direct, diffuse


def test_homogeneous_canopy_run_band_vs_oracle(ctx, canopy):
    s, g, o = _pair(ctx, canopy)
    for req in _requests(s.P):
        Rg = native.run_band(g, req)
        Ro = o.run_band(req)
        assert rel_err(Rg, Ro, floor=1e-2).max() <= 1e-4, req.bands[0].label
        # leaves absorb: the canopy intercepts a sizeable share of the unit flux
        nl = int((s.twosided == 1).sum())
        assert Ro[:nl].mean() > 0.05 and Ro[nl:].mean() > 0.05


def test_homogeneous_canopy_diffuse_hits_bit_exact(ctx, canopy):
    _, g, o = _pair(ctx, canopy)
    hg = g.diffuse_hits(8, 21)
    ho = o.diffuse_hits(8, 21)
    assert np.array_equal(hg, ho), f"{np.count_nonzero(hg != ho)} hit words differ"
    # the masks matter: without them rays stop on the transparent corners of the leaves
    s = canopy[0]
    plain = OracleScene.from_arrays(s).diffuse_hits(8, 21)
    assert np.count_nonzero(plain != ho) > 0
