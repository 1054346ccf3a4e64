"""Launch shapes (waves per workgroup, workgroups per CU, work-queue chunk) and the autotuner.

Results must not depend on the launch shape: every primitive is reduced by one wave in a fixed order,
so any shape gives bit-identical fluxes."""
import numpy as np
import pytest

from helios_amd import native
from tests.helpers import small_canopy, two_band_request

pytestmark = pytest.mark.gpu



@pytest.fixture(scope="module")
def initial(ctx):
    """The context's launch shapes and grid density as this module found them (the library's defaults, or
    whatever an earlier module left); every test restores them."""
    return [ctx.get_launch_cfg(k) for k in range(3)], ctx.grid_density()


def _restore(ctx, initial):
    shapes, density = initial
    for k in range(3):
        ctx.set_launch_cfg(k, *shapes[k], grid_density=density if k == 0 else 0)


@pytest.mark.parametrize("shape", [(1, 2, 1), (2, 8, 4), (4, 16, 16), (1, 1, 3)])
def test_results_independent_of_launch_shape(ctx, initial, shape):
    s = small_canopy(5, 6000)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = two_band_request(s, depth=2, nrays=36)
    R0 = native.run_band(g, req)
    try:
        for k in range(3):
            ctx.set_launch_cfg(k, *shape)
            assert ctx.get_launch_cfg(k) == shape
        R1 = native.run_band(g, req)
    finally:
        _restore(ctx, initial)
    assert np.array_equal(R0, R1)


def test_autotune_installs_fastest_shapes(ctx, initial):
    s = small_canopy(5, 20000)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    req = two_band_request(s, depth=2, nrays=36)
    R0 = native.run_band(g, req)
    try:
        res = ctx.autotune(g, req, nsuns=8)
        # 4 grid densities + 25 direct + 25 diffuse + 13 gather shapes (non-periodic scene: grids apply)
        assert res["candidates"] == 67 and res["sample_prims"] > 0
        assert res["grid_density"] in (1, 2, 4, 8) and ctx.grid_density() == res["grid_density"]
        for k, name in enumerate(("direct", "diffuse", "gather")):
            r = res[name]
            assert 0 < r["ms"] <= r["default_ms"] + 1e-9
            assert ctx.get_launch_cfg(k) == (r["waves_per_block"], r["blocks_per_cu"], r["chunk"])
        R1 = native.run_band(g, req)
        # a batch base without a sun of its own is tuned on an oblique default direction (grids still apply)
        zero_dir = native.BandRequest(s.P, req.bands, rho=req.rho, tau=req.tau, direct_rays=36, diffuse_rays=36,
                                      source_dirs=[[0.0, 0.0, 0.0]], source_flux=[[0.0, 0.0]])
        assert ctx.autotune(g, zero_dir, nsuns=8)["candidates"] == 67
    finally:
        _restore(ctx, initial)
    assert np.array_equal(R0, R1)


def test_launch_cfg_validation(ctx, initial):
    with pytest.raises(native.RTError, match="waves_per_block"):
        ctx.set_launch_cfg(0, 3, 4, 1)
    with pytest.raises(native.RTError, match="blocks_per_cu"):
        ctx.set_launch_cfg(1, 4, 0, 1)
    with pytest.raises(native.RTError, match="chunk"):
        ctx.set_launch_cfg(2, 4, 4, 1000)
    # rejected shapes leave the installed ones untouched
    assert [ctx.get_launch_cfg(k) for k in range(3)] == initial[0]
