"""The allocation-free LiDAR cast (rt_lidar_cast_rays_noalloc) against rt_lidar_cast_rays, bit for bit, through
the C ABI: two calls in a row on the same scene, the second smaller than the first (the shared scratch keeps
the first call's contents), plus the size limit.  The caller's device arrays are made with the HIP runtime the
library links (test plumbing)."""
import math

import numpy as np
import pytest

from helios_amd import native, scenes
from tests.helpers import DeviceArray

pytestmark = pytest.mark.gpu


def _scan(n, origin, seed):
    rng = np.random.default_rng(seed)
    th = np.arccos(rng.uniform(-0.6, 0.9, n))
    ph = rng.uniform(0.0, 2.0 * math.pi, n)
    D = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1).astype(np.float32)
    O = np.tile(np.asarray(origin, np.float32), (n, 1))
    O += rng.normal(0.0, 0.02, O.shape).astype(np.float32)
    return O, D


def _noalloc(g, O, D, miss_t):
    n = O.shape[0]
    dO, dD = DeviceArray(O), DeviceArray(D)
    # outputs pre-filled with garbage: every element must be written by the call
    ht = DeviceArray(np.full(n, -7.0, np.float32))
    hf = DeviceArray(np.full(n, -7.0, np.float32))
    hi = DeviceArray(np.full(n, -7, np.int32))
    g.lidar_cast_rays_noalloc(dO.addr, dD.addr, n, miss_t, ht.addr, hf.addr, hi.addr)
    out = ht.get(), hf.get(), hi.get()
    for a in (dO, dD, ht, hf, hi):
        a.free()
    return out


def test_noalloc_lidar_cast_equals_allocating_path(ctx):
    s = scenes.random_canopy(31, 40000)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    miss_t = 1001.0
    for n, seed in ((300000, 1), (70001, 2)):  # the second call is smaller
        O, D = _scan(n, (0.0, 0.0, 0.8), seed)
        ref_t, ref_f, ref_i = g.lidar_cast_rays(O, D, 1, miss_t)
        t, f, i = _noalloc(g, O, D, miss_t)
        assert np.array_equal(i, ref_i), (n, int((i != ref_i).sum()))
        assert np.array_equal(t, ref_t) and np.array_equal(f, ref_f)
        hit = ref_i >= 0
        assert 0.05 < hit.mean() < 0.95


def test_noalloc_lidar_cast_rejects_oversized_calls(ctx):
    s = scenes.random_canopy(31, 2000)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    with pytest.raises(native.RTError, match="at most"):
        # the size check comes before any pointer is used
        g.lidar_cast_rays_noalloc(16, 16, native.NOALLOC_MAX_RAYS + 1, 10.0, 16, 16, 16)
