"""The C ABI library builds, loads and exports every declared entry point; no CPU fallback (CPU)."""
import os
import re

import pytest

from helios_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "helios_rt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("rt_create", "rt_scene_create", "rt_cast_rays_soa", "rt_run_band", "rt_run_band_batch"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = native.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    missing = [s for s in declared_symbols() if s not in native.SIGNATURES]
    assert not missing, missing


def test_no_silent_cpu_fallback():
    # with no device (this container) the context fails loudly with a reason
    lib = native.load_library()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(native.RTError, match="no HIP device|gfx950"):
        native.Context(0)


def _footprint(P, nsuns, nsources):
    req = native.BandRequest(1, [native.BandSpec("PAR", 1, False, 100.0)], rho=[[0.1]], tau=[[0.05]],
                             direct_rays=4, diffuse_rays=4, source_dirs=[[0.0, 0.0, 1.0]] * nsources,
                             source_flux=[[1000.0]] * nsources)
    return native.solve_footprint(P, req, nsuns)


def test_direct_work_items_are_bounded():
    # suns x sources x primitives of one direct launch is counted in 32 bits: a batch that would wrap it is
    # rejected (host-only check, the same one run_band makes before any launch), one that fits is sized
    limit = 0xFFFFFFFF - (1 << 25)
    P = 1 << 26
    assert 63 * P <= limit < 64 * P
    assert _footprint(P, 63, 1) > 0
    with pytest.raises(native.RTError, match="work items"):
        _footprint(P, 64, 1)
    with pytest.raises(native.RTError, match="work items"):
        _footprint(P, 16, 4)
