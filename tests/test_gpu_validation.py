"""Geometry the kernels would not trace is rejected at the boundary, by name, instead of silently dropping out
of the scene (GPU: scene creation needs a device context)."""
import numpy as np
import pytest

from helios_amd import native
from tests import closed_forms as cf

pytestmark = pytest.mark.gpu


def _two(types):
    T = np.stack([cf.patch_T((0, 0, 0), (1, 0, 0), (0, 1, 0)), cf.patch_T((0, 0, 1), (1, 0, 0), (0, 1, 0))])
    return T, np.asarray(types, np.uint32), np.zeros(2, np.uint32)


@pytest.mark.parametrize("types,match", [([0, 2], "primitive 1 is a disk"), ([0, 5], "unknown type"),
                                         ([9, 0], "unknown type")])
def test_untraced_primitive_types_are_rejected(ctx, types, match):
    T, ty, tw = _two(types)
    with pytest.raises(native.RTError, match=match):
        ctx.scene(T, ty, tw)
    be = native.Backend(ctx)
    with pytest.raises(native.RTError, match=match):
        be.update_geometry(T, ty, tw)


def test_traced_types_are_accepted(ctx):
    T, ty, tw = _two([0, 3])
    assert ctx.scene(T, ty, tw).info()["nleaves"] == 2
    T, ty, tw = _two([4, 1])  # a voxel is a leaf too (castRays hits it)
    assert ctx.scene(T, ty, tw).info()["nleaves"] == 2
    T, ty, tw = _two([1, 0])
    be = native.Backend(ctx)
    be.update_geometry(T, ty, tw)
    be.build_acceleration_structure()
