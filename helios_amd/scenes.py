"""Seeded synthetic scenes (ctypes binding of ``libhelios_scenes.so``) for the five workloads."""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhelios_scenes.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        f32p, u32p = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
        u32 = C.c_uint32
        _lib.sg_config1.argtypes = [u32, f32p, u32p, u32p]
        _lib.sg_patch_canopy.argtypes = [u32, u32, C.c_float, C.c_float, C.c_float, u32, u32, f32p, u32p, u32p]
        _lib.sg_config2.argtypes = [u32, u32, f32p, u32p, u32p]
        _lib.sg_orchard.argtypes = [u32, u32, u32, u32, u32, u32, C.c_float, f32p, u32p, u32p]
        _lib.sg_random_canopy.argtypes = [u32, u32, f32p, u32p, u32p]
        _lib.sg_rays.argtypes = [u32, C.c_uint64, f32p, f32p, f32p, f32p]
        _lib.sg_sincos_2pi.argtypes = [C.c_float, f32p, f32p]
        for n in ("sg_config1", "sg_patch_canopy", "sg_config2", "sg_orchard", "sg_random_canopy"):
            getattr(_lib, n).restype = u32
        _lib.sg_rays.restype = None
    return _lib


@dataclass
class SceneArrays:
    transforms: np.ndarray  # [P][16] float32
    types: np.ndarray       # [P] uint32
    twosided: np.ndarray    # [P] uint32
    periodic: tuple = (False, False)

    @property
    def P(self) -> int:
        return int(self.types.shape[0])

    def areas(self) -> np.ndarray:
        T = self.transforms.reshape(-1, 16).astype(np.float64)
        patch = (self.types == 0) | (self.types == 3)
        a = np.where(patch[:, None], T[:, [0, 4, 8]], T[:, [1, 5, 9]])
        b = np.where(patch[:, None], T[:, [1, 5, 9]], T[:, [0, 4, 8]] + T[:, [1, 5, 9]])
        cr = np.linalg.norm(np.cross(a, b), axis=1)
        return np.where(patch, cr, np.where(self.types == 1, 0.5 * cr, 0.0))


def _gen(fn, *args) -> SceneArrays:
    lib = _load()
    P = fn(*args, None, None, None)
    T = np.zeros((P, 16), np.float32)
    ty = np.zeros(P, np.uint32)
    tw = np.zeros(P, np.uint32)
    f32p, u32p = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
    n = fn(*args, T.ctypes.data_as(f32p), ty.ctypes.data_as(u32p), tw.ctypes.data_as(u32p))
    assert n == P
    return SceneArrays(T, ty, tw)


def config1(seed: int = 1) -> SceneArrays:
    """1,600 square leaves (0.1 m) + one 2x2 m ground patch."""
    return _gen(_load().sg_config1, seed)


def patch_canopy(seed: int, nleaves: int, w: float, D: float, h: float, ng: int, periodic=False,
                 inset=False) -> SceneArrays:
    """Square leaves of side w, centres uniform in a D x D x h volume, over a D x D ground of ng x ng
    patches; inset keeps each leaf inside the ground's footprint (exact periodic cell)."""
    s = _gen(_load().sg_patch_canopy, seed, nleaves, w, D, h, ng, 1 if inset else 0)
    s.periodic = (periodic, periodic)
    return s


def config2(seed: int = 2, nleaves: int = 200_000) -> SceneArrays:
    """200k isosceles leaf triangles (4-16 cm^2) in a periodic 10x10x2 m canopy + 400 ground patches."""
    s = _gen(_load().sg_config2, seed, nleaves)
    s.periodic = (True, True)
    return s


def orchard(seed: int = 3, nx: int = 5, ny: int = 5, leaves_per_tree: int = 38_000, tube_tris: int = 2_000,
            tiles: int = 1, jitter: float = 0.0) -> SceneArrays:
    """Orchard of nx*ny trees (tubes + leaf triangles) over a ground of patches; tiles > 1 repeats it in x."""
    return _gen(_load().sg_orchard, seed, nx, ny, leaves_per_tree, tube_tris, tiles, jitter)


def random_canopy(seed: int, P: int) -> SceneArrays:
    """Mixed patch/triangle canopy of about P primitives (serving requests)."""
    return _gen(_load().sg_random_canopy, seed, P)


def rays(seed: int, n: int, lo, hi):
    """n rays with origins uniform in the box [lo, hi] and directions uniform on the sphere."""
    lib = _load()
    O = np.zeros((n, 3), np.float32)
    D = np.zeros((n, 3), np.float32)
    f32p = C.POINTER(C.c_float)
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    lib.sg_rays(seed, n, lo.ctypes.data_as(f32p), hi.ctypes.data_as(f32p), O.ctypes.data_as(f32p),
                D.ctypes.data_as(f32p))
    return O, D


def sincos_2pi(r: float):
    lib = _load()
    s, c = C.c_float(), C.c_float()
    lib.sg_sincos_2pi(r, C.byref(s), C.byref(c))
    return s.value, c.value
