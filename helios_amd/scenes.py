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
        _lib.sg_mixed_canopy.argtypes = [u32, u32, f32p, u32p, u32p]
        _lib.sg_rays.argtypes = [u32, C.c_uint64, f32p, f32p, f32p, f32p]
        _lib.sg_sincos_2pi.argtypes = [C.c_float, f32p, f32p]
        for n in ("sg_config1", "sg_patch_canopy", "sg_config2", "sg_orchard", "sg_random_canopy", "sg_mixed_canopy"):
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


def mixed_canopy(seed: int, P: int) -> SceneArrays:
    """Canopy of about P primitives whose leaves cycle through square patches, triangles and 6-triangle disks
    (Helios disks are triangle fans), over a one-sided ground (config 4's mixed scenes)."""
    return _gen(_load().sg_mixed_canopy, seed, P)


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


def homogeneous_canopy(seed: int = 1, D: float = 20.0, lai: float = 1.0, h: float = 3.0, w_leaf: float = 0.05,
                       mask_res: int = 800):
    """The radiation homogeneous-canopy benchmark scene (benchmarks/radiation_homogeneous_canopy/main.cpp):
    square leaves of side w_leaf made of 2 x 2 sub-patches, each textured with its quarter of a disk mask
    (the harness uses an 800 x 800 disk.png; here an analytic disk of mask_res texels), enough leaves for
    the leaf area index over a D x D periodic domain (leaf area = the mask's solid area), positions
    uniform in x, y and in z over [w_leaf / 2, w_leaf / 2 + h], spherical leaf angles (elevation
    acos(1 - u), azimuth 2 pi u'), rotated about y by -elevation then about z by the azimuth; a one-sided
    ground tile of 100 x 100 patches.  Returns (scene, masks, mask_ids, uv, uv_ids)."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:mask_res, 0:mask_res]
    r = (mask_res - 1) / 2.0
    disk = ((xx - r) ** 2 + (yy - r) ** 2 <= (mask_res / 2.0) ** 2).astype(np.uint8)
    solid = float(disk.mean())
    n_leaves = int(round(lai * D * D / (w_leaf * w_leaf * solid)))
    u = rng.uniform(size=(n_leaves, 5))
    pos = np.stack([(-0.5 + u[:, 0]) * D, (-0.5 + u[:, 1]) * D, 0.5 * w_leaf + u[:, 2] * h], 1)
    elev = np.arccos(1.0 - u[:, 3])
    az = 2.0 * np.pi * u[:, 4]
    # R = Rz(az) Ry(-elev)
    ce, se, ca, sa = np.cos(-elev), np.sin(-elev), np.cos(az), np.sin(az)
    R = np.zeros((n_leaves, 3, 3))
    R[:, 0, 0], R[:, 0, 1], R[:, 0, 2] = ca * ce, -sa, ca * se
    R[:, 1, 0], R[:, 1, 1], R[:, 1, 2] = sa * ce, ca, sa * se
    R[:, 2, 0], R[:, 2, 1], R[:, 2, 2] = -se, 0.0, ce
    q = w_leaf / 2.0
    Ts, uvs = [], []
    for jj in range(2):
        for ii in range(2):
            c_l = np.array([(-0.5 + (ii + 0.5) / 2) * w_leaf, (-0.5 + (jj + 0.5) / 2) * w_leaf, 0.0])
            ex = R @ np.array([q, 0.0, 0.0])
            ey = R @ np.array([0.0, q, 0.0])
            ez = np.cross(ex, ey)
            ez /= np.linalg.norm(ez, axis=1, keepdims=True)
            c = pos + R @ c_l
            T = np.zeros((n_leaves, 16))
            T[:, 0], T[:, 4], T[:, 8] = ex[:, 0], ex[:, 1], ex[:, 2]
            T[:, 1], T[:, 5], T[:, 9] = ey[:, 0], ey[:, 1], ey[:, 2]
            T[:, 2], T[:, 6], T[:, 10] = ez[:, 0], ez[:, 1], ez[:, 2]
            T[:, 3], T[:, 7], T[:, 11] = c[:, 0], c[:, 1], c[:, 2]
            T[:, 15] = 1.0
            Ts.append(T)
            u0, u1, v0, v1 = ii / 2, (ii + 1) / 2, jj / 2, (jj + 1) / 2
            uvs.append(np.tile(np.array([[u0, v0], [u1, v0], [u1, v1], [u0, v1]]), (n_leaves, 1, 1)))
    leaf_T = np.stack(Ts, 1).reshape(-1, 16)   # leaf k's sub-patches are consecutive
    leaf_uv = np.stack(uvs, 1).reshape(-1, 4, 2)
    ng = 100
    g = (np.arange(ng) + 0.5) / ng * D - D / 2
    gx, gy = np.meshgrid(g, g, indexing="xy")
    GT = np.zeros((ng * ng, 16))
    GT[:, 0] = D / ng
    GT[:, 5] = D / ng
    GT[:, 10] = 1.0
    GT[:, 3], GT[:, 7] = gx.reshape(-1), gy.reshape(-1)
    GT[:, 15] = 1.0
    T = np.concatenate([leaf_T, GT]).astype(np.float32)
    P = T.shape[0]
    nl = leaf_T.shape[0]
    types = np.zeros(P, np.uint32)
    twosided = np.r_[np.ones(nl, np.uint32), np.zeros(ng * ng, np.uint32)]
    mask_ids = np.r_[np.zeros(nl, np.int32), np.full(ng * ng, -1, np.int32)]
    uv = np.concatenate([leaf_uv, np.zeros((ng * ng, 4, 2))]).astype(np.float32)
    uv_ids = np.r_[np.arange(nl, dtype=np.int32), np.full(ng * ng, -1, np.int32)]
    return SceneArrays(T, types, twosided, (True, True)), [disk], mask_ids, uv, uv_ids
