"""ctypes binding of ``libhelios_rt.so`` (the C ABI declared in ``include/helios_rt.h``).

There is no fallback: if the library is missing or no gfx950 device is present, every entry point
raises :class:`RTError` naming the reason.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhelios_rt.so")

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class RTError(RuntimeError):
    pass


# radiation source types (RayTracingTypes.h:189-196)
SOURCE_COLLIMATED, SOURCE_SPHERE, SOURCE_SUN_SPHERE, SOURCE_RECTANGLE, SOURCE_DISK = 0, 1, 2, 3, 4
NOALLOC_MAX_RAYS = 1 << 20  # RT_NOALLOC_MAX_RAYS


class rt_band_request(C.Structure):
    _fields_ = [
        ("nbands", C.c_uint32), ("nsources", C.c_uint32),
        ("direct_rays", C.c_uint32), ("diffuse_rays", C.c_uint32),
        ("band_depth", u32p), ("emission", u32p),
        ("source_dir", f32p), ("source_flux", f32p), ("diffuse_flux", f32p),
        ("rho", f32p), ("tau", f32p), ("eps", f32p), ("temperature", f32p),
        ("direct_seed", C.c_uint32), ("diffuse_seed", u32p), ("schedule", C.c_uint32),
        ("source_type", u32p), ("source_width", f32p), ("source_rotation", f32p),
        ("diffuse_extinction", f32p), ("diffuse_peak_dir", f32p), ("diffuse_dist_norm", f32p), ("sky_params", f32p),
        ("is_glass", C.POINTER(C.c_uint8)), ("glass_n", f32p), ("glass_KL", f32p),
    ]


class rt_sun_batch(C.Structure):
    _fields_ = [("nsuns", C.c_uint32), ("source_dir", f32p), ("source_flux", f32p),
                ("diffuse_flux", f32p), ("direct_seed", u32p), ("diffuse_peak_dir", f32p)]


class rt_scene_info(C.Structure):
    _fields_ = [("nprims", C.c_uint32), ("nleaves", C.c_uint32), ("depth", C.c_uint32), ("nbbox", C.c_uint32),
                ("domain", C.c_float * 4), ("device_bytes", C.c_uint64), ("builder", C.c_uint32),
                ("depth4", C.c_uint32)]


class rt_launch_cfg(C.Structure):
    _fields_ = [("waves_per_block", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("chunk", C.c_uint32),
                ("grid_density", C.c_uint32)]


class rt_tune_result(C.Structure):
    _fields_ = [("cfg", rt_launch_cfg * 3), ("ms", C.c_double * 3), ("default_ms", C.c_double * 3),
                ("candidates", C.c_uint32), ("sample_prims", C.c_uint32)]


class rt_timing(C.Structure):
    _fields_ = [("direct_ms", C.c_double), ("diffuse_ms", C.c_double), ("gather_ms", C.c_double),
                ("other_ms", C.c_double), ("direct_rays", C.c_uint64), ("diffuse_rays", C.c_uint64),
                ("direct_launches", C.c_uint32), ("diffuse_launches", C.c_uint32), ("direct_grid", C.c_uint32),
                ("grid_entries", C.c_uint32)]


class rt_launch_params(C.Structure):
    _fields_ = [("launch_offset", C.c_uint32), ("launch_count", C.c_uint32), ("rays_per_primitive", C.c_uint32),
                ("random_seed", C.c_uint32), ("num_bands_global", C.c_uint32), ("num_bands_launch", C.c_uint32),
                ("band_map", u32p), ("band_active", C.POINTER(C.c_uint8)), ("launch_face", C.c_uint32),
                ("diffuse_flux", f32p), ("radiation_out_top", f32p), ("radiation_out_bottom", f32p),
                ("diffuse_extinction", f32p), ("diffuse_peak_dir", f32p), ("diffuse_dist_norm", f32p),
                ("specular_reflection_enabled", C.c_uint32)]


class rt_camera_params(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 2), ("focal_length", C.c_float),
                ("lens_diameter", C.c_float), ("fov_aspect", C.c_float), ("viewplane_length", C.c_float),
                ("resolution", C.c_uint32 * 2), ("resolution_full", C.c_uint32 * 2), ("pixel_offset", C.c_uint32 * 2),
                ("antialiasing_samples", C.c_uint32), ("random_seed", C.c_uint32),
                ("specular_reflection_enabled", C.c_uint32)]


class rt_server_cfg(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("max_batch_suns", C.c_uint32), ("memory_budget", C.c_uint64)]


class rt_solve_sun(C.Structure):
    _fields_ = [("source_dir", f32p), ("source_flux", f32p), ("diffuse_flux", f32p), ("direct_seed", C.c_uint32)]


class rt_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("batches", C.c_uint64), ("batched_requests", C.c_uint64), ("max_batch", C.c_uint32),
                ("queued", C.c_uint32), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("memory_budget", C.c_uint64), ("peak_workspace_bytes", C.c_uint64),
                ("resident_model_bytes", C.c_uint64)]


# name -> (restype, argtypes); the ABI test checks that every declared symbol is exported
SIGNATURES = {
    "rt_last_error": (C.c_char_p, []),
    "rt_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rt_destroy": (C.c_int, [C.c_void_p]),
    "rt_backend_name": (C.c_char_p, [C.c_void_p]),
    "rt_query_memory": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "rt_scene_create": (C.c_int, [C.c_void_p, C.c_uint32, f32p, u32p, u32p, u32p, C.c_int, C.c_int,
                                  C.POINTER(C.c_void_p)]),
    "rt_scene_destroy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "rt_scene_get_info": (C.c_int, [C.c_void_p, C.POINTER(rt_scene_info)]),
    "rt_cast_rays_soa_packets": (C.c_int, [C.c_void_p, C.c_void_p, f32p, f32p, C.c_size_t, C.c_size_t, C.c_float, f32p,
                                           f32p, u32p]),
    "rt_lidar_cast_rays": (C.c_int, [C.c_void_p, C.c_void_p, f32p, f32p, C.c_size_t, C.c_size_t, C.c_float, f32p, f32p,
                                     C.POINTER(C.c_int32)]),
    "rt_scene_set_masks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint8), u32p,
                                     C.POINTER(C.c_int32), f32p, C.POINTER(C.c_int32)]),
    "rt_update_texture_masks": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8), u32p,
                                          C.POINTER(C.c_int32), f32p, C.POINTER(C.c_int32)]),
    "rt_cast_rays_soa": (C.c_int, [C.c_void_p, C.c_void_p, f32p, f32p, C.c_size_t, C.c_float, f32p, f32p, u32p]),
    "rt_lidar_cast_rays_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_float,
                                             C.c_void_p, C.c_void_p, C.c_void_p]),
    "rt_cast_rays_soa_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_float,
                                          C.c_void_p, C.c_void_p, C.c_void_p]),
    "rt_run_band": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(rt_band_request), f32p]),
    "rt_run_band_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(rt_band_request), C.POINTER(rt_sun_batch),
                                    C.c_void_p, C.c_int]),
    "rt_energy_residual_fp64": (C.c_int, [C.c_void_p, f64p, C.c_uint32]),
    "rt_set_energy_accounting": (C.c_int, [C.c_void_p, C.c_int]),
    "rt_set_predicates": (C.c_int, [C.c_void_p, C.c_int]),
    "rt_energy_passes_fp64": (C.c_int, [C.c_void_p, f64p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]),
    "rt_set_launch_cfg": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(rt_launch_cfg)]),
    "rt_get_launch_cfg": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(rt_launch_cfg)]),
    "rt_enable_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "rt_set_direct_grids": (C.c_int, [C.c_void_p, C.c_int]),
    "rt_set_counting": (C.c_int, [C.c_void_p, C.c_int]),
    "rt_get_counters": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "rt_get_timing": (C.c_int, [C.c_void_p, C.POINTER(rt_timing)]),
    "rt_debug_diffuse_hits": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, u32p]),
    "rt_debug_direct_sums": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, f32p]),
    "rt_debug_direct_sums_src": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, u32p, f32p,
                                           f32p, f32p]),
    "rt_debug_diffuse_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, f32p]),
    "rt_debug_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, f32p, u32p, C.c_float, C.c_int, u32p, f32p]),
    "rt_update_geometry": (C.c_int, [C.c_void_p, C.c_uint32, f32p, u32p, u32p, u32p, C.c_int, C.c_int]),
    "rt_build_acceleration_structure": (C.c_int, [C.c_void_p]),
    "rt_update_materials": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, f32p]),
    "rt_update_sources": (C.c_int, [C.c_void_p, C.c_uint32, u32p, f32p, f32p, f32p]),
    "rt_zero_radiation_buffers": (C.c_int, [C.c_void_p, C.c_uint32]),
    "rt_zero_scatter_buffers": (C.c_int, [C.c_void_p]),
    "rt_copy_scatter_to_radiation": (C.c_int, [C.c_void_p]),
    "rt_upload_radiation_out": (C.c_int, [C.c_void_p, f32p, f32p, C.c_size_t]),
    "rt_upload_source_fluxes": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "rt_update_specular_materials": (C.c_int, [C.c_void_p, C.c_uint32, f32p, f32p]),
    "rt_update_glass_materials": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8), f32p, f32p]),
    "rt_upload_source_fluxes_cam": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "rt_get_radiation_specular": (C.c_int, [C.c_void_p, f32p, C.c_size_t]),
    "rt_launch_direct_rays": (C.c_int, [C.c_void_p, C.POINTER(rt_launch_params)]),
    "rt_launch_diffuse_rays": (C.c_int, [C.c_void_p, C.POINTER(rt_launch_params)]),
    "rt_get_radiation_results": (C.c_int, [C.c_void_p, f32p, f32p, f32p, f32p, f32p]),
    "rt_get_backend_info": (C.c_int, [C.c_void_p, u32p]),
    "rt_update_sky_model": (C.c_int, [C.c_void_p, C.c_uint32, f32p, f32p, f32p, f32p, C.c_float, f32p, u32p]),
    "rt_launch_camera_rays": (C.c_int, [C.c_void_p, C.POINTER(rt_camera_params)]),
    "rt_zero_camera_pixel_buffers": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "rt_launch_pixel_label_rays": (C.c_int, [C.c_void_p, C.POINTER(rt_camera_params)]),
    "rt_get_camera_results": (C.c_int, [C.c_void_p, f32p, u32p, f32p, C.c_uint32, C.c_uint32]),
    "rt_server_create": (C.c_int, [C.c_void_p, C.POINTER(rt_server_cfg), C.POINTER(C.c_void_p)]),
    "rt_server_destroy": (C.c_int, [C.c_void_p]),
    "rt_server_add_model": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(rt_band_request), u32p]),
    "rt_server_submit": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(rt_solve_sun), f32p, C.POINTER(C.c_uint64)]),
    "rt_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double)]),
    "rt_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(rt_server_stats)]),
    "rt_server_reset_stats": (C.c_int, [C.c_void_p]),
    "rt_solve_footprint": (C.c_int, [C.c_uint32, C.POINTER(rt_band_request), C.c_uint32, C.POINTER(C.c_uint64)]),
    "rt_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(rt_band_request), C.c_uint32,
                              C.POINTER(rt_tune_result)]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libhelios_rt.so and bind every declared signature (raises RTError if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RTError(f"{path} is not built: run __graft_entry__.build() (no CPU fallback exists)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise RTError(_lib.rt_last_error().decode())


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _pf(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(f32p)


def _pu(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(u32p)


@dataclass
class BandSpec:
    """Per-band settings of one runBand call (names follow RadiationModel's setters)."""
    label: str
    scattering_depth: int = 0
    emission: bool = False
    diffuse_flux: float = 0.0


class Context:
    """One gfx950 device context (RayTracingBackend equivalent)."""

    def __init__(self, device: int = 0):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.rt_create(device, C.byref(h)))
        self._h = h
        self.lib = lib

    @property
    def name(self) -> str:
        return self.lib.rt_backend_name(self._h).decode()

    def close(self):
        if self._h:
            self.lib.rt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def memory(self):
        f, t = C.c_size_t(), C.c_size_t()
        _check(self.lib.rt_query_memory(self._h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def scene(self, transforms, types, twosided, uuids=None, periodic=(False, False)) -> "Scene":
        return Scene(self, transforms, types, twosided, uuids, periodic)

    def set_launch_cfg(self, kernel: int, waves_per_block: int, blocks_per_cu: int, chunk: int = 1,
                       grid_density: int = 0):
        cfg = rt_launch_cfg(waves_per_block, blocks_per_cu, chunk, grid_density)
        _check(self.lib.rt_set_launch_cfg(self._h, kernel, C.byref(cfg)))

    def get_launch_cfg(self, kernel: int):
        cfg = rt_launch_cfg()
        _check(self.lib.rt_get_launch_cfg(self._h, kernel, C.byref(cfg)))
        return cfg.waves_per_block, cfg.blocks_per_cu, cfg.chunk

    def grid_density(self) -> int:
        cfg = rt_launch_cfg()
        _check(self.lib.rt_get_launch_cfg(self._h, 0, C.byref(cfg)))
        return cfg.grid_density

    def autotune(self, scene: "Scene", req: "BandRequest", nsuns: int = 1) -> dict:
        """Time candidate launch shapes on a sample of `scene` and install the fastest per kernel class."""
        r = rt_tune_result()
        cs = req.c_struct()
        _check(self.lib.rt_autotune(self._h, scene._h, C.byref(cs), nsuns, C.byref(r)))
        names = ("direct", "diffuse", "gather")
        return {"candidates": r.candidates, "sample_prims": r.sample_prims,
                **{names[k]: {"waves_per_block": r.cfg[k].waves_per_block, "blocks_per_cu": r.cfg[k].blocks_per_cu,
                              "chunk": r.cfg[k].chunk, "ms": round(r.ms[k], 3),
                              "default_ms": round(r.default_ms[k], 3)} for k in range(3)},
                "grid_density": r.cfg[0].grid_density}

    def set_predicates(self, mode: str):
        """Intersection predicates of later traces: "reference" (parity mode, default) or "watertight"."""
        _check(self.lib.rt_set_predicates(self._h, {"reference": 0, "watertight": 1}[mode]))

    def enable_timing(self, on: bool = True):
        _check(self.lib.rt_enable_timing(self._h, 1 if on else 0))

    def set_direct_grids(self, on: bool):
        """Occlusion grids for collimated direct rays (default on) or BVH traversal only; same results."""
        _check(self.lib.rt_set_direct_grids(self._h, 1 if on else 0))

    def timing(self) -> dict:
        t = rt_timing()
        _check(self.lib.rt_get_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in rt_timing._fields_}

    def set_counting(self, on: bool = True):
        """Zero and enable/disable the traversal work counters."""
        _check(self.lib.rt_set_counting(self._h, 1 if on else 0))

    def counters(self) -> dict:
        """BVH node fetches, primitive tests and rays of the direct, diffuse, castRays and camera kernels."""
        out = (C.c_uint64 * 12)()
        _check(self.lib.rt_get_counters(self._h, out))
        v = list(out)
        return {k: {"nodes": v[3 * i], "prims": v[3 * i + 1], "rays": v[3 * i + 2]}
                for i, k in enumerate(("direct", "diffuse", "cast", "camera"))}


class Scene:
    """A resident scene (geometry + LBVH in HBM)."""

    def __init__(self, ctx: Context, transforms, types, twosided, uuids=None, periodic=(False, False)):
        self.ctx = ctx
        T = _f32(transforms).reshape(-1, 16)
        self.P = T.shape[0]
        ty = _u32(types)
        tw = _u32(twosided)
        uu = None if uuids is None else _u32(uuids)
        h = C.c_void_p()
        _check(ctx.lib.rt_scene_create(ctx._h, self.P, _pf(T), _pu(ty), _pu(uu), _pu(tw), int(periodic[0]),
                                       int(periodic[1]), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            self.ctx.lib.rt_scene_destroy(self.ctx._h, self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = rt_scene_info()
        _check(self.ctx.lib.rt_scene_get_info(self._h, C.byref(i)))
        return {"nprims": i.nprims, "nleaves": i.nleaves, "depth": i.depth, "nbbox": i.nbbox,
                "domain": list(i.domain), "device_bytes": i.device_bytes,
                "builder": "ploc" if i.builder == 1 else "lbvh", "depth4": i.depth4}

    def set_masks(self, masks, mask_ids, uv=None, uv_ids=None):
        """Texture masks: masks = list of 2-D arrays [height][width] (nonzero = opaque, row 0 = top),
        mask_ids [P] (-1 = none), uv [P][4][2] vertex UVs, uv_ids [P] (< 0: parametric patch UVs)."""
        args = _mask_args(masks, mask_ids, uv, uv_ids, self.P)
        _check(self.ctx.lib.rt_scene_set_masks(self.ctx._h, self._h, *args[:-1]))

    # ---- CollisionDetection::castRaysSoA ----
    def cast_rays(self, origins, dirs, max_dist: float = -1.0):
        O = _f32(origins).reshape(-1, 3)
        D = _f32(dirs).reshape(-1, 3)
        n = O.shape[0]
        t = np.full(n, np.nan, np.float32)
        nrm = np.full((n, 3), np.nan, np.float32)
        uu = np.zeros(n, np.uint32)
        _check(self.ctx.lib.rt_cast_rays_soa(self.ctx._h, self._h, _pf(O), _pf(D), n, max_dist, _pf(t), _pf(nrm),
                                             _pu(uu)))
        return t, nrm, uu

    def cast_rays_packets(self, origins, dirs, packet_size: int, max_dist: float = -1.0):
        """castRaysSoA_packets: pulses of packet_size contiguous rays (results equal cast_rays)."""
        O = _f32(origins).reshape(-1, 3)
        D = _f32(dirs).reshape(-1, 3)
        n = O.shape[0]
        t = np.full(n, np.nan, np.float32)
        nrm = np.full((n, 3), np.nan, np.float32)
        uu = np.zeros(n, np.uint32)
        _check(self.ctx.lib.rt_cast_rays_soa_packets(self.ctx._h, self._h, _pf(O), _pf(D), n, int(packet_size), max_dist,
                                                     _pf(t), _pf(nrm), _pu(uu)))
        return t, nrm, uu

    def lidar_cast_rays(self, origins, dirs, packet_size: int, miss_t: float):
        """LiDARcloud::castRaysUnified: (hit_t, hit_fnorm, hit_id) per ray."""
        O = _f32(origins).reshape(-1, 3)
        D = _f32(dirs).reshape(-1, 3)
        n = O.shape[0]
        t = np.zeros(n, np.float32)
        f = np.zeros(n, np.float32)
        ids = np.zeros(n, np.int32)
        _check(self.ctx.lib.rt_lidar_cast_rays(self.ctx._h, self._h, _pf(O), _pf(D), n, int(packet_size), float(miss_t),
                                               _pf(t), _pf(f), ids.ctypes.data_as(C.POINTER(C.c_int32))))
        return t, f, ids

    def lidar_cast_rays_noalloc(self, d_origins: int, d_dirs: int, n: int, miss_t: float, d_hit_t: int,
                                d_hit_fnorm: int, d_hit_id: int):
        """rt_lidar_cast_rays with device arrays and no allocation (at most NOALLOC_MAX_RAYS rays per call)."""
        _check(self.ctx.lib.rt_lidar_cast_rays_noalloc(self.ctx._h, self._h, C.c_void_p(d_origins), C.c_void_p(d_dirs),
                                                       n, float(miss_t), C.c_void_p(d_hit_t), C.c_void_p(d_hit_fnorm),
                                                       C.c_void_p(d_hit_id)))

    def cast_rays_device(self, d_origins: int, d_dirs: int, n: int, max_dist: float, d_t: int, d_n: int, d_u: int):
        _check(self.ctx.lib.rt_cast_rays_soa_device(self.ctx._h, self._h, C.c_void_p(d_origins), C.c_void_p(d_dirs),
                                                    n, max_dist, C.c_void_p(d_t), C.c_void_p(d_n), C.c_void_p(d_u)))

    # ---- parity hooks ----
    def diffuse_hits(self, n: int, seed: int) -> np.ndarray:
        out = np.zeros((self.P, 2, n * n), np.uint32)
        _check(self.ctx.lib.rt_debug_diffuse_hits(self.ctx._h, self._h, n, seed, _pu(out)))
        return out

    def diffuse_rays(self, n: int, seed: int) -> np.ndarray:
        out = np.zeros((self.P, 2, n * n, 6), np.float32)
        _check(self.ctx.lib.rt_debug_diffuse_rays(self.ctx._h, self._h, n, seed, _pf(out)))
        return out

    def trace(self, rays, skip, tmin: float, brute: bool = False):
        R = _f32(rays).reshape(-1, 6)
        S = _u32(skip).reshape(-1)
        n = R.shape[0]
        h = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        _check(self.ctx.lib.rt_debug_trace(self.ctx._h, self._h, n, _pf(R), _pu(S), tmin, int(brute), _pu(h), _pf(t)))
        return h, t

    def direct_sums_src(self, d: int, seed: int, positions, types=None, widths=None, rotations=None) -> np.ndarray:
        """One direct pass for sources of any type -> [P][S][2] strength sums (debug/parity access)."""
        X = _f32(positions).reshape(-1, 3)
        S = X.shape[0]
        ty = None if types is None else _u32(types).reshape(S)
        W = None if widths is None else _f32(widths).reshape(S, 2)
        R = None if rotations is None else _f32(rotations).reshape(S, 3)
        out = np.zeros((self.P, S, 2), np.float32)
        _check(self.ctx.lib.rt_debug_direct_sums_src(self.ctx._h, self._h, d, seed, S, _pf(X), _pu(ty), _pf(W), _pf(R),
                                                      _pf(out)))
        return out

    def direct_sums(self, d: int, seed: int, src_dirs) -> np.ndarray:
        S = _f32(src_dirs).reshape(-1, 3)
        out = np.zeros((self.P, S.shape[0], 2), np.float32)
        _check(self.ctx.lib.rt_debug_direct_sums(self.ctx._h, self._h, d, seed, S.shape[0], _pf(S), _pf(out)))
        return out


def _mask_args(masks, mask_ids, uv, uv_ids, P):
    tex = np.concatenate([np.ascontiguousarray(m, np.uint8).reshape(-1) != 0 for m in masks]).astype(np.uint8) \
        if len(masks) else np.zeros(1, np.uint8)
    sizes = _u32([[np.shape(m)[1], np.shape(m)[0]] for m in masks] if len(masks) else np.zeros((1, 2)))
    mid = np.ascontiguousarray(mask_ids, np.int32).reshape(P)
    U = None if uv is None else _f32(uv).reshape(P, 8)
    uid = None if uv_ids is None else np.ascontiguousarray(uv_ids, np.int32).reshape(P)
    p8 = tex.ctypes.data_as(C.POINTER(C.c_uint8))
    pi = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))
    keep = (tex, sizes, mid, U, uid)
    return (len(masks), p8, _pu(sizes), pi(mid), _pf(U), pi(uid), keep)


class Backend:
    """The call-by-call backend interface of one context (helios::RayTracingBackend,
    plugins/radiation/include/RayTracingBackend.h:43-312); method names follow the reference's."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.lib = ctx.lib
        self.P = 0
        self.B_launch = 0

    def update_geometry(self, transforms, types, twosided, uuids=None, periodic=(False, False)):
        T = _f32(transforms).reshape(-1, 16)
        self._geom_P = T.shape[0]
        _check(self.lib.rt_update_geometry(self.ctx._h, T.shape[0], _pf(T), _pu(_u32(types)),
                                           _pu(None if uuids is None else _u32(uuids)), _pu(_u32(twosided)),
                                           int(periodic[0]), int(periodic[1])))

    def update_texture_masks(self, masks, mask_ids, uv=None, uv_ids=None):
        """Texture masks of the staged geometry (as Scene.set_masks), applied by the next build."""
        args = _mask_args(masks, mask_ids, uv, uv_ids, self._geom_P)
        _check(self.lib.rt_update_texture_masks(self.ctx._h, self._geom_P, *args[:-1]))

    def build_acceleration_structure(self):
        _check(self.lib.rt_build_acceleration_structure(self.ctx._h))
        self.P = self._geom_P

    def update_materials(self, rho, tau):
        """rho, tau: [S][P][B_global] (or [P][B_global] for one source)."""
        R = _f32(rho)
        Tt = _f32(tau)
        if R.ndim == 2:
            R, Tt = R[None], Tt[None]
        S, P, B = R.shape
        _check(self.lib.rt_update_materials(self.ctx._h, S, P, B, _pf(R), _pf(Tt)))

    def update_sources(self, positions, types=None, widths=None, rotations=None):
        """positions [S][3]: direction towards a collimated source, else the source position; widths [S][2]
        and rotations [S][3] as RayTracingSource::width / ::rotation (needed for non-collimated types)."""
        D = _f32(positions).reshape(-1, 3)
        S = D.shape[0]
        ty = _u32(np.zeros(S) if types is None else types)
        W = None if widths is None else _f32(widths).reshape(S, 2)
        R = None if rotations is None else _f32(rotations).reshape(S, 3)
        _check(self.lib.rt_update_sources(self.ctx._h, S, _pu(ty), _pf(D), _pf(W), _pf(R)))

    def zero_radiation_buffers(self, launch_band_count: int):
        _check(self.lib.rt_zero_radiation_buffers(self.ctx._h, launch_band_count))
        self.B_launch = launch_band_count

    def zero_scatter_buffers(self):
        _check(self.lib.rt_zero_scatter_buffers(self.ctx._h))

    def copy_scatter_to_radiation(self):
        _check(self.lib.rt_copy_scatter_to_radiation(self.ctx._h))

    def upload_radiation_out(self, top, bottom):
        t, b = _f32(top).reshape(-1), _f32(bottom).reshape(-1)
        if t.size != b.size:
            raise RTError("upload_radiation_out: top and bottom differ in size")
        _check(self.lib.rt_upload_radiation_out(self.ctx._h, _pf(t), _pf(b), t.size))

    def upload_source_fluxes(self, fluxes):
        f = _f32(fluxes).reshape(-1)
        _check(self.lib.rt_upload_source_fluxes(self.ctx._h, _pf(f), f.size))

    def update_glass_materials(self, is_glass, glass_n=None, glass_KL=None):
        """Translucent covers: is_glass, glass_n (refractive index), glass_KL (absorption K*L, None = 0), each
        [P][B_global]; is_glass None removes every cover."""
        if is_glass is None:
            _check(self.lib.rt_update_glass_materials(self.ctx._h, self.P, self.info()["nbands_global"], None, None,
                                                      None))
            return
        g = np.ascontiguousarray(is_glass, np.uint8)
        P, B = g.shape
        gn = _f32(glass_n).reshape(P, B)
        gkl = None if glass_KL is None else _f32(glass_KL).reshape(P, B)
        _check(self.lib.rt_update_glass_materials(self.ctx._h, P, B, g.ctypes.data_as(C.POINTER(C.c_uint8)), _pf(gn),
                                                  _pf(gkl)))

    def update_specular_materials(self, exponent, scale=None):
        """specular_exponent [P] (not positive: no specular term) and specular_scale [P] (None = 0)."""
        e = _f32(exponent).reshape(-1)
        sc = None if scale is None else _f32(scale).reshape(-1)
        _check(self.lib.rt_update_specular_materials(self.ctx._h, e.size, _pf(e), _pf(sc)))

    def upload_source_fluxes_cam(self, weights):
        """The camera's spectral weight of each source and launch band, [S][B_launch]."""
        w = _f32(weights).reshape(-1)
        _check(self.lib.rt_upload_source_fluxes_cam(self.ctx._h, _pf(w), w.size))

    def get_radiation_specular(self) -> np.ndarray:
        """radiation_specular [S][P][B_launch] accumulated by direct launches with specular on."""
        S = self.info()["nsources"]
        out = np.zeros((S, self.P, self.B_launch), np.float32)
        _check(self.lib.rt_get_radiation_specular(self.ctx._h, _pf(out), out.size))
        return out

    def _params(self, rays, seed, nbands_global, band_map=None, band_active=None, launch_face=1,
                diffuse_flux=None, radiation_out=None, offset=0, count=None, sky=None, specular=0):
        keep = []
        bm = None if band_map is None else _u32(band_map)
        ba = None if band_active is None else np.ascontiguousarray(band_active, dtype=np.uint8)
        df = None if diffuse_flux is None else _f32(diffuse_flux)
        ot = ob = None
        if radiation_out is not None:
            ot, ob = _f32(radiation_out[0]).reshape(-1), _f32(radiation_out[1]).reshape(-1)
        sk = [None, None, None] if sky is None else [None if a is None else _f32(a).reshape(-1) for a in sky]
        keep += [bm, ba, df, ot, ob] + sk
        p = rt_launch_params(int(offset), int(self.P - offset if count is None else count), int(rays), int(seed),
                             int(nbands_global), int(self.B_launch), _pu(bm),
                             None if ba is None else ba.ctypes.data_as(C.POINTER(C.c_uint8)), int(launch_face),
                             _pf(df), _pf(ot), _pf(ob), _pf(sk[0]), _pf(sk[1]), _pf(sk[2]), int(specular))
        return p, keep

    def launch_direct_rays(self, rays_per_primitive: int, seed: int, nbands_global: int, band_map=None,
                           offset: int = 0, count: Optional[int] = None, specular: int = 0):
        """specular: non-zero also adds each source's delivered radiation to radiation_specular."""
        p, keep = self._params(rays_per_primitive, seed, nbands_global, band_map, offset=offset, count=count,
                               specular=specular)
        _check(self.lib.rt_launch_direct_rays(self.ctx._h, C.byref(p)))

    def launch_diffuse_rays(self, rays_per_primitive: int, seed: int, nbands_global: int, launch_face: int,
                            diffuse_flux=None, band_map=None, band_active=None, radiation_out=None,
                            offset: int = 0, count: Optional[int] = None, sky_extinction=None, sky_peak_dir=None,
                            sky_norm=None):
        """sky_*: per launch band power-law K, peak direction [B][3] and normalisation (the Prague
        parameters come from update_sky_model)."""
        sky = None if sky_extinction is None and sky_peak_dir is None else (sky_extinction, sky_peak_dir, sky_norm)
        p, keep = self._params(rays_per_primitive, seed, nbands_global, band_map, band_active, launch_face,
                               diffuse_flux, radiation_out, offset, count, sky)
        _check(self.lib.rt_launch_diffuse_rays(self.ctx._h, C.byref(p)))

    def get_radiation_results(self) -> dict:
        shape = (self.P, self.B_launch)
        out = {k: np.zeros(shape, np.float32) for k in ("radiation_in", "radiation_out_top", "radiation_out_bottom",
                                                         "scatter_buff_top", "scatter_buff_bottom")}
        _check(self.lib.rt_get_radiation_results(self.ctx._h, _pf(out["radiation_in"]), _pf(out["radiation_out_top"]),
                                                 _pf(out["radiation_out_bottom"]), _pf(out["scatter_buff_top"]),
                                                 _pf(out["scatter_buff_bottom"])))
        return out

    def info(self) -> dict:
        o = np.zeros(4, np.uint32)
        _check(self.lib.rt_get_backend_info(self.ctx._h, _pu(o)))
        return {"nprims": int(o[0]), "nbands_global": int(o[1]), "nbands_launch": int(o[2]), "nsources": int(o[3])}

    # ---- cameras (RayTracingBackend.h:129-210) ----
    def update_sky_model(self, sky_params, camera_sky_radiance, sun_direction, solar_disk_radiance,
                         solar_disk_cos_angle: float, camera_diffuse_flux=None, band_emission_flag=None):
        """Per launch band: sky_params [B][4] (Prague circumsolar strength, width in degrees, horizon
        brightness, normalisation; normalisation <= 0 = isotropic), camera_sky_radiance [B],
        solar_disk_radiance [B], camera_diffuse_flux [B], band_emission_flag [B]."""
        prm = _f32(sky_params).reshape(-1, 4)
        B = prm.shape[0]
        srad = _f32(camera_sky_radiance).reshape(B)
        sun = _f32(sun_direction).reshape(3)
        solar = _f32(solar_disk_radiance).reshape(B)
        df = None if camera_diffuse_flux is None else _f32(camera_diffuse_flux).reshape(B)
        em = None if band_emission_flag is None else _u32(band_emission_flag).reshape(B)
        _check(self.lib.rt_update_sky_model(self.ctx._h, B, _pf(prm), _pf(srad), _pf(sun), _pf(solar),
                                            float(solar_disk_cos_angle), _pf(df), _pu(em)))

    @staticmethod
    def camera_params(camera: "Camera", tile=None, seed: int = 0, specular: int = 0) -> rt_camera_params:
        """Launch parameters of one tile (x0, y0, w, h) of a camera (the whole image when tile is None);
        specular: 0 off, 1 Blinn-Phong term with scale 1, 2 with the primitives' specular_scale."""
        fx, fy = camera.resolution
        x0, y0, w, h = (0, 0, fx, fy) if tile is None else tile
        return rt_camera_params((C.c_float * 3)(*camera.position), (C.c_float * 2)(*camera.direction),
                                camera.focal_length, camera.lens_diameter, camera.fov_aspect, camera.viewplane_length,
                                (C.c_uint32 * 2)(w, h), (C.c_uint32 * 2)(fx, fy), (C.c_uint32 * 2)(x0, y0),
                                camera.antialiasing_samples, int(seed), int(specular))

    def launch_camera_rays(self, camera: "Camera", seed: int = 0, tile=None, specular: int = 0):
        p = self.camera_params(camera, tile, seed, specular)
        _check(self.lib.rt_launch_camera_rays(self.ctx._h, C.byref(p)))

    def zero_camera_pixel_buffers(self, resolution):
        _check(self.lib.rt_zero_camera_pixel_buffers(self.ctx._h, int(resolution[0]), int(resolution[1])))

    def launch_pixel_label_rays(self, camera: "Camera", tile=None):
        p = self.camera_params(camera, tile)
        _check(self.lib.rt_launch_pixel_label_rays(self.ctx._h, C.byref(p)))

    def get_camera_results(self, resolution) -> dict:
        """pixel_data [H*W][B_launch] (pixel = row * W + column), pixel_labels [H*W], pixel_depths [H*W]."""
        W, H = int(resolution[0]), int(resolution[1])
        out = {"pixel_data": np.zeros((W * H, self.B_launch), np.float32),
               "pixel_labels": np.zeros(W * H, np.uint32), "pixel_depths": np.zeros(W * H, np.float32)}
        _check(self.lib.rt_get_camera_results(self.ctx._h, _pf(out["pixel_data"]), _pu(out["pixel_labels"]),
                                              _pf(out["pixel_depths"]), W, H))
        return out


@dataclass
class Camera:
    """A radiation camera's launch fields (RayTracingTypes.h:221-235): position, direction (elevation,
    azimuth in radians), focal length (working distance), lens diameter (0 = pinhole), HFOV/VFOV aspect,
    viewplane length, resolution (width, height) and antialiasing samples per pixel."""
    position: tuple
    direction: tuple
    resolution: tuple
    focal_length: float = 1.0
    lens_diameter: float = 0.0
    fov_aspect: float = 1.0
    viewplane_length: float = 1.0
    antialiasing_samples: int = 1


class Server:
    """Request server over one context (continuous batching with memory-footprint admission)."""

    def __init__(self, ctx: Context, slots: int = 4, max_batch_suns: int = 64, memory_budget: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = rt_server_cfg(slots, max_batch_suns, memory_budget)
        h = C.c_void_p()
        _check(self.lib.rt_server_create(ctx._h, C.byref(cfg), C.byref(h)))
        self._h = h
        self._models = {}
        self._pending = {}

    def close(self):
        if self._h:
            self.lib.rt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_model(self, scene: "Scene", req: "BandRequest") -> int:
        mid = C.c_uint32()
        cs = req.c_struct()
        _check(self.lib.rt_server_add_model(self._h, scene._h, C.byref(cs), C.byref(mid)))
        self._models[mid.value] = (scene.P, req.B, req.sdir.shape[0])
        return mid.value

    def submit(self, model: int, source_dirs, source_flux, diffuse_flux, direct_seed: int) -> int:
        # an unknown model id still goes to the library, which rejects it
        P, B, S = self._models.get(model, (1, len(np.ravel(diffuse_flux)), len(np.ravel(source_dirs)) // 3))
        D = _f32(source_dirs).reshape(S, 3)
        F = _f32(source_flux).reshape(S, B)
        Q = _f32(diffuse_flux).reshape(B)
        out = np.zeros((P, B), np.float32)
        sun = rt_solve_sun(_pf(D), _pf(F), _pf(Q), int(direct_seed))
        t = C.c_uint64()
        _check(self.lib.rt_server_submit(self._h, model, C.byref(sun), _pf(out), C.byref(t)))
        self._pending[t.value] = out
        return t.value

    def wait(self, ticket: int):
        """-> (radiation_flux [P][B], latency ms)"""
        lat = C.c_double()
        out = self._pending.pop(ticket)
        _check(self.lib.rt_server_wait(self._h, ticket, C.byref(lat)))
        return out, lat.value

    def stats(self) -> dict:
        st = rt_server_stats()
        _check(self.lib.rt_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in rt_server_stats._fields_}

    def reset_stats(self):
        _check(self.lib.rt_server_reset_stats(self._h))


class BandRequest:
    """All inputs of one runBand call except the per-sun part (owns its numpy buffers)."""

    def __init__(self, P: int, bands: Sequence[BandSpec], rho, tau, direct_rays: int, diffuse_rays: int,
                 source_dirs=None, source_flux=None, direct_seed: int = 1, diffuse_seeds=None, eps=None,
                 temperature=None, schedule: int = 0, source_types=None, source_widths=None,
                 source_rotations=None, sky_extinction=None, sky_peak_dir=None, sky_norm=None, sky_prague=None,
                 is_glass=None, glass_n=None, glass_KL=None):
        """source_dirs [S][3]: direction towards a collimated source, else the source position;
        source_types [S] (SOURCE_*; None = all collimated), source_widths [S][2] and source_rotations [S][3]
        as RayTracingSource::width / ::rotation.  Anisotropic sky per band: sky_extinction [B] (power-law K),
        sky_peak_dir [B][3], sky_norm [B] (power-law normalisation), sky_prague [B][4].  Translucent covers:
        is_glass [P][B] (a primitive glass in every band is a cover), glass_n [P][B] refractive index,
        glass_KL [P][B] absorption product (None = 0)."""
        B = len(bands)
        self.P, self.B = P, B
        self.bands = list(bands)
        self.depth = _u32([b.scattering_depth for b in bands])
        self.emission = _u32([1 if b.emission else 0 for b in bands])
        self.dflux = _f32([b.diffuse_flux for b in bands])
        maxd = int(self.depth.max()) if B else 0
        if diffuse_seeds is None:
            diffuse_seeds = [1000 + k for k in range(maxd + 1)]
        self.dseeds = _u32(list(diffuse_seeds) + [0] * max(0, maxd + 1 - len(diffuse_seeds)))
        self.rho = _f32(rho).reshape(P, B)
        self.tau = _f32(tau).reshape(P, B)
        self.eps = None if eps is None else _f32(eps).reshape(P, B)
        self.temp = None if temperature is None else _f32(temperature).reshape(P)
        self.sdir = _f32(source_dirs if source_dirs is not None else np.zeros((0, 3))).reshape(-1, 3)
        S = self.sdir.shape[0]
        self.sflux = _f32(source_flux if source_flux is not None else np.zeros((S, B))).reshape(S, B)
        self.direct_seed = int(direct_seed)
        self.direct_rays, self.diffuse_rays = int(direct_rays), int(diffuse_rays)
        self.schedule = int(schedule)
        self.stype = None if source_types is None else _u32(source_types).reshape(S)
        self.swidth = None if source_widths is None else _f32(source_widths).reshape(S, 2)
        self.srot = None if source_rotations is None else _f32(source_rotations).reshape(S, 3)
        self.sky_ext = None if sky_extinction is None else _f32(sky_extinction).reshape(B)
        self.sky_peak = None if sky_peak_dir is None else _f32(sky_peak_dir).reshape(B, 3)
        self.sky_norm = None if sky_norm is None else _f32(sky_norm).reshape(B)
        self.sky_prague = None if sky_prague is None else _f32(sky_prague).reshape(B, 4)
        self.is_glass = None if is_glass is None else np.ascontiguousarray(is_glass, np.uint8).reshape(P, B)
        self.glass_n = None if glass_n is None else _f32(glass_n).reshape(P, B)
        self.glass_KL = None if is_glass is None else (np.zeros((P, B), np.float32) if glass_KL is None
                                                         else _f32(glass_KL).reshape(P, B))

    def c_struct(self) -> rt_band_request:
        return rt_band_request(
            self.B, self.sdir.shape[0], self.direct_rays, self.diffuse_rays, _pu(self.depth), _pu(self.emission),
            _pf(self.sdir), _pf(self.sflux), _pf(self.dflux), _pf(self.rho), _pf(self.tau), _pf(self.eps),
            _pf(self.temp), self.direct_seed, _pu(self.dseeds), self.schedule, _pu(self.stype), _pf(self.swidth),
            _pf(self.srot), _pf(self.sky_ext), _pf(self.sky_peak), _pf(self.sky_norm), _pf(self.sky_prague),
            None if self.is_glass is None else self.is_glass.ctypes.data_as(C.POINTER(C.c_uint8)), _pf(self.glass_n),
            _pf(self.glass_KL))


def run_band(scene: Scene, req: BandRequest) -> np.ndarray:
    """Whole runBand on the device -> radiation_flux [P][B]."""
    R = np.zeros((scene.P, req.B), np.float32)
    cs = req.c_struct()
    _check(scene.ctx.lib.rt_run_band(scene.ctx._h, scene._h, C.byref(cs), _pf(R)))
    return R


def run_band_batch(scene: Scene, req: BandRequest, sun_dirs, sun_flux, sky_flux, direct_seeds,
                   out_device_ptr: Optional[int] = None, sky_peak_dirs=None) -> Optional[np.ndarray]:
    """K suns sharing bands/materials/diffuse seeds -> radiation_flux [K][P][B]; sky_peak_dirs [K][B][3]
    (anisotropic sky peak of each sun, None = the request's)."""
    D = _f32(sun_dirs)
    K = D.shape[0]
    D = D.reshape(K, -1, 3)
    F = _f32(sun_flux).reshape(K, D.shape[1], req.B)
    Q = _f32(sky_flux).reshape(K, req.B)
    seeds = _u32(direct_seeds).reshape(K)
    PK = None if sky_peak_dirs is None else _f32(sky_peak_dirs).reshape(K, req.B, 3)
    sb = rt_sun_batch(K, _pf(D), _pf(F), _pf(Q), _pu(seeds), _pf(PK))
    cs = req.c_struct()
    if out_device_ptr is not None:
        _check(scene.ctx.lib.rt_run_band_batch(scene.ctx._h, scene._h, C.byref(cs), C.byref(sb),
                                               C.c_void_p(out_device_ptr), 1))
        return None
    R = np.zeros((K, scene.P, req.B), np.float32)
    _check(scene.ctx.lib.rt_run_band_batch(scene.ctx._h, scene._h, C.byref(cs), C.byref(sb),
                                           R.ctypes.data_as(C.c_void_p), 0))
    return R


def solve_footprint(P: int, req: "BandRequest", nsuns: int) -> int:
    """Device bytes a server slot needs for a batch of nsuns suns of req over P primitives."""
    lib = load_library()
    out = C.c_uint64()
    cs = req.c_struct()
    _check(lib.rt_solve_footprint(P, C.byref(cs), nsuns, C.byref(out)))
    return out.value


ENERGY_ROWS = ("source", "sky", "emitted", "absorbed", "escaped", "leftover", "residual")
ENERGY_PASS_ROWS = ("radiated", "arrived", "sky", "escaped")


def set_energy_accounting(ctx: Context, on: bool = True):
    """FP64 energy bookkeeping of later solves on ctx (rt_set_energy_accounting)."""
    _check(ctx.lib.rt_set_energy_accounting(ctx._h, 1 if on else 0))


def energy_fp64(ctx: Context, nchannels: int) -> dict:
    """Totals of the last solve (it must have run with the bookkeeping on): {row name: [C] float64}, rows as
    ENERGY_ROWS, plus "passes": [npass][4][C] (ENERGY_PASS_ROWS per diffuse/scattering pass)."""
    out = np.zeros((len(ENERGY_ROWS), nchannels), np.float64)
    _check(ctx.lib.rt_energy_residual_fp64(ctx._h, out.ctypes.data_as(f64p), nchannels))
    npass = C.c_uint32()
    _check(ctx.lib.rt_energy_passes_fp64(ctx._h, None, nchannels, 0, C.byref(npass)))
    passes = np.zeros((npass.value, len(ENERGY_PASS_ROWS), nchannels), np.float64)
    if npass.value:
        _check(ctx.lib.rt_energy_passes_fp64(ctx._h, passes.ctypes.data_as(f64p), nchannels, npass.value,
                                             C.byref(npass)))
    res = {k: out[i] for i, k in enumerate(ENERGY_ROWS)}
    res["passes"] = passes
    return res


def sun_direction(zenith_deg: float, azimuth_deg: float) -> np.ndarray:
    """Direction towards the sun (Helios sphere2cart: azimuth from +y towards +x,
    core/src/global.cpp:833-835)."""
    z = math.radians(zenith_deg)
    a = math.radians(azimuth_deg)
    return np.array([math.sin(z) * math.sin(a), math.sin(z) * math.cos(a), math.cos(z)], np.float32)
