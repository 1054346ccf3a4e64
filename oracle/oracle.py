"""ctypes binding of the CPU oracle (``oracle/librt_oracle.so``).

Test infrastructure only: imported by tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librt_oracle.so")
f32p, u32p, f64p = C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.POINTER(C.c_double)


class orc_band_params(C.Structure):
    _fields_ = [("nbands", C.c_uint32), ("nsources", C.c_uint32), ("direct_rays", C.c_uint32),
                ("diffuse_rays", C.c_uint32), ("band_depth", u32p), ("emission", u32p), ("source_dir", f32p),
                ("source_flux", f32p), ("diffuse_flux", f32p), ("rho", f32p), ("tau", f32p), ("eps", f32p),
                ("temperature", f32p), ("direct_seed", C.c_uint32), ("diffuse_seed", u32p),
                ("schedule", C.c_uint32), ("source_type", u32p), ("source_width", f32p), ("source_rotation", f32p),
                ("diffuse_extinction", f32p), ("diffuse_peak_dir", f32p), ("diffuse_dist_norm", f32p),
                ("sky_params", f32p), ("is_glass", C.POINTER(C.c_uint8)), ("glass_n", f32p), ("glass_KL", f32p)]


class orc_camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 2), ("focal_length", C.c_float),
                ("lens_diameter", C.c_float), ("fov_aspect", C.c_float), ("viewplane_length", C.c_float),
                ("resolution", C.c_uint32 * 2), ("resolution_full", C.c_uint32 * 2), ("pixel_offset", C.c_uint32 * 2),
                ("antialiasing_samples", C.c_uint32), ("random_seed", C.c_uint32)]


class orc_camera_light(C.Structure):
    _fields_ = [("nbands", C.c_uint32), ("nsources", C.c_uint32), ("out_top", f32p), ("out_bottom", f32p),
                ("source_flux", f32p), ("source_pos", f32p), ("source_type", u32p), ("source_width", f32p),
                ("source_rotation", f32p), ("sky_params", f32p), ("sky_radiance", f32p), ("solar_radiance", f32p),
                ("camera_diffuse_flux", f32p), ("emission", u32p), ("sun", C.c_float * 3), ("solar_cos", C.c_float),
                ("spec", f32p), ("spec_weight", f32p), ("spec_exp", f32p), ("spec_scale", f32p),
                ("spec_mode", C.c_uint32)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_tea.restype = C.c_uint32
        L.orc_tea.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p]
        L.orc_tea16.restype = C.c_uint32
        L.orc_tea16.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_lcg_next.restype = C.c_uint32
        L.orc_lcg_next.argtypes = [C.c_uint32]
        L.orc_rnd.restype = C.c_float
        L.orc_rnd.argtypes = [u32p]
        L.orc_sincos_2pi.argtypes = [C.c_float, f32p, f32p]
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.c_uint32, f32p, u32p, u32p, u32p, C.c_int, C.c_int]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_scene_geometry.argtypes = [C.c_void_p, f32p, f32p, f32p, f32p, u32p]
        L.orc_trace.restype = C.c_uint32
        L.orc_trace.argtypes = [C.c_void_p, f32p, f32p, C.c_float, C.c_uint32, C.c_int, f32p]
        L.orc_trace_periodic.restype = C.c_uint32
        L.orc_trace_periodic.argtypes = [C.c_void_p, f32p, f32p, C.c_float, C.c_uint32]
        L.orc_direct_ray.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, f32p]
        L.orc_diffuse_ray.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, f32p, f32p]
        L.orc_diffuse_hits.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, u32p]
        L.orc_glass_tra.argtypes = [C.c_float, C.c_float, C.c_float, f32p]
        L.orc_set_predicates.argtypes = [C.c_int]
        L.orc_direct_sums.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, f64p]
        L.orc_diffuse_hits_range.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p]
        L.orc_direct_sums_range.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, C.c_uint32,
                                            C.c_uint32, f64p]
        L.orc_cast_rays.argtypes = [C.c_void_p, C.c_uint64, f32p, f32p, C.c_float, f32p, f32p, u32p, C.c_int]
        L.orc_run_band.restype = C.c_int
        L.orc_run_band.argtypes = [C.c_void_p, C.POINTER(orc_band_params), f64p, f64p]
        L.orc_direct_sums_src.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, u32p, f32p, f32p,
                                          f64p]
        L.orc_rotation_matrix.argtypes = [f32p, f32p]
        L.orc_areas.argtypes = [C.c_void_p, f64p]
        L.orc_scene_set_masks.restype = C.c_int
        L.orc_scene_set_masks.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint8), u32p, C.POINTER(C.c_int32),
                                          f32p, C.POINTER(C.c_int32)]
        L.orc_camera_pixels.argtypes = [C.c_void_p, C.POINTER(orc_camera), C.POINTER(orc_camera_light), f64p]
        L.orc_pixel_labels.argtypes = [C.c_void_p, C.POINTER(orc_camera), u32p, f32p]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _u(a):
    return np.ascontiguousarray(a, np.uint32)


class predicates:
    """Context manager: the oracle's intersection predicates inside the block ("reference" or "watertight")."""

    def __init__(self, mode: str):
        self.mode = {"reference": 0, "watertight": 1}[mode]

    def __enter__(self):
        lib().orc_set_predicates(self.mode)
        return self

    def __exit__(self, *exc):
        lib().orc_set_predicates(0)


def glass_tra(cos_theta, n, KL=0.0):
    """(tau, rho, alpha) of a glass sheet at incidence cosine cos_theta (orc_glass_tra)."""
    out = np.zeros(3, np.float32)
    lib().orc_glass_tra(float(cos_theta), float(n), float(KL), out.ctypes.data_as(f32p))
    return tuple(float(x) for x in out)


def tea(v0, v1, rounds=16, key=(0xa341316c, 0xc8013ea4, 0xad90777d, 0x7e95761e)):
    k = (C.c_uint32 * 4)(*key)
    out1 = C.c_uint32()
    r0 = lib().orc_tea(v0, v1, rounds, k, C.byref(out1))
    return r0, out1.value


def tea16(v0, v1):
    return lib().orc_tea16(v0, v1)


def lcg_next(x):
    return lib().orc_lcg_next(x)


def rnd_stream(state, n):
    s = C.c_uint32(state)
    return [lib().orc_rnd(C.byref(s)) for _ in range(n)]


def sincos_2pi(r):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos_2pi(r, C.byref(s), C.byref(c))
    return s.value, c.value


class OracleScene:
    def __init__(self, transforms, types, twosided, uuids=None, periodic=(False, False)):
        self.T = _f(transforms).reshape(-1, 16)
        self.types = _u(types)
        self.twos = _u(twosided)
        self.uuids = None if uuids is None else _u(uuids)
        self.P = self.T.shape[0]
        self.h = lib().orc_scene_create(self.P, self.T.ctypes.data_as(f32p), self.types.ctypes.data_as(u32p),
                                        self.twos.ctypes.data_as(u32p),
                                        None if self.uuids is None else self.uuids.ctypes.data_as(u32p),
                                        int(periodic[0]), int(periodic[1]))

    @classmethod
    def from_arrays(cls, s):
        return cls(s.transforms, s.types, s.twosided, None, s.periodic)

    def __del__(self):
        try:
            if self.h:
                lib().orc_scene_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def geometry(self):
        v = np.zeros((self.P, 4, 3), np.float32)
        n = np.zeros((self.P, 3), np.float32)
        fr = np.zeros((self.P, 4), np.float32)
        dom = np.zeros(4, np.float32)
        nb = C.c_uint32()
        lib().orc_scene_geometry(self.h, v.ctypes.data_as(f32p), n.ctypes.data_as(f32p), fr.ctypes.data_as(f32p),
                                 dom.ctypes.data_as(f32p), C.byref(nb))
        return v, n, fr, dom, nb.value

    def areas(self):
        a = np.zeros(self.P, np.float64)
        lib().orc_areas(self.h, a.ctypes.data_as(f64p))
        return a

    def trace(self, o, d, tmin=1e-5, skip=0xFFFFFFFF, brute=False):
        o, d = _f(o), _f(d)
        t = C.c_float()
        h = lib().orc_trace(self.h, o.ctypes.data_as(f32p), d.ctypes.data_as(f32p), tmin, skip, int(brute),
                            C.byref(t))
        return h, t.value

    def trace_periodic(self, o, d, tmin, skip):
        o, d = _f(o), _f(d)
        return lib().orc_trace_periodic(self.h, o.ctypes.data_as(f32p), d.ctypes.data_as(f32p), tmin, skip)

    def direct_ray(self, prim, lx, ly, d, seed):
        o = np.zeros(3, np.float32)
        lib().orc_direct_ray(self.h, prim, lx, ly, d, seed, o.ctypes.data_as(f32p))
        return o

    def diffuse_ray(self, prim, lx, ly, n, face, seed):
        o = np.zeros(3, np.float32)
        d = np.zeros(3, np.float32)
        lib().orc_diffuse_ray(self.h, prim, lx, ly, n, face, seed, o.ctypes.data_as(f32p), d.ctypes.data_as(f32p))
        return o, d

    def diffuse_hits(self, n, seed):
        out = np.zeros((self.P, 2, n * n), np.uint32)
        lib().orc_diffuse_hits(self.h, n, seed, out.ctypes.data_as(u32p))
        return out

    def diffuse_hits_range(self, n, seed, p0, p1):
        """diffuse_hits() of primitives [p0, p1) only: [p1 - p0][2][n*n]."""
        out = np.zeros((p1 - p0, 2, n * n), np.uint32)
        lib().orc_diffuse_hits_range(self.h, n, seed, p0, p1, out.ctypes.data_as(u32p))
        return out

    def direct_sums_range(self, d, seed, src_dirs, p0, p1):
        """direct_sums() of primitives [p0, p1) only: [p1 - p0][S][2]."""
        S = _f(src_dirs).reshape(-1, 3)
        out = np.zeros((p1 - p0, S.shape[0], 2), np.float64)
        lib().orc_direct_sums_range(self.h, d, seed, S.shape[0], S.ctypes.data_as(f32p), p0, p1,
                                    out.ctypes.data_as(f64p))
        return out

    def direct_sums(self, d, seed, src_dirs):
        S = _f(src_dirs).reshape(-1, 3)
        out = np.zeros((self.P, S.shape[0], 2), np.float64)
        lib().orc_direct_sums(self.h, d, seed, S.shape[0], S.ctypes.data_as(f32p), out.ctypes.data_as(f64p))
        return out

    def direct_sums_src(self, d, seed, positions, types=None, widths=None, rotations=None):
        """[P][S][2] direct strength sums for sources of any type (FP64 accumulation)."""
        X = _f(positions).reshape(-1, 3)
        S = X.shape[0]
        ty = None if types is None else _u(types).reshape(S)
        W = None if widths is None else _f(widths).reshape(S, 2)
        R = None if rotations is None else _f(rotations).reshape(S, 3)
        out = np.zeros((self.P, S, 2), np.float64)
        lib().orc_direct_sums_src(self.h, d, seed, S, X.ctypes.data_as(f32p),
                                  None if ty is None else ty.ctypes.data_as(u32p),
                                  None if W is None else W.ctypes.data_as(f32p),
                                  None if R is None else R.ctypes.data_as(f32p), out.ctypes.data_as(f64p))
        return out

    def set_masks(self, masks, mask_ids, uv=None, uv_ids=None):
        """Texture masks: list of 2-D arrays [height][width] (nonzero = opaque, row 0 = top), mask_ids [P]
        (-1 = none), uv [P][4][2], uv_ids [P] (< 0: parametric patch UVs)."""
        tex = np.concatenate([np.asarray(m).reshape(-1) != 0 for m in masks]).astype(np.uint8) if len(masks) \
            else np.zeros(1, np.uint8)
        sizes = _u([[np.shape(m)[1], np.shape(m)[0]] for m in masks] if len(masks) else np.zeros((1, 2)))
        mid = np.ascontiguousarray(mask_ids, np.int32).reshape(self.P)
        U = None if uv is None else _f(uv).reshape(self.P, 8)
        uid = None if uv_ids is None else np.ascontiguousarray(uv_ids, np.int32).reshape(self.P)
        pi = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))
        self._masks = (tex, sizes, mid, U, uid)
        rc = lib().orc_scene_set_masks(self.h, len(masks), tex.ctypes.data_as(C.POINTER(C.c_uint8)),
                                       sizes.ctypes.data_as(u32p), pi(mid), None if U is None else U.ctypes.data_as(f32p),
                                       pi(uid))
        if rc != 0:
            raise ValueError("mask id out of range")

    @staticmethod
    def _camera(cam, tile, seed):
        fx, fy = cam.resolution
        x0, y0, w, h = (0, 0, fx, fy) if tile is None else tile
        return orc_camera((C.c_float * 3)(*cam.position), (C.c_float * 2)(*cam.direction), cam.focal_length,
                          cam.lens_diameter, cam.fov_aspect, cam.viewplane_length, (C.c_uint32 * 2)(w, h),
                          (C.c_uint32 * 2)(fx, fy), (C.c_uint32 * 2)(x0, y0), cam.antialiasing_samples, int(seed))

    def camera_pixels(self, cam, out_top, out_bottom, source_flux, source_pos, sky, seed=0, tile=None,
                      source_types=None, source_widths=None, source_rotations=None, pixels=None, specular=None):
        """Camera radiation [B][H*W] (FP64 sums of the FP32 samples), added to `pixels` when given.
        sky: dict(params [B][4], radiance [B], sun [3], solar [B], solar_cos, diffuse_flux [B], emission [B]).
        specular: dict(spec [S][P][B], weight [S][B], exponent [P], scale [P] or None, mode 1|2), or None."""
        ot, ob = _f(out_top).reshape(self.P, -1), _f(out_bottom).reshape(self.P, -1)
        B = ot.shape[1]
        X = _f(source_pos).reshape(-1, 3)
        S = X.shape[0]
        F = _f(source_flux).reshape(S, B) if S else np.zeros((1, B), np.float32)
        ty = None if source_types is None else _u(source_types).reshape(S)
        W = None if source_widths is None else _f(source_widths).reshape(S, 2)
        R = None if source_rotations is None else _f(source_rotations).reshape(S, 3)
        prm = _f(sky["params"]).reshape(B, 4)
        srad, solar = _f(sky["radiance"]).reshape(B), _f(sky["solar"]).reshape(B)
        df = _f(sky.get("diffuse_flux", np.zeros(B))).reshape(B)
        em = _u(sky.get("emission", np.zeros(B))).reshape(B)
        pt = lambda a: None if a is None else a.ctypes.data_as(f32p)
        L = orc_camera_light(B, S, pt(ot), pt(ob), pt(F), pt(X), None if ty is None else ty.ctypes.data_as(u32p),
                             pt(W), pt(R), pt(prm), pt(srad), pt(solar), pt(df), em.ctypes.data_as(u32p),
                             (C.c_float * 3)(*[float(v) for v in sky["sun"]]), float(sky["solar_cos"]))
        keep = []
        if specular is not None:
            sp = _f(specular["spec"]).reshape(S, self.P, B)
            sw = _f(specular["weight"]).reshape(S, B)
            se = _f(specular["exponent"]).reshape(self.P)
            ss = _f(specular["scale"] if specular.get("scale") is not None else np.zeros(self.P)).reshape(self.P)
            keep = [sp, sw, se, ss]
            L.spec, L.spec_weight, L.spec_exp, L.spec_scale = pt(sp), pt(sw), pt(se), pt(ss)
            L.spec_mode = int(specular["mode"])
        fx, fy = cam.resolution
        out = np.zeros((B, fx * fy), np.float64) if pixels is None else pixels
        c = self._camera(cam, tile, seed)
        lib().orc_camera_pixels(self.h, C.byref(c), C.byref(L), out.ctypes.data_as(f64p))
        return out

    def pixel_labels(self, cam, tile=None, labels=None, depths=None):
        """(labels [H*W] uint32, depths [H*W] float32) of the pixel-centre rays."""
        fx, fy = cam.resolution
        lab = np.zeros(fx * fy, np.uint32) if labels is None else labels
        dep = np.zeros(fx * fy, np.float32) if depths is None else depths
        c = self._camera(cam, tile, 0)
        lib().orc_pixel_labels(self.h, C.byref(c), lab.ctypes.data_as(u32p), dep.ctypes.data_as(f32p))
        return lab, dep

    def cast_rays(self, O, D, max_dist=-1.0, brute=False):
        O = _f(O).reshape(-1, 3)
        D = _f(D).reshape(-1, 3)
        n = O.shape[0]
        t = np.full(n, np.nan, np.float32)
        nrm = np.full((n, 3), np.nan, np.float32)
        u = np.zeros(n, np.uint32)
        lib().orc_cast_rays(self.h, n, O.ctypes.data_as(f32p), D.ctypes.data_as(f32p), max_dist,
                            t.ctypes.data_as(f32p), nrm.ctypes.data_as(f32p), u.ctypes.data_as(u32p), int(brute))
        return t, nrm, u

    def run_band(self, req, diag=False):
        """req: helios_amd.native.BandRequest-like object (same fields)."""
        P, B = self.P, req.B
        keep = dict(depth=_u(req.depth), emission=_u(req.emission), sdir=_f(req.sdir), sflux=_f(req.sflux),
                    dflux=_f(req.dflux), rho=_f(req.rho), tau=_f(req.tau),
                    eps=None if req.eps is None else _f(req.eps),
                    temp=None if req.temp is None else _f(req.temp), dseeds=_u(req.dseeds),
                    stype=None if getattr(req, "stype", None) is None else _u(req.stype),
                    swidth=None if getattr(req, "swidth", None) is None else _f(req.swidth),
                    srot=None if getattr(req, "srot", None) is None else _f(req.srot),
                    kext=None if getattr(req, "sky_ext", None) is None else _f(req.sky_ext),
                    kpeak=None if getattr(req, "sky_peak", None) is None else _f(req.sky_peak),
                    knorm=None if getattr(req, "sky_norm", None) is None else _f(req.sky_norm),
                    kprague=None if getattr(req, "sky_prague", None) is None else _f(req.sky_prague),
                    glass=None if getattr(req, "is_glass", None) is None else
                    np.ascontiguousarray(req.is_glass, np.uint8),
                    gn=None if getattr(req, "glass_n", None) is None else _f(req.glass_n),
                    gkl=None if getattr(req, "glass_KL", None) is None else _f(req.glass_KL))
        p = orc_band_params(
            B, keep["sdir"].reshape(-1, 3).shape[0], req.direct_rays, req.diffuse_rays,
            keep["depth"].ctypes.data_as(u32p), keep["emission"].ctypes.data_as(u32p),
            keep["sdir"].ctypes.data_as(f32p), keep["sflux"].ctypes.data_as(f32p), keep["dflux"].ctypes.data_as(f32p),
            keep["rho"].ctypes.data_as(f32p), keep["tau"].ctypes.data_as(f32p),
            None if keep["eps"] is None else keep["eps"].ctypes.data_as(f32p),
            None if keep["temp"] is None else keep["temp"].ctypes.data_as(f32p),
            req.direct_seed, keep["dseeds"].ctypes.data_as(u32p), req.schedule,
            None if keep["stype"] is None else keep["stype"].ctypes.data_as(u32p),
            None if keep["swidth"] is None else keep["swidth"].ctypes.data_as(f32p),
            None if keep["srot"] is None else keep["srot"].ctypes.data_as(f32p),
            *[None if keep[k] is None else keep[k].ctypes.data_as(f32p) for k in ("kext", "kpeak", "knorm", "kprague")],
            None if keep["glass"] is None else keep["glass"].ctypes.data_as(C.POINTER(C.c_uint8)),
            *[None if keep[k] is None else keep[k].ctypes.data_as(f32p) for k in ("gn", "gkl")])
        R = np.zeros((P, B), np.float64)
        dg = np.zeros((3, B), np.float64)
        rc = lib().orc_run_band(self.h, C.byref(p), R.ctypes.data_as(f64p), dg.ctypes.data_as(f64p))
        assert rc == 0
        return (R, dg) if diag else R
