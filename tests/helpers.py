"""Shared builders for tests (scenes, requests)."""
from __future__ import annotations

import numpy as np

from helios_amd import native, scenes


def config1_request(P, depth=3, schedule=0, seed=7, dseeds=(11, 12, 13, 14), nrays=100, sun=(30.0, 0.0)):
    """Config 1 band PAR: sun zenith 30 deg az 0 at 1000 W/m^2, sky 100 W/m^2, leaf rho .10 tau .05,
    ground rho .15."""
    rho = np.r_[np.full(P - 1, 0.10), [0.15]].reshape(-1, 1)
    tau = np.r_[np.full(P - 1, 0.05), [0.0]].reshape(-1, 1)
    return native.BandRequest(P, [native.BandSpec("PAR", depth, False, 100.0)], rho=rho, tau=tau,
                              direct_rays=nrays, diffuse_rays=nrays, source_dirs=[native.sun_direction(*sun)],
                              source_flux=[[1000.0]], direct_seed=seed, diffuse_seeds=list(dseeds), schedule=schedule)


def two_band_request(s, depth=2, schedule=0, nrays=64, emission_lw=False):
    """Leaves: PAR (rho .10 tau .05), NIR (rho .45 tau .40); ground rho .15/.30."""
    P = s.P
    leaf = s.twosided == 1
    rho = np.stack([np.where(leaf, 0.45, 0.30), np.where(leaf, 0.10, 0.15)], 1)
    tau = np.stack([np.where(leaf, 0.40, 0.0), np.where(leaf, 0.05, 0.0)], 1)
    bands = [native.BandSpec("NIR", depth, False, 60.0), native.BandSpec("PAR", depth, False, 120.0)]
    return native.BandRequest(P, bands, rho=rho, tau=tau, direct_rays=nrays, diffuse_rays=nrays,
                              source_dirs=[native.sun_direction(40.0, 120.0)], source_flux=[[350.0, 500.0]],
                              direct_seed=21, diffuse_seeds=[31 + k for k in range(depth + 1)], schedule=schedule)


def small_canopy(seed=5, nleaves=3000, periodic=False):
    return scenes.patch_canopy(seed, nleaves, 0.06, 4.0, 1.5, 8, periodic=periodic)


def rel_err(a, b, floor=1e-3):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


class DeviceArray:
    """A device copy of a numpy array made with the HIP runtime the library itself links (libamdhip64.so.7),
    for tests of entry points that take device pointers.  Test plumbing only."""

    _hip = None

    @classmethod
    def _rt(cls):
        if cls._hip is None:
            import ctypes as C
            native.load_library()  # the library has loaded the runtime already; this returns the same one
            hip = C.CDLL("libamdhip64.so.7")
            hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            hip.hipFree.argtypes = [C.c_void_p]
            hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip = hip
        return cls._hip

    def __init__(self, a):
        import ctypes as C
        self.a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        hip = self._rt()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.a.nbytes, 4)) == 0
        assert hip.hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, 1) == 0  # host to device

    @property
    def addr(self) -> int:
        return int(self.ptr.value)

    def get(self) -> np.ndarray:
        out = np.empty_like(self.a)
        assert self._rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr.value:
            self._rt().hipFree(self.ptr)
            self.ptr.value = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
