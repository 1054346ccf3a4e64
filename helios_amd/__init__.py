"""MI355X (gfx950) radiation-transfer backend for the Helios RadiationModel / CollisionDetection path.

The product is ``libhelios_rt.so`` (C ABI in ``include/helios_rt.h``); this package holds its
ctypes binding (``native``), the seeded scene generators (``scenes``), the request server
(``server``) and the build script (``build``).
"""
from .native import (BandRequest, BandSpec, Context, RTError, Scene, energy_fp64, load_library, run_band,
                     run_band_batch, sun_direction)

__all__ = ["BandRequest", "BandSpec", "Context", "RTError", "Scene", "energy_fp64", "load_library", "run_band",
           "run_band_batch", "sun_direction"]
