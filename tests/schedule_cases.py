"""Scenes and requests whose closed forms pin the runBand pass schedule (RadiationModel.cpp:3984-4356).

The reference schedule's copy condition at the first scattering iteration (`s > 0 || (emission && rundirect)`,
RadiationModel.cpp:4286) decides whether the primary diffuse pass's scatter reaches radiation_out.  The cases
below are built so that each branch gives a different closed form:

* `squares_sw`: the reference's 90-degree common-edge squares (selfTest.cpp:585-668) with the shortwave band
  alone plus an isotropic sky -- direct + diffuse + scattering with emission off, the branch configs 1-3's
  shortwave bands run (no copy at s = 0);
* `squares_lw_sw`: the same squares with the emitting longwave band launched beside the shortwave one, as the
  reference test runs them (emission on + direct: the copy happens at s = 0);
* `gray_rectangles`: the gray parallel rectangles (selfTest.cpp:731-807), emission only, five scattering
  iterations -- the multi-bounce case, where both schedules agree.
"""
from __future__ import annotations

import numpy as np

from helios_amd import native
from tests import closed_forms as cf

Q, QD, RHO0 = 1000.0, 100.0, 0.3
TEMP = 300.0


def squares_scene():
    return cf.common_edge_squares()


def squares_sw(schedule: int, rays: int = 1000000) -> native.BandRequest:
    """Band SW only: sun at the zenith (Q), sky QD, patch 0 reflectivity RHO0, patch 1 black, depth 1."""
    rho = np.array([[RHO0], [0.0]])
    tau = np.zeros((2, 1))
    return native.BandRequest(2, [native.BandSpec("SW", 1, False, QD)], rho=rho, tau=tau, direct_rays=rays // 100,
                              diffuse_rays=rays, source_dirs=[[0.0, 0.0, 1.0]], source_flux=[[Q]], direct_seed=5,
                              diffuse_seeds=[17, 18], schedule=schedule)


def squares_sw_exact(schedule: int):
    _, _, _, F = squares_scene()
    return cf.common_edge_squares_sw(Q, QD, RHO0, F, schedule)


def squares_lw_sw(schedule: int, rays: int = 1000000) -> native.BandRequest:
    """Bands LW (emitting, depth 0; patch 0 at 300 K, patch 1 at 0 K, emissivity 1) and SW (emission off, depth
    1, sun Q, no sky), launched together as in the reference test."""
    rho = np.array([[0.0, RHO0], [0.0, 0.0]])
    tau = np.zeros((2, 2))
    bands = [native.BandSpec("LW", 0, True, 0.0), native.BandSpec("SW", 1, False, 0.0)]
    return native.BandRequest(2, bands, rho=rho, tau=tau, direct_rays=rays // 100, diffuse_rays=rays,
                              source_dirs=[[0.0, 0.0, 1.0]], source_flux=[[0.0, Q]], direct_seed=5,
                              diffuse_seeds=[27, 28], eps=np.ones((2, 2)), temperature=[TEMP, 0.0],
                              schedule=schedule)


def squares_lw_sw_exact():
    """[patch][band] under both schedules: with emission in the launch the reference copies at s = 0, so the
    reflected sun reaches patch 1 once; patch 1 absorbs F sigma T^4 of patch 0's emission, patch 0 nothing."""
    _, _, _, F = squares_scene()
    return np.array([[0.0, (1 - RHO0) * Q], [F * cf.SIGMA * TEMP ** 4, F * RHO0 * Q]])


def gray_rectangles(schedule: int, rays: int = 1000000, eps: float = 0.6, depth: int = 5):
    T, ty, tw, F = cf.parallel_rectangles()
    req = native.BandRequest(2, [native.BandSpec("LW", depth, True, 0.0)], rho=np.full((2, 1), 1.0 - eps),
                             tau=np.zeros((2, 1)), direct_rays=0, diffuse_rays=rays,
                             diffuse_seeds=[41 + k for k in range(depth + 1)], eps=np.full((2, 1), eps),
                             temperature=[TEMP, TEMP], schedule=schedule)
    return (T, ty, tw), req, cf.gray_parallel_rectangles_lw(eps, TEMP, F)
