"""Translucent (glass) covers in the oracle (CPU), pinned by the Fresnel + Bouguer slab model derived
independently and by the reference's glass tests (plugins/radiation/tests/selfTest.cpp:9595-9880), whose
scenes make every direct ray cross the covers at one known angle, so the expected fluxes are exact."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleScene, glass_tra
from tests import glass_cases as gc

F = gc.F_SUN


@pytest.mark.parametrize("KL", [0.0, 0.05, 0.3])
def test_glass_function_is_the_fresnel_slab(KL):
    for deg in (0.0, 0.03, 10.0, 30.0, 45.0, 60.0, 75.0, 84.0):
        c = math.cos(math.radians(deg))
        got = glass_tra(c, 1.5, KL)
        want = gc.fresnel_slab(c, 1.5, KL)
        assert np.allclose(got, want, atol=3e-6), (deg, got, want)
        assert abs(sum(got) - 1.0) <= 1e-6


def test_glass_function_reference_values():
    # selfTest.cpp:9595-9617: r0 = 0.04 at n = 1.5, tau(0) = (1 - r0) / (1 + r0); monotone towards grazing;
    # Bouguer absorption at KL = 0.05
    t0 = glass_tra(1.0, 1.5, 0.0)
    assert abs(t0[0] - 0.96 / 1.04) <= 1e-6 and t0[2] == 0.0
    taus = [glass_tra(math.cos(math.radians(a)), 1.5, 0.0)[0] for a in (0, 60, 75)]
    assert taus[0] > taus[1] > taus[2] and taus[2] < 0.7
    tA = glass_tra(1.0, 1.5, 0.05)
    assert tA[0] < t0[0] and tA[2] > 0.01


def _run(scene, req):
    return OracleScene(*scene).run_band(req)


def test_normal_incidence_transmittance():
    sc = gc.cover_over_receiver()
    R = _run(sc, gc.request(2, [1, 0]))
    assert abs(R[1, 0] - F * glass_tra(1.0, 1.5)[0]) <= 1e-4 * F


def test_oblique_transmittance():
    th = math.radians(30.0)
    sc = gc.cover_over_receiver(cover_size=8.0)
    R = _run(sc, gc.request(2, [1, 0], sun=(math.sin(th), 0.0, math.cos(th)), direct_rays=20000))
    assert abs(R[1, 0] - F * math.cos(th) * gc.fresnel_slab(math.cos(th), 1.5)[0]) <= 1e-4 * F


def test_double_layer_is_multiplicative():
    R = _run(gc.double_layer(), gc.request(3, [1, 1, 0]))
    assert abs(R[2, 0] - F * glass_tra(1.0, 1.5)[0] ** 2) <= 1e-4 * F


def test_mixed_band_cover_is_opaque_and_inactive_cover_blocks():
    sc = gc.cover_over_receiver()
    R = _run(sc, gc.request(2, [[1, 0], [0, 0]], bands=("LW", "SW")))  # glass in one band only
    assert R[1, 0] == 0.0 and R[1, 1] == 0.0
    R = _run(sc, gc.request(2, [0, 0], depth=0))  # no glass at all: an opaque black cover
    assert R[1, 0] == 0.0


def test_window_preserves_beam_direction():
    th = math.radians(45.0)
    R = _run(gc.window(th), gc.request(5, [0, 0, 1, 0, 0], sun=(math.sin(th), 0.0, math.cos(th)),
                                       direct_rays=40000))
    beam = F * math.cos(th) * gc.fresnel_slab(math.cos(th), 1.5)[0]
    assert abs(R[3, 0] - beam) <= 1e-4 * F  # every in-path ray crosses the window at 45 degrees
    assert R[4, 0] == 0.0                   # every ray of the receiver below the window meets the ceiling


def test_cover_takes_its_absorbed_and_reflected_shares():
    # depth 0: nothing is handed on, so radiation_flux = radiation_in + scatter.  The cover absorbs the sun
    # itself (its own rays, black as an origin) and takes alpha + rho of every receiver ray that crosses it,
    # pushed as the receiver's flux density: F (1 + alpha + rho) = F (2 - tau)
    sc = gc.cover_over_receiver()
    tau = glass_tra(1.0, 1.5, 0.05)[0]
    R = _run(sc, gc.request(2, [1, 0], KL=0.05, depth=0))
    assert abs(R[1, 0] - F * tau) <= 1e-4 * F
    assert abs(R[0, 0] - F * (2.0 - tau)) <= 1e-4 * F


def test_sky_through_a_cover():
    # a cover far wider than the receiver's view: the sky reaches the receiver through it at every angle, and
    # a hit behind a cover is not attenuated (the receiver's only surface hit is the cover, passed through)
    Qd = 100.0
    sc = gc.scene([gc.horiz((0, 0, 1), 4000, 4000), gc.horiz((0, 0, 0), 2, 2, tw=0)])
    R = _run(sc, gc.request(2, [1, 0], F=0.0, sky=Qd, diffuse_rays=250000, depth=0))
    assert abs(R[1, 0] - Qd * gc.sky_through_cover()) <= 2e-3 * Qd
