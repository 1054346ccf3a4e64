"""The oracle's runBand against the reference's closed-form tests (CPU).

plugins/radiation/tests/selfTest.cpp: Simple Direct (:518-545), Black Parallel Rectangles (:670-729),
Second Law Equilibrium (:1061-1132), Homogeneous Canopy / Beer's law with periodic boundaries
(:1372-1468, :1625-1717).  Single seeded runs with stratified rays replace the reference's ensembles."""
import math

import numpy as np

from helios_amd import native, scenes
from oracle.oracle import OracleScene
from tests import closed_forms as cf


def req1(P, band, rho, tau, d=0, n=0, sun=None, F=None, depth=0, emission=False, dflux=0.0, eps=None, temp=None,
         seeds=(1,), schedule=0):
    return native.BandRequest(P, [native.BandSpec(band, depth, emission, dflux)], rho=rho, tau=tau, direct_rays=d,
                              diffuse_rays=n, source_dirs=None if sun is None else [sun],
                              source_flux=None if F is None else [[F]], direct_seed=3,
                              diffuse_seeds=list(seeds) + [100 + k for k in range(depth + 1)], eps=eps,
                              temperature=temp, schedule=schedule)


def test_simple_direct():
    T = cf.patch_T((0, 0, 0), (1, 0, 0), (0, 1, 0))[None]
    o = OracleScene(T, [0], [0])
    R = o.run_band(req1(1, "SW", [[0.0]], [[0.0]], d=10000, sun=[0, 0, 1], F=1000.0))
    assert abs(R[0, 0] - 1000.0) / 1000.0 <= 0.01
    assert abs(R[0, 0] - 1000.0) < 1e-3  # every ray is unoccluded: exact up to rounding


def test_direct_cosine_law():
    T = cf.patch_T((0, 0, 0), (1, 0, 0), (0, 1, 0))[None]
    o = OracleScene(T, [0], [0])
    for z in (0.0, 30.0, 60.0, 85.0):
        R = o.run_band(req1(1, "SW", [[0.0]], [[0.0]], d=100, sun=native.sun_direction(z, 40.0), F=500.0))
        assert abs(R[0, 0] - 500.0 * math.cos(math.radians(z))) < 1e-3


def test_black_parallel_rectangles():
    T, ty, tw, F12 = cf.parallel_rectangles()
    o = OracleScene(T, ty, tw)
    R = o.run_band(req1(2, "SW", np.zeros((2, 1)), np.zeros((2, 1)), n=40000, dflux=1.0))
    exact = 1.0 - F12
    assert np.max(np.abs(R[:, 0] - exact)) / exact <= 0.005


def test_second_law_enclosure():
    T, ty, tw = cf.closed_box()
    P = T.shape[0]
    eps = 0.8
    o = OracleScene(T, ty, tw)
    R = o.run_band(req1(P, "LW", np.full((P, 1), 1 - eps), np.zeros((P, 1)), n=10000, depth=5, emission=True,
                        eps=np.full((P, 1), eps), temp=np.full(P, 300.0)))
    ref = eps * cf.SIGMA * 300.0 ** 4
    assert np.mean(np.abs(R[:, 0] - ref) / ref) <= 0.005


def _beer_canopy():
    # LAI 2 of 7.5 cm leaves in a periodic 6 x 6 m domain (the reference uses 50 x 50 m; periodic
    # wrapping removes the edges, the smaller cell leaves ~0.5% realisation noise), ground of
    # 10 x 10 one-sided patches
    D, w, lai = 6.0, 0.075, 2.0
    nleaves = int(round(lai * D * D / w / w))
    s = scenes.patch_canopy(11, nleaves, w, D, 3.0, 10, periodic=True, inset=True)
    return s, D, lai


def test_beer_law_periodic_canopy_direct_and_diffuse():
    s, D, lai_nominal = _beer_canopy()
    o = OracleScene.from_arrays(s)
    A = o.areas()
    leaf = s.twosided == 1
    # the periodic cell is the vertex bounding box (leaves overhang the ground), so the realised LAI
    # and the leaf-based interception are per unit area of that cell
    dom = o.geometry()[3]
    cell = float(dom[1] - dom[0]) * float(dom[3] - dom[2])
    lai = A[leaf].sum() / cell
    th = 0.2 * math.pi
    P = s.P
    Rd = o.run_band(req1(P, "direct", np.zeros((P, 1)), np.zeros((P, 1)), d=100,
                         sun=native.sun_direction(math.degrees(th), 0.0), F=1.0 / math.cos(th)))
    Rf = o.run_band(req1(P, "diffuse", np.zeros((P, 1)), np.zeros((P, 1)), n=400, dflux=1.0))
    for R, exact in ((Rd, cf.beer_direct(lai, th)), (Rf, cf.beer_diffuse(lai))):
        leaf_int = float(np.sum(R[leaf, 0] * A[leaf])) / cell
        ground_int = 1.0 - float(np.sum(R[~leaf, 0] * A[~leaf])) / float(np.sum(A[~leaf]))
        assert abs(leaf_int - exact) <= 0.01, (leaf_int, exact)
        assert abs(ground_int - exact) <= 0.01, (ground_int, exact)


def test_schedules_agree_without_direct():
    # emission-only enclosure: the reference schedule already propagates every pass, so both
    # schedules must give identical results (RadiationModel.cpp:4225-4234, 4286-4288)
    T, ty, tw = cf.closed_box(m=3)
    P = T.shape[0]
    o = OracleScene(T, ty, tw)
    kw = dict(n=2500, depth=3, emission=True, eps=np.full((P, 1), 0.5), temp=np.full(P, 310.0))
    R0 = o.run_band(req1(P, "LW", np.full((P, 1), 0.5), np.zeros((P, 1)), schedule=0, **kw))
    R1 = o.run_band(req1(P, "LW", np.full((P, 1), 0.5), np.zeros((P, 1)), schedule=1, **kw))
    assert np.array_equal(R0, R1)


def test_reference_schedule_double_counts_first_scatter():
    # direct + diffuse + scattering with emission off (configs 1-3's shortwave bands): the reference
    # schedule gathers the direct-scattered flux in the primary pass AND again at s = 0, and drops the
    # primary pass's own scatter (RadiationModel.cpp:4188-4189, 4286-4289).  On a lone patch under a
    # sky there is nothing to gather, so both schedules give F cos + Q (1 - rho) exactly.
    T = cf.patch_T((0, 0, 0), (1, 0, 0), (0, 1, 0))[None]
    o = OracleScene(T, [0], [0])
    for sched in (0, 1):
        R = o.run_band(req1(1, "SW", [[0.3]], [[0.0]], d=16, n=16, sun=[0, 0, 1], F=100.0, depth=2, dflux=50.0,
                            schedule=sched))
        assert abs(R[0, 0] - (100.0 * 0.7 + 50.0 * 0.7)) < 1e-3
