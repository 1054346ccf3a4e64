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


def beer_scene(s):
    """Config 2's leaves over a ground that covers the whole periodic cell, with the leaves that dip below
    z = 0 lifted clear of it.  As generated, config 2's ground (10 x 10 m) is narrower than the periodic cell
    (the bounding box of all primitives, RadiationModel.cpp periodic domain, including leaf overhang) and ~1%
    of the leaves reach below the ground: a missed ray counts as sky in every direction, so leaves see sky
    through the strips and below the ground.  Beer's law describes the clean canopy."""
    leaf = s.twosided == 1
    assert np.all(s.types[leaf] == 1)
    T = s.transforms.reshape(-1, 16).astype(np.float64)[leaf]
    loc = ((0.0, 0.0), (0.0, 1.0), (1.0, 1.0))  # triangle vertices in the transform's unit frame
    vx = np.stack([T[:, 0] * x + T[:, 1] * y + T[:, 3] for x, y in loc])
    vy = np.stack([T[:, 4] * x + T[:, 5] * y + T[:, 7] for x, y in loc])
    vz = np.stack([T[:, 8] * x + T[:, 9] * y + T[:, 11] for x, y in loc])
    dom = (min(vx.min(), -5.0), max(vx.max(), 5.0), min(vy.min(), -5.0), max(vy.max(), 5.0))
    zmin = vz.min(0)
    T[:, 11] += np.where(zmin < 0.0, 1e-4 - zmin, 0.0)
    gx = np.linspace(dom[0], dom[1], 21)
    gy = np.linspace(dom[2], dom[3], 21)
    G = np.zeros((400, 16))
    for i in range(20):
        for j in range(20):
            G[20 * i + j, [0, 5, 10, 15]] = (gx[i + 1] - gx[i], gy[j + 1] - gy[j], 1.0, 1.0)
            G[20 * i + j, [3, 7]] = (0.5 * (gx[i] + gx[i + 1]), 0.5 * (gy[j] + gy[j + 1]))
    n = T.shape[0]
    out = scenes.SceneArrays(np.concatenate([T, G]).astype(np.float32), np.r_[np.ones(n), np.zeros(400)].astype(
        np.uint32), np.r_[np.ones(n), np.zeros(400)].astype(np.uint32))
    out.periodic = (True, True)
    return out, (dom[1] - dom[0]) * (dom[3] - dom[2])


def beer_config2(run, s, A, cell, n_diffuse, tol=0.006, check_leaves=True):
    """Leaf interception (per unit of the periodic cell) and ground-based interception of a canopy with black
    leaves and ground, for direct rays at two zeniths and for an isotropic sky, against Beer's law at the
    realised LAI (selfTest.cpp:1372-1468, whose tolerance is 0.01).  The Monte Carlo error is ~3e-4; what
    remains is the realisation's own coverage, which strays from Beer's mean: for a Boolean model of 200k
    leaves of ~5e-4 m^2 shadow over 100 m^2 its standard deviation is ~1.5e-3 (seen: up to 3.0e-3)."""
    leaf = s.twosided == 1
    lai = A[leaf].sum() / cell
    P = s.P
    zero = np.zeros((P, 1))
    out = []
    for th_deg in (20.0, 50.0):
        th = math.radians(th_deg)
        R = run(req1(P, "direct", zero, zero, d=100, sun=native.sun_direction(th_deg, 30.0), F=1.0 / math.cos(th)))
        out.append((R, cf.beer_direct(lai, th)))
    R = run(req1(P, "diffuse", zero, zero, n=n_diffuse, dflux=1.0))
    out.append((R, cf.beer_diffuse(lai)))
    res = []
    for R, exact in out:
        leaf_int = float(np.sum(R[leaf, 0] * A[leaf])) / cell
        ground_int = 1.0 - float(np.sum(R[~leaf, 0] * A[~leaf])) / float(np.sum(A[~leaf]))
        if check_leaves:
            assert abs(leaf_int - exact) <= tol, (leaf_int, exact)
        assert abs(ground_int - exact) <= tol, (ground_int, exact)
        res.append((leaf_int, ground_int, exact))
    return lai, res


def test_beer_law_config2_canopy():
    # a quarter of config 2's leaves (realised LAI ~0.43) keeps the CPU run short (seen: up to 3.4e-3)
    s, cell = beer_scene(scenes.config2(2, 50_000))
    o = OracleScene.from_arrays(s)
    dom = o.geometry()[3]
    assert abs(float(dom[1] - dom[0]) * float(dom[3] - dom[2]) - cell) <= 1e-3
    lai, _ = beer_config2(o.run_band, s, o.areas(), cell, 100)
    assert 0.4 < lai < 0.46
