"""The five workload configurations as ready-to-run requests (shared by bench.py and the tests).

Configs (one GPU):
  1  1,600 leaf patches + ground, PAR band, d = n = 10, depth 3
  2  200,000 leaf triangles, periodic 10x10x2 m canopy, NIR + PAR, d = n = 32, depth 5
  3  orchard of 25 trees (1,000,000 triangles + 1,600 ground patches), 64 suns x (LW, NIR, PAR),
     d = n = 15, depth 4 -- the diurnal series
  4  request stream of canopies with log-uniform 10k-2M primitives, 1-4 bands, own sun
  5  orchard tiled 5x (5 M primitives), castRays of uniform rays
"""
from __future__ import annotations

import math

import numpy as np

from . import native, scenes

DIURNAL_BANDS = ("LW", "NIR", "PAR")


def optical_arrays(s: scenes.SceneArrays, bands):
    """rho/tau/eps [P][B] for leaves (twosided 1), tubes (one-sided triangles) and ground (patches)."""
    leaf = s.twosided == 1
    ground = (s.twosided == 0) & (s.types == 0)
    P, B = s.P, len(bands)
    rho = np.zeros((P, B), np.float32)
    tau = np.zeros((P, B), np.float32)
    eps = np.ones((P, B), np.float32)
    for b, lab in enumerate(bands):
        if lab == "PAR":
            rho[:, b] = np.where(leaf, 0.10, np.where(ground, 0.15, 0.12))
            tau[:, b] = np.where(leaf, 0.05, 0.0)
        elif lab == "NIR":
            rho[:, b] = np.where(leaf, 0.45, np.where(ground, 0.30, 0.35))
            tau[:, b] = np.where(leaf, 0.40, 0.0)
        elif lab == "LW":
            eps[:, b] = 0.95
            rho[:, b] = 1.0 - eps[:, b]  # eps + rho + tau = 1 (R/src/RadiationModel.cpp:6296-6299)
            tau[:, b] = 0.0
    return rho, tau, eps


def config1_request(s, seed=7):
    rho, tau, _ = optical_arrays(s, ["PAR"])
    rho[-1, 0] = 0.15
    return native.BandRequest(s.P, [native.BandSpec("PAR", 3, False, 100.0)], rho=rho, tau=tau, direct_rays=100,
                              diffuse_rays=100, source_dirs=[native.sun_direction(30.0, 0.0)],
                              source_flux=[[1000.0]], direct_seed=seed, diffuse_seeds=[11, 12, 13, 14])


def config2_request(s, rays=1000, depth=5, black=False):
    bands = ["NIR", "PAR"]
    rho, tau, _ = optical_arrays(s, bands)
    if black:
        rho[:] = 0.0
        tau[:] = 0.0
    spec = [native.BandSpec("NIR", depth, False, 80.0), native.BandSpec("PAR", depth, False, 120.0)]
    return native.BandRequest(s.P, spec, rho=rho, tau=tau, direct_rays=rays, diffuse_rays=rays,
                              source_dirs=[native.sun_direction(35.0, 150.0)], source_flux=[[400.0, 550.0]],
                              direct_seed=2024, diffuse_seeds=[3000 + k for k in range(depth + 1)])


def diurnal_suns(K=64, seed=3):
    """K sun positions (zenith 10-85 deg, azimuth 60-300 deg) on a seeded schedule, with per-sun
    direct and sky fluxes for (LW, NIR, PAR)."""
    rng = np.random.default_rng(seed)
    zen = np.sort(10.0 + 75.0 * rng.random(K))[::-1]
    az = np.linspace(60.0, 300.0, K) + rng.uniform(-2.0, 2.0, K)
    dirs = np.stack([native.sun_direction(z, a) for z, a in zip(zen, az)])[:, None, :]
    cz = np.cos(np.radians(zen))
    atten = (0.3 + 0.7 * cz ** 0.5).astype(np.float32)
    flux = np.stack([np.zeros(K), 450.0 * atten, 520.0 * atten], 1).astype(np.float32)[:, None, :]
    sky = np.stack([np.full(K, 330.0), 70.0 + 30.0 * (1 - cz), 110.0 + 40.0 * (1 - cz)], 1).astype(np.float32)
    seeds = (10_000 + np.arange(K)).astype(np.uint32)
    return dirs.astype(np.float32), flux, sky, seeds


def diurnal_request(s, rays=200, depth=4):
    """Shared part of the diurnal series: bands LW (emission, eps .95), NIR, PAR; materials; diffuse seeds."""
    rho, tau, eps = optical_arrays(s, DIURNAL_BANDS)
    temp = np.where(s.twosided == 1, 298.0, 293.0).astype(np.float32)
    spec = [native.BandSpec("LW", depth, True, 330.0), native.BandSpec("NIR", depth, False, 80.0),
            native.BandSpec("PAR", depth, False, 120.0)]
    return native.BandRequest(s.P, spec, rho=rho, tau=tau, eps=eps, temperature=temp, direct_rays=rays,
                              diffuse_rays=rays, source_dirs=np.zeros((1, 3), np.float32),
                              source_flux=np.zeros((1, 3), np.float32), direct_seed=0,
                              diffuse_seeds=[20_000 + k for k in range(depth + 1)])


def single_sun_request(base: native.BandRequest, dirs, flux, sky, seed):
    """The plain (one-solve-per-launch) request of one sun of a batch."""
    spec = [native.BandSpec(b.label, b.scattering_depth, b.emission, float(q)) for b, q in zip(base.bands, sky)]
    return native.BandRequest(base.P, spec, rho=base.rho, tau=base.tau, eps=base.eps, temperature=base.temp,
                              direct_rays=base.direct_rays, diffuse_rays=base.diffuse_rays, source_dirs=dirs,
                              source_flux=flux, direct_seed=int(seed), diffuse_seeds=list(base.dseeds),
                              schedule=base.schedule)


def rays_per_solve(s, req) -> dict:
    """Traced rays of one plain solve: P*d^2*S (direct) + passes * n^2 * sum(faces) (diffuse+scatter)."""
    traced = (s.types == 0) | (s.types == 1) | (s.types == 3)
    d = math.ceil(math.sqrt(req.direct_rays))
    n = math.ceil(math.sqrt(req.diffuse_rays))
    faces = int(np.sum(traced)) + int(np.sum(traced & (s.twosided != 0)))
    depth = int(req.depth.max())
    S = req.sdir.shape[0]
    direct = int(np.sum(traced)) * d * d * S
    diffuse = (1 + depth) * n * n * faces
    return {"direct": direct, "diffuse": diffuse, "total": direct + diffuse, "d": d, "n": n, "passes": 1 + depth}


def serve_models(seed: int = 4, nscenes: int = 6, pmin: float = 1e4, pmax: float = 2e6, rays: int = 100,
                 depth: int = 2):
    """Config 4 models: seeded mixed canopies (square patches, triangles and triangle-fan disks) with P
    log-uniform in [pmin, pmax], each with 1-4 bands of its own optical properties and sky flux, one
    collimated source, d = n = ceil(sqrt(rays)), scattering depth `depth`.  Returns
    [(SceneArrays, BandRequest)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nscenes):
        P = int(round(10 ** rng.uniform(math.log10(pmin), math.log10(pmax))))
        s = scenes.mixed_canopy(100 + i, P)
        nb = int(rng.integers(1, 5))
        leaf = s.twosided == 1
        rho = np.zeros((s.P, nb), np.float32)
        tau = np.zeros((s.P, nb), np.float32)
        spec = []
        for b in range(nb):
            r, t = rng.uniform(0.05, 0.45), rng.uniform(0.0, 0.4)
            rho[:, b] = np.where(leaf, r, 0.15)
            tau[:, b] = np.where(leaf, t, 0.0)
            spec.append(native.BandSpec(f"band{b}", depth, False, float(rng.uniform(50, 200))))
        req = native.BandRequest(s.P, spec, rho=rho, tau=tau, direct_rays=rays, diffuse_rays=rays,
                                 source_dirs=[[0.0, 0.0, 1.0]], source_flux=[[500.0] * nb], direct_seed=1,
                                 diffuse_seeds=[7000 + 10 * i + k for k in range(depth + 1)])
        out.append((s, req))
    return out


def serve_requests(models, n: int, seed: int = 5):
    """n requests: (model index, sun dir [1][3], source flux [1][B], sky flux [B], direct seed), models
    drawn uniformly, suns uniform in zenith 10-80 deg and azimuth 0-360 deg."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m = int(rng.integers(len(models)))
        req = models[m][1]
        z, a = math.radians(rng.uniform(10, 80)), math.radians(rng.uniform(0, 360))
        d = np.array([[math.sin(z) * math.sin(a), math.sin(z) * math.cos(a), math.cos(z)]], np.float32)
        out.append((m, d, (req.sflux * rng.uniform(0.5, 1.5)).astype(np.float32),
                    (req.dflux * rng.uniform(0.5, 1.5)).astype(np.float32), int(rng.integers(1, 2**31))))
    return out
