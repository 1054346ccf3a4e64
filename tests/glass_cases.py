"""Scenes, requests and closed forms of translucent (glass) covers, after the reference's glass tests
(plugins/radiation/tests/selfTest.cpp:9595-9880), shared by the oracle and GPU tests."""
from __future__ import annotations

import math

import numpy as np

from helios_amd import native
from tests import closed_forms as cf

F_SUN = 1000.0


def fresnel_slab(cos_theta, n, KL=0.0):
    """(tau, rho, alpha) of a dielectric slab in air, derived here from the Fresnel equations in cosine form
    (r_s = ((cos i - n cos t) / (cos i + n cos t))^2, r_p = ((cos t - n cos i) / (cos t + n cos i))^2), the
    Bouguer factor exp(-KL / cos t) and the geometric series of internal reflections, per polarisation,
    averaged -- an independent statement of the model the reference uses (Duffie & Beckman)."""
    c = min(1.0, max(1e-4, cos_theta))
    s_t = math.sqrt(max(0.0, 1.0 - c * c)) / n
    ct = math.sqrt(max(0.0, 1.0 - s_t * s_t))
    rs = ((c - n * ct) / (c + n * ct)) ** 2
    rp = ((ct - n * c) / (ct + n * c)) ** 2
    ta = math.exp(-KL / ct)
    T = R = 0.0
    for r in (rs, rp):
        den = 1.0 - (r * ta) ** 2
        T += 0.5 * ta * (1 - r) ** 2 / den
        R += 0.5 * (r + r * (1 - r) ** 2 * ta * ta / den)
    return T, R, 1.0 - T - R


def scene(prims):
    """prims: list of (center, ex, ey, twosided) patches -> (T, types, twosided)."""
    T = np.stack([cf.patch_T(c, ex, ey) for c, ex, ey, _ in prims])
    return T, np.zeros(len(prims), np.uint32), np.array([tw for *_r, tw in prims], np.uint32)


def horiz(c, w, h, tw=1):
    return (c, (w, 0, 0), (0, h, 0), tw)


def request(P, glass, n_glass=1.5, KL=0.0, sun=(0.0, 0.0, 1.0), F=F_SUN, sky=0.0, depth=1, bands=("SW",),
            direct_rays=10000, diffuse_rays=2500, rho=None, tau=None, schedule=0):
    """glass[P][B] flags; every non-glass primitive black unless rho/tau given."""
    B = len(bands)
    g = np.asarray(glass, np.uint8).reshape(P, B)
    spec = [native.BandSpec(b, depth, False, sky) for b in bands]
    return native.BandRequest(P, spec, rho=np.zeros((P, B)) if rho is None else rho,
                              tau=np.zeros((P, B)) if tau is None else tau, direct_rays=direct_rays,
                              diffuse_rays=diffuse_rays, source_dirs=[sun], source_flux=[[F] * B], direct_seed=3,
                              diffuse_seeds=[11 + k for k in range(depth + 1)], is_glass=g,
                              glass_n=np.full((P, B), n_glass), glass_KL=np.full((P, B), KL), schedule=schedule)


def cover_over_receiver(cover_size=4.0):
    """Cover (two-sided) at z = 1 over a one-sided 2 x 2 black receiver at z = 0 (selfTest.cpp:9621-9650)."""
    return scene([horiz((0, 0, 1), cover_size, cover_size), horiz((0, 0, 0), 2, 2, tw=0)])


def double_layer():
    return scene([horiz((0, 0, 2), 4, 4), horiz((0, 0, 1), 4, 4), horiz((0, 0, 0), 2, 2, tw=0)])


def window(theta=math.radians(45.0), dz=2.0):
    """Opaque ceiling with a glass window (selfTest.cpp:9719-9778): [ceil_left, ceil_right, window, in-path
    receiver at x = -dz tan(theta), receiver below the window]."""
    shift = dz * math.tan(theta)
    return scene([horiz((-3.0, 0, dz), 4, 6), horiz((3.0, 0, dz), 4, 6), horiz((0, 0, dz), 2, 6),
                  horiz((-shift, 0, 0), 1, 1, tw=0), horiz((0, 0, 0), 1, 1, tw=0)])


def sky_through_cover(n_glass=1.5, nq=4000):
    """Isotropic sky seen through a cover that spans the receiver's whole upper hemisphere (up to 89.99
    degrees): cos-weighted hemisphere average of the transmittance, by midpoint quadrature."""
    tot = 0.0
    for k in range(nq):
        th = (k + 0.5) * (0.5 * math.pi / nq)
        tot += fresnel_slab(math.cos(th), n_glass)[0] * 2.0 * math.sin(th) * math.cos(th) * (0.5 * math.pi / nq)
    return tot
