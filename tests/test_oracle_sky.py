"""The oracle's anisotropic diffuse sky (sky_distribution.glsl:40-116, diffuse_raygen.comp:613-625) against
quadrature (CPU).

A diffuse ray leaves a primitive with a cosine-weighted direction and, when it reaches the sky, carries
fd(direction) x the sky flux / N, so the flux a lone black patch absorbs is sky flux x the integral of
fd (n.w) / pi over its hemisphere.  Normalising fd so that the upper-hemisphere integral of fd cos(theta)
/ pi is 1 (what the normalisation parameters are for), a horizontal patch absorbs the sky flux whatever
the pattern; a tilted one absorbs the quadrature value of its own hemisphere, which depends on where the
pattern's peak is -- a sign, frame or unit slip moves it by far more than the Monte Carlo error."""
import math

import numpy as np

from helios_amd import native
from oracle.oracle import OracleScene
from tests import closed_forms as cf

PEAK = np.array([math.sin(math.radians(35)) * math.cos(math.radians(20)),
                 math.sin(math.radians(35)) * math.sin(math.radians(20)), math.cos(math.radians(35))])


def _grid(nt=400, npp=800, lower=False):
    th = (np.arange(nt) + 0.5) / nt * (math.pi if lower else math.pi / 2)
    ph = (np.arange(npp) + 0.5) / npp * 2 * math.pi
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    w = np.sin(T) * (th[1] - th[0]) * (ph[1] - ph[0])
    d = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1)
    return d, w


def power_law(d, K, norm):
    psi = np.maximum(np.arccos(np.clip(d @ PEAK, -1, 1)), math.pi / 180)
    return psi ** -K * norm


def prague(d, prm):
    gamma = np.degrees(np.arccos(np.clip(d @ PEAK, -1, 1)))
    return (1 + prm[0] * np.exp(-gamma / prm[1])) * (1 + (prm[2] - 1) * (1 - np.maximum(d[..., 2], 0))) * prm[3] * math.pi


def _norms(K, prm):
    d, w = _grid()
    cz = d[..., 2]
    pl = math.pi / float((power_law(d, K, 1.0) * cz * w).sum())
    pr = 1.0 / float((prague(d, list(prm[:3]) + [1.0 / math.pi]) * cz * w).sum())
    return pl, pr


def _patch_flux(normal_tilt_deg, K, prm, n=48):
    a = math.radians(normal_tilt_deg)
    ex = (1.0, 0.0, 0.0)
    ey = (0.0, math.cos(a), math.sin(a))
    o = OracleScene(cf.patch_T((0, 0, 0), ex, ey)[None], [0], [0])
    bands = [native.BandSpec("PL", 0, False, 100.0), native.BandSpec("PR", 0, False, 100.0)]
    pl, pr = _norms(K, prm)
    req = native.BandRequest(1, bands, rho=np.zeros((1, 2)), tau=np.zeros((1, 2)), direct_rays=1, diffuse_rays=n * n,
                             diffuse_seeds=[5], sky_extinction=[K, 0.0], sky_peak_dir=[PEAK, PEAK],
                             sky_norm=[pl, 0.0], sky_prague=[[0, 1, 1, 0], list(prm[:3]) + [pr]])
    return o.run_band(req)[0], (pl, pr), np.cross(ex, ey)


def test_normalised_skies_deliver_the_sky_flux_to_a_horizontal_patch():
    prm = (3.0, 25.0, 1.8)
    R, _, _ = _patch_flux(0.0, 0.7, prm)
    assert abs(R[0] - 100.0) < 3.0, R
    assert abs(R[1] - 100.0) < 2.0, R


def test_tilted_patch_matches_hemisphere_quadrature():
    prm = (3.0, 25.0, 1.8)
    for tilt in (60.0, 120.0):
        R, (pl, pr), nrm = _patch_flux(tilt, 0.7, prm)
        d, w = _grid(lower=True)
        c = d @ nrm
        m = c > 0
        e_pl = 100.0 * float((power_law(d, 0.7, pl) * c * w)[m].sum()) / math.pi
        e_pr = 100.0 * float((prague(d, list(prm) + [pr]) * c * w)[m].sum()) / math.pi
        assert abs(R[0] - e_pl) < 0.04 * e_pl, (tilt, R[0], e_pl)
        assert abs(R[1] - e_pr) < 0.03 * e_pr, (tilt, R[1], e_pr)
    # the peak lies towards +y: a normal tilted towards +y (tilt < 0) receives more than one tilted to -y
    toward, _, n1 = _patch_flux(-60.0, 0.7, prm)
    away, _, n2 = _patch_flux(60.0, 0.7, prm)
    assert n1[1] > 0 > n2[1]
    assert toward[0] > 1.1 * away[0] and toward[1] > away[1]
