"""Scenes and analytic answers of the reference's closed-form radiation tests
(plugins/radiation/tests/selfTest.cpp), restated with explicit transforms."""
from __future__ import annotations

import math

import numpy as np

SIGMA = 5.6703744e-8


def patch_T(center, ex, ey):
    """Row-major transform mapping the canonical patch (-0.5..0.5)^2 onto a patch spanned by ex, ey."""
    T = np.zeros(16, np.float32)
    T[0], T[4], T[8] = ex
    T[1], T[5], T[9] = ey
    T[2], T[6], T[10] = np.cross(ex, ey) / max(np.linalg.norm(np.cross(ex, ey)), 1e-30)
    T[3], T[7], T[11] = center
    T[15] = 1.0
    return T


def voxel_T(center, size, rotation=0.0):
    """Row-major transform of a voxel as Context::addVoxel forms it (core/src/Context_primitive.cpp:256-269):
    scale by size, rotate by ``rotation`` about z, translate to center; the unit cube is (-0.5..0.5)^3."""
    c, s = np.cos(rotation), np.sin(rotation)
    T = np.zeros(16, np.float32)
    T[0], T[1], T[4], T[5] = c * size[0], -s * size[1], s * size[0], c * size[1]
    T[10] = size[2]
    T[3], T[7], T[11] = center
    T[15] = 1.0
    return T


def voxel_box(T):
    """FP64 axis-aligned box of a voxel's eight transformed corners."""
    M = np.asarray(T, np.float64).reshape(4, 4)
    C = np.array([[x, y, z, 1.0] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)])
    W = C @ M.T
    return W[:, :3].min(0), W[:, :3].max(0)


def tri_T(v0, v1, v2):
    v0, v1, v2 = (np.asarray(v, np.float32) for v in (v0, v1, v2))
    T = np.zeros(16, np.float32)
    c0, c1 = v2 - v1, v1 - v0
    n = np.cross(c0, c1)
    n = n / max(np.linalg.norm(n), 1e-30)
    T[0], T[4], T[8] = c0
    T[1], T[5], T[9] = c1
    T[2], T[6], T[10] = n
    T[3], T[7], T[11] = v0
    T[15] = 1.0
    return T


def parallel_rectangles(a=1.0, b=2.0, c=0.5):
    """Two one-sided a x b rectangles facing each other at distance c (selfTest.cpp:670-729)."""
    T = np.stack([patch_T((0, 0, 0), (a, 0, 0), (0, b, 0)),           # faces +z
                  patch_T((0, 0, c), (0, b, 0), (a, 0, 0))])           # faces -z
    X, Y = a / c, b / c
    X2, Y2 = X * X, Y * Y
    F12 = 2.0 / (math.pi * X * Y) * (math.log(math.sqrt((1 + X2) * (1 + Y2) / (1 + X2 + Y2)))
                                     + X * math.sqrt(1 + Y2) * math.atan(X / math.sqrt(1 + Y2))
                                     + Y * math.sqrt(1 + X2) * math.atan(Y / math.sqrt(1 + X2))
                                     - X * math.atan(X) - Y * math.atan(Y))
    return T, np.zeros(2, np.uint32), np.zeros(2, np.uint32), F12


def closed_box(L=10.0, m=5):
    """Closed cube of side L, each face split m x m, patches facing inwards (selfTest.cpp:1061-1132)."""
    h, s = L / 2, L / m
    Ts = []
    e = np.eye(3, dtype=np.float32) * s
    faces = [  # (axis, sign of position, ex, ey) with cross(ex, ey) pointing inwards
        (2, -1, e[0], e[1]), (2, +1, e[1], e[0]), (0, -1, e[1], e[2]), (0, +1, e[2], e[1]),
        (1, -1, e[2], e[0]), (1, +1, e[0], e[2])]
    for axis, sgn, ex, ey in faces:
        for j in range(m):
            for i in range(m):
                u = -h + (i + 0.5) * s
                v = -h + (j + 0.5) * s
                c = np.zeros(3, np.float32)
                c[axis] = sgn * h
                # place (u, v) along the directions of ex, ey
                c = c + (ex / s) * u + (ey / s) * v
                Ts.append(patch_T(c, ex, ey))
    T = np.stack(Ts)
    return T, np.zeros(len(Ts), np.uint32), np.zeros(len(Ts), np.uint32)


def beer_direct(lai, theta_s):
    """Fraction of collimated flux intercepted by a spherical-LAD canopy (selfTest.cpp:1462)."""
    return 1.0 - math.exp(-0.5 * lai / math.cos(theta_s))


def beer_diffuse(lai, N=50):
    """Fraction of isotropic diffuse flux intercepted (midpoint rule, selfTest.cpp:1453-1460)."""
    dth = 0.5 * math.pi / N
    tot = 0.0
    for i in range(N):
        th = (i + 0.5) * dth
        tot += 2.0 * (1.0 - math.exp(-0.5 * lai / math.cos(th))) * math.cos(th) * math.sin(th) * dth
    return tot


def perpendicular_plates_F(W=1.0, H=1.0):
    """View factor from a 1 x W plate to a perpendicular 1 x H plate sharing the edge of length 1 (both
    lengths in units of the common edge; Howell's catalogue, case C-14).  0.20004 for two unit squares."""
    A, B = W * W, H * H
    Cc = A + B
    t = W * math.atan(1.0 / W) + H * math.atan(1.0 / H) - math.sqrt(Cc) * math.atan(1.0 / math.sqrt(Cc))
    lg = 0.25 * math.log(((1 + A) * (1 + B) / (1 + Cc)) * (A * (1 + Cc) / ((1 + A) * Cc)) ** A
                         * (B * (1 + Cc) / ((1 + B) * Cc)) ** B)
    return (t + lg) / (math.pi * W)


def common_edge_squares():
    """Two one-sided unit squares at 90 degrees sharing an edge (selfTest.cpp:585-668): patch 0 horizontal at
    z = 0 facing +z, patch 1 in the plane x = 0.5 facing -x (towards the space above patch 0)."""
    T = np.stack([patch_T((0, 0, 0), (1, 0, 0), (0, 1, 0)),
                  patch_T((0.5, 0, 0.5), (0, 0, 1), (0, 1, 0))])
    return T, np.zeros(2, np.uint32), np.zeros(2, np.uint32), perpendicular_plates_F()


def common_edge_squares_sw(Q, Qd, rho0, F, schedule):
    """Absorbed shortwave flux of the two squares above, sun at the zenith (flux Q, patch 1 edge-on), isotropic
    sky Qd, patch 0 reflectivity rho0, patch 1 black, scattering depth 1, emission off.

    Patch 0 absorbs (1 - rho0) Q of the sun and (1 - rho0)(1 - F) Qd of the sky under both schedules.  Patch 1
    sees the sky through 1 - F and the reflected sun rho0 Q of patch 0 through F.  The reference schedule
    (RadiationModel.cpp:4188-4189, 4286-4289: no copy at s = 0 when emission is off) gathers that reflected sun
    in the primary pass and again in the first scattering pass, and drops what patch 0 reflects of the sky; the
    conserving schedule gathers the reflected sun once and propagates rho0 (1 - F) Qd, of which patch 1 sees F."""
    R0 = (1 - rho0) * Q + (1 - rho0) * (1 - F) * Qd
    if schedule == 0:
        R1 = 2 * F * rho0 * Q + (1 - F) * Qd
    else:
        R1 = F * rho0 * Q + (1 - F) * Qd + F * rho0 * (1 - F) * Qd
    return R0, R1


def gray_parallel_rectangles_lw(eps, T, F):
    """Absorbed longwave flux of each of two facing gray one-sided rectangles at equal temperature T with
    emissivity eps and reflectivity 1 - eps (selfTest.cpp:731-807; the reference's closed form reduces to this
    for T0 = T1): each receives F J of the other's radiosity J = eps sigma T^4 / (1 - (1 - eps) F) and absorbs
    the fraction eps of it."""
    J = eps * SIGMA * T ** 4 / (1.0 - (1.0 - eps) * F)
    return eps * F * J
