// rt_device.h -- FP32 ray generation and intersection arithmetic shared by every gfx950 kernel.
//
// Determinism contract: the whole library is built with -ffp-contract=off and these helpers use
// correctly rounded division and square root, so a ray's origin, direction and intersection
// decisions are bit-identical to a plain IEEE-754 binary32 evaluation of the same expressions in
// the same order.  That is what lets the hit tables be compared bit for bit with a CPU evaluation.
//
// Reference semantics (Helios tree, R = plugins/radiation, CD = plugins/collisiondetection):
//   RNG            R/shaders/common/random.glsl:10-45
//   sample points  R/shaders/direct_raygen.comp:186-224, R/shaders/diffuse_raygen.comp:197-234
//   hemisphere     R/shaders/diffuse_raygen.comp:461-482, rotate_to_normal :303-327
//   predicates     R/shaders/common/intersections.glsl:26-112, CD/src/CollisionDetection_RayTracing.cpp:1793-1884
//   periodic wrap  R/shaders/common/periodic_boundary.glsl:28-112
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RT_HD __host__ __device__ __forceinline__
#define RT_MISS 0xFFFFFFFFu

struct v3f { float x, y, z; };

RT_HD v3f mk3(float x, float y, float z) { v3f r; r.x = x; r.y = y; r.z = z; return r; }
RT_HD v3f sub3(v3f a, v3f b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_HD v3f add3(v3f a, v3f b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_HD v3f mul3(v3f a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
RT_HD float dot3(v3f a, v3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RT_HD v3f cross3(v3f a, v3f b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// correctly rounded division / square root on both sides
RT_HD float rdiv(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __fdiv_rn(a, b);
#else
    return a / b;
#endif
}
// correctly rounded square root.  On gfx950 __builtin_sqrtf lowers to v_sqrt_f32 followed by a
// +-1 ulp residual correction (correct rounding); HIP's __fsqrt_rn lowers to the bare 1-ulp v_sqrt_f32.
RT_HD float rsqrt_cr(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_sqrtf(a);
#else
    return sqrtf(a);
#endif
}

// normalize by multiplying with 1/sqrt(|v|^2) (core/include/helios_vector_types.h:891-900)
RT_HD v3f nrm3(v3f a) {
    float m2 = a.x * a.x + a.y * a.y + a.z * a.z;
    float inv = rdiv(1.0f, rsqrt_cr(m2));
    return mul3(a, inv);
}

// row-major 4x4 affine point transform, terms summed left to right
RT_HD v3f xform3(const float *t, float px, float py, float pz) {
    return mk3(t[0] * px + t[1] * py + t[2] * pz + t[3],
               t[4] * px + t[5] * py + t[6] * pz + t[7],
               t[8] * px + t[9] * py + t[10] * pz + t[11]);
}

// ---- RNG: 16-round TEA seeding + 24-bit LCG ----
RT_HD uint32_t tea16(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

RT_HD float rnd24(uint32_t &st) {
    st = 1664525u * st + 1013904223u;
    return (float)(st & 0x00FFFFFFu) * (1.0f / 16777216.0f);  // exact: 24-bit integer times 2^-24
}

// sin/cos of 2*pi*r: quadrant reduction (exact) then Taylor polynomials on [-pi/4, pi/4]
RT_HD void sincos_2pi(float r, float &s, float &c) {
    float u = r * 4.0f;
    float q = floorf(u + 0.5f);
    float x = u - q;
    float a = x * 1.57079632679489662f;
    float a2 = a * a;
    float ps = a + a * a2 * (-0.166666666666666667f + a2 * (0.00833333333333333333f +
                                a2 * (-0.000198412698412698413f + a2 * 2.75573192239858907e-6f)));
    float pc = 1.0f + a2 * (-0.5f + a2 * (0.0416666666666666667f + a2 * (-0.00138888888888888889f +
                                a2 * (2.48015873015873016e-5f + a2 * -2.75573192239858907e-7f))));
    int qi = ((int)q) & 3;
    if (qi == 0) { s = ps; c = pc; }
    else if (qi == 1) { s = pc; c = -ps; }
    else if (qi == 2) { s = -ps; c = -pc; }
    else { s = -pc; c = ps; }
}

// ---- primitive record (64 B, leaf order) ----
//   r0 = v0.xyz, bits(prim_index | PATCH_BIT)
//   r1 = a|e1 .xyz, |a|^2
//   r2 = b|e2 .xyz, |b|^2
//   r3 = patch plane normal .xyz, 0
// A voxel (type 4) carries RT_VOXEL_BIT: r0 = its box's lower corner, r1 = the upper corner.
#define RT_PATCH_BIT 0x80000000u
// A primitive with a texture mask carries RT_MASK_BIT in its tag and, in r3.w, the float4 index (from the
// record base) of its mask descriptor; the index bits of a tag are RT_ID_BITS.
#define RT_MASK_BIT 0x40000000u
#define RT_VOXEL_BIT 0x20000000u
#define RT_ID_BITS 0x1FFFFFFFu

// castRays' voxel test (CD/src/CollisionDetection_RayTracing.cpp:1625-1680): slab method over the voxel's
// axis-aligned box [lo, hi] with the reference's arithmetic -- a direction component below 1e-8 in magnitude
// is parallel (a miss when the origin lies outside that slab), else t1, t2 = (lo - o) / d, (hi - o) / d;
// the entry and exit distances are the largest t1 and the smallest t2; the hit distance is the entry if
// it is at least 1e-6, else the exit (an origin inside the box); a hit needs exit >= 0 and distance >= 1e-6
// (the closest-hit test takes a voxel at exactly 1e-6, a surface only beyond it).  The radiation passes do
// not intersect voxels (R/shaders/common/intersections.glsl:128-133).
RT_HD bool voxel_hit_cd(const float4 lo, const float4 hi, v3f o, v3f d, float &t_out) {
    const float dc[3] = {d.x, d.y, d.z}, oc[3] = {o.x, o.y, o.z};
    const float lc[3] = {lo.x, lo.y, lo.z}, hc[3] = {hi.x, hi.y, hi.z};
    float tn = -3.402823466e+38f, tf = 3.402823466e+38f;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (fabsf(dc[a]) < 1e-8f) {
            ok = ok && !(oc[a] < lc[a] || oc[a] > hc[a]);
        } else {
            float t1 = rdiv(lc[a] - oc[a], dc[a]), t2 = rdiv(hc[a] - oc[a], dc[a]);
            if (t1 > t2) { const float x = t1; t1 = t2; t2 = x; }
            tn = fmaxf(tn, t1);
            tf = fminf(tf, t2);
            ok = ok && !(tn > tf);
        }
    }
    const float dist = tn >= 1e-6f ? tn : tf;
    t_out = dist;
    return ok && tf >= 0.0f && dist >= 1e-6f;
}

// Outward normal of the voxel face a castRays hit lies on (CollisionDetection_RayTracing.cpp:1685-1711):
// the axis along which the hit point is relatively farthest from the box centre.
RT_HD v3f voxel_normal_cd(const float4 lo, const float4 hi, v3f p) {
    const v3f c = mk3((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
    const v3f e = mk3((hi.x - lo.x) * 0.5f, (hi.y - lo.y) * 0.5f, (hi.z - lo.z) * 0.5f);
    const v3f l = sub3(p, c);
    const float rx = rdiv(fabsf(l.x), e.x), ry = rdiv(fabsf(l.y), e.y), rz = rdiv(fabsf(l.z), e.z);
    if (rx >= ry && rx >= rz) return mk3(l.x > 0.f ? 1.0f : -1.0f, 0.f, 0.f);
    if (ry >= rz) return mk3(0.f, l.y > 0.f ? 1.0f : -1.0f, 0.f);
    return mk3(0.f, 0.f, l.z > 0.f ? 1.0f : -1.0f);
}

// Radiation predicates (intersections.glsl) and CD predicates share the arithmetic; only the
// acceptance thresholds differ.  Written as predicates (every quantity computed, the tests combined at the
// end) rather than early returns: same floating-point operations and decisions, but no divergent control
// flow, which on a wave costs exec-mask bookkeeping per test.  t_out is written in every case and is
// meaningful only when the function returns true.
// ua, ub: the hit's d.a and d.b (patch) or barycentric u, v (triangle), for texture masks.
// Watertight ray / triangle test (Woop, Benthin and Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1),
// 2013): the vertices are moved into a frame where the ray runs along +z through the origin (the largest
// direction component becomes z, x and y swapped to keep the winding, then sheared), and the 2-D edge functions
// U, V, W of the sheared vertices decide containment.  They are computed the same way for an edge shared by two
// triangles with the same vertices, so a ray crossing that edge hits at least one of them; an edge function
// of exactly 0 is recomputed in double, and points on an edge count as inside.  Both faces are hit.  Returns
// the scaled distance T and determinant det (t = T / det) and the barycentrics of a and b.
RT_HD bool wt_tri(v3f o, v3f d, v3f a, v3f b, v3f c, float &t, float &ub, float &uc) {
    const float ax_ = fabsf(d.x), ay_ = fabsf(d.y), az_ = fabsf(d.z);
    int kz = ax_ > ay_ ? (ax_ > az_ ? 0 : 2) : (ay_ > az_ ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    const float dz = kz == 0 ? d.x : (kz == 1 ? d.y : d.z);
    if (dz < 0.0f) { const int tmp = kx; kx = ky; ky = tmp; }
    auto comp = [](v3f v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); };
    const float dx = comp(d, kx), dy = comp(d, ky);
    const float Sx = dx / dz, Sy = dy / dz, Sz = 1.0f / dz;
    const v3f A = sub3(a, o), B = sub3(b, o), C = sub3(c, o);
    const float Ax = comp(A, kx) - Sx * comp(A, kz), Ay = comp(A, ky) - Sy * comp(A, kz);
    const float Bx = comp(B, kx) - Sx * comp(B, kz), By = comp(B, ky) - Sy * comp(B, kz);
    const float Cx = comp(C, kx) - Sx * comp(C, kz), Cy = comp(C, ky) - Sy * comp(C, kz);
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float T = U * (Sz * comp(A, kz)) + V * (Sz * comp(B, kz)) + W * (Sz * comp(C, kz));
    const float inv = 1.0f / det;
    t = T * inv;
    ub = V * inv;
    uc = W * inv;
    return true;
}

template <bool CD, bool WT = false>
RT_HD bool isect_rec_uv(const float4 r0, const float4 r1, const float4 r2, const float4 r3, bool patch, v3f o, v3f d,
                        float &t_out, float &ua, float &ub) {
    v3f v0 = mk3(r0.x, r0.y, r0.z);
    v3f e1 = mk3(r1.x, r1.y, r1.z);
    v3f e2 = mk3(r2.x, r2.y, r2.z);
    if (WT) {
        // watertight predicates: a triangle (v0, v0 + e1, v0 + e2) and a patch as the triangles
        // (v0, v0 + a, v0 + a + b) and (v0, v0 + a + b, v0 + b), closed edges, both faces; the distance window
        // is the reference's ((1e-8, 1e8) for radiation rays, t > 1e-5 for castRays)
        float t = 0.f, u = 0.f, v = 0.f;
        bool hit;
        if (patch) {
            const v3f p1 = add3(v0, e1), p2 = add3(p1, e2), p3 = add3(v0, e2);
            hit = wt_tri(o, d, v0, p1, p2, t, u, v);
            if (!hit) hit = wt_tri(o, d, v0, p2, p3, t, u, v);
            const v3f dd = sub3(add3(o, mul3(d, t)), v0);
            u = dot3(dd, e1);
            v = dot3(dd, e2);
        } else {
            hit = wt_tri(o, d, v0, add3(v0, e1), add3(v0, e2), t, u, v);
        }
        hit = hit && (CD ? (t > 1e-5f) : (t > 1e-8f && t < 1e8f));
        t_out = t;
        ua = u;
        ub = v;
        return hit;
    }
    if (patch) {
        v3f n = mk3(r3.x, r3.y, r3.z);
        float denom = dot3(d, n);
        bool ok = CD ? (fabsf(denom) > 1e-8f) : !(fabsf(denom) < 1e-8f);
        float t = rdiv(dot3(sub3(v0, o), n), denom);
        ok = ok && (CD ? (t > 1e-8f && t < 1e8f) : !(t < 1e-8f || t > 1e8f));
        v3f p = add3(o, mul3(d, t));
        v3f dd = sub3(p, v0);
        float da = dot3(dd, e1), db = dot3(dd, e2);
        bool in;
        if (CD) in = (da >= -1e-8f && da <= r1.w + 1e-8f && db >= -1e-8f && db <= r2.w + 1e-8f);
        else in = (da > 0.0f && da < r1.w && db > 0.0f && db < r2.w);
        t_out = t;
        ua = da;
        ub = db;
        return ok && in;
    } else {
        const float EPS = CD ? 1e-5f : 1e-8f;
        v3f h = cross3(d, e2);
        float a = dot3(e1, h);
        bool ok = CD ? !(a > -EPS && a < EPS) : !(fabsf(a) < EPS);
        float f = rdiv(1.0f, a);
        v3f s = sub3(o, v0);
        float u = f * dot3(s, h);
        ok = ok && (CD ? !(u < -EPS || u > 1.0f + EPS) : !(u < 0.0f || u > 1.0f));
        v3f q = cross3(s, e1);
        float v = f * dot3(d, q);
        ok = ok && (CD ? !(v < -EPS || u + v > 1.0f + EPS) : !(v < 0.0f || u + v > 1.0f));
        float t = f * dot3(e2, q);
        ok = ok && (CD ? (t > EPS) : (t > 1e-8f && t < 1e8f));
        t_out = t;
        ua = u;
        ub = v;
        return ok;
    }
}

template <bool CD, bool WT = false>
RT_HD bool isect_rec(const float4 r0, const float4 r1, const float4 r2, const float4 r3, bool patch, v3f o, v3f d,
                     float &t_out) {
    float ua, ub;
    return isect_rec_uv<CD, WT>(r0, r1, r2, r3, patch, o, d, t_out, ua, ub);
}

// ---- texture masks (R/shaders/common/texture_mask.glsl:20-66) ----
// Descriptor of a masked primitive, three float4 at the record base + its index:
//   q0 = bits(texel byte offset from the record base, width, height, mode), mode 0 = patch with default
//        UVs, 1 = patch with custom UVs, 2 = triangle
//   q1 = uv0.xy, uv1.xy;  q2 = uv2.xy, uv3.xy
// Texels are one byte each, row-major from the top row; 0 = transparent.
RT_HD bool mask_texel(const float4 *base, float4 q0, int tx, int ty) {
    const uint32_t off = __float_as_uint(q0.x);
    const int sx = (int)__float_as_uint(q0.y), sy = (int)__float_as_uint(q0.z);
    tx = tx < 0 ? 0 : (tx > sx - 1 ? sx - 1 : tx);
    ty = ty < 0 ? 0 : (ty > sy - 1 ? sy - 1 : ty);
    const unsigned char *tex = (const unsigned char *)base + off;
    return tex[(size_t)ty * (size_t)sx + (size_t)tx] != 0;
}

// float -> int as GLSL int(): truncation, saturated (clamped to the texture right after)
RT_HD int mask_int(float x) { return x >= 2147483520.0f ? 2147483520 : (x > -2147483520.0f ? (int)x : -2147483520); }

// check_texture_mask for a patch at parametric (u, v) in [0,1]^2
RT_HD bool mask_patch_uv(const float4 *base, float4 q0, float4 q1, float4 q2, float u, float v) {
    const float fx = (float)((int)__float_as_uint(q0.y) - 1), fy = (float)((int)__float_as_uint(q0.z) - 1);
    int tx, ty;
    if (__float_as_uint(q0.w) == 0u) {
        tx = mask_int(floorf(fx * u));
        ty = mask_int(floorf(fy * (1.0f - v)));
    } else {
        const float du = q1.z - q1.x;   // uv1.x - uv0.x
        const float dv = q2.y - q1.w;   // uv2.y - uv1.y
        tx = mask_int(floorf(fx * (q1.x + u * du)));
        ty = mask_int(floorf(fy * ((1.0f - q1.y) - v * dv)));
    }
    return mask_texel(base, q0, tx, ty);
}

// check_texture_mask for a triangle at barycentric (u, v): interpolated UVs, flipped v, rounded
RT_HD bool mask_tri_uv(const float4 *base, float4 q0, float4 q1, float4 q2, float u, float v) {
    const float fx = (float)((int)__float_as_uint(q0.y) - 1), fy = (float)((int)__float_as_uint(q0.z) - 1);
    const float tu = (q1.x + u * (q1.z - q1.x)) + v * (q2.x - q1.x);
    const float tv = 1.0f - ((q1.y + u * (q1.w - q1.y)) + v * (q2.y - q1.y));
    return mask_texel(base, q0, mask_int(roundf(fx * fabsf(tu))), mask_int(roundf(fy * fabsf(tv))));
}

// Mask test of an accepted intersection: patch (ua, ub) = (d.a, d.b) -> u = d.a / |a|^2 with |a| the
// rounded length (intersections.glsl:64-66); triangle (ua, ub) = barycentric.
RT_HD bool mask_hit_opaque(const float4 *base, uint32_t didx, bool patch, float ua, float ub, float aa, float bb) {
    const float4 *q = base + didx;
    const float4 q0 = q[0], q1 = q[1], q2 = q[2];
    if (patch) {
        const float am = rsqrt_cr(aa), bm = rsqrt_cr(bb);
        return mask_patch_uv(base, q0, q1, q2, rdiv(ua, am * am), rdiv(ub, bm * bm));
    }
    return mask_tri_uv(base, q0, q1, q2, ua, ub);
}

// Direct-pass sample point in primitive-local coordinates (direct_raygen.comp:360-428)
RT_HD void direct_sample(uint32_t prim, uint32_t lx, uint32_t ly, uint32_t d, uint32_t seed, bool patch, float &spx,
                         float &spy) {
    uint32_t Nrays = d * d;
    uint32_t st = tea16(d * ly + lx + Nrays * prim, seed);
    float Rx = rnd24(st), Ry = rnd24(st);
    if (patch) {
        float fd = (float)d;
        spx = (-0.5f + rdiv((float)lx, fd)) + rdiv(Rx, fd);
        spy = (-0.5f + rdiv((float)ly, fd)) + rdiv(Ry, fd);
    } else {
        if (Rx < Ry) { spx = Rx; spy = Ry; } else { spx = Ry; spy = Rx; }
    }
}

// direct_sample() that also returns the RNG state after its two draws: the next draws belong to the
// non-collimated sources of the request, in source order (direct_raygen.comp:482-561)
RT_HD void direct_sample_st(uint32_t prim, uint32_t lx, uint32_t ly, uint32_t d, uint32_t seed, bool patch, float &spx,
                            float &spy, uint32_t &st) {
    uint32_t Nrays = d * d;
    st = tea16(d * ly + lx + Nrays * prim, seed);
    float Rx = rnd24(st), Ry = rnd24(st);
    if (patch) {
        float fd = (float)d;
        spx = (-0.5f + rdiv((float)lx, fd)) + rdiv(Rx, fd);
        spy = (-0.5f + rdiv((float)ly, fd)) + rdiv(Ry, fd);
    } else {
        if (Rx < Ry) { spx = Rx; spy = Ry; } else { spx = Ry; spy = Rx; }
    }
}

// ---- non-collimated sources (direct_raygen.comp:495-561) ----
// Source types: 0 collimated, 1 sphere, 2 sun sphere (as 1), 3 rectangle, 4 disk.  A source carries its
// position, width (sphere diameter | rectangle w, h | disk radius) and a rotation matrix m[9] (rows) built
// on the host from the ZYX Euler angles (transforms.glsl:49-64).  sin(acos(x)) is evaluated as
// sqrt(1 - x^2) and sin/cos of angles that are fractions of 2*pi with sincos_2pi(), so the FP32 rays are
// reproducible bit for bit; mathematically they are the reference's expressions.
#define RT_PI_F 3.14159265359f

// calc_sphere_strength (direct_raygen.comp:305-320): 10x10 quadrature over the sphere's visible hemisphere
RT_HD float sphere_strength(v3f n, v3f rd, float mag, float diam, uint32_t Nrays) {
    const float inv_n = rdiv(1.0f, (float)Nrays);
    const float ndr = fabsf(dot3(n, rd));
    const float mag2 = mag * mag;
    float s = 0.0f;
    for (int j = 0; j < 10; j++) {
        float sp, cp;
        sincos_2pi(rdiv((float)j + 0.5f, 10.0f), sp, cp);
        for (int i = 0; i < 10; i++) {
            float ct = 1.0f - rdiv(2.0f * ((float)i + 0.5f), 10.0f);
            float sth = rsqrt_cr(fmaxf(0.0f, 1.0f - ct * ct));
            v3f L = mk3(sth * cp, sth * sp, ct);
            float ld = dot3(L, rd);
            if (ld < 0.0f) s += rdiv(rdiv((inv_n * ndr) * fabsf(ld), mag2), 100.0f) * diam * diam;
        }
    }
    return s;
}

// One direct ray from origin o (primitive normal n) towards a non-collimated source: draws from st, sets the
// unit direction and the strength; false when the rectangle/disk back-face test rejects the ray.
RT_HD bool source_ray(uint32_t type, const float *pos, const float *width, const float *m, v3f o, v3f n,
                      uint32_t Nrays, uint32_t &st, v3f &dir, float &strength) {
    const v3f p = mk3(pos[0], pos[1], pos[2]);
    const float inv_n = rdiv(1.0f, (float)Nrays);
    if (type == 1u || type == 2u) {
        float u1 = rnd24(st), u2 = rnd24(st);
        float ct = 1.0f - 2.0f * u1;
        float sth = rsqrt_cr(fmaxf(0.0f, 1.0f - ct * ct));
        float sp, cp;
        sincos_2pi(u2, sp, cp);
        float radius = width[0] * 0.5f;
        v3f sph = mk3(radius * (sth * cp), radius * (sth * sp), radius * ct);
        v3f rd = sub3(add3(sph, p), o);
        float mag = rsqrt_cr(dot3(rd, rd));
        dir = nrm3(rd);
        strength = sphere_strength(n, dir, mag, width[0], Nrays);
        return true;
    }
    float lx, ly;
    float scale = 1.0f;
    if (type == 3u) {
        float Rx = rnd24(st), Ry = rnd24(st);
        lx = (-0.5f + Rx) * width[0];
        ly = (-0.5f + Ry) * width[1];
    } else {
        // concentric square-to-disk map (direct_raygen.comp:227-266); phi = k * pi/4 = 2*pi * (k/8)
        float Rx = rnd24(st), Ry = rnd24(st);
        float x = -1.0f + 2.0f * Rx, y = -1.0f + 2.0f * Ry;
        float r, k;
        if (x > -y) {
            if (x > y) { r = x; k = rdiv(y, x); }
            else { r = y; k = 2.0f - rdiv(x, y); }
        } else {
            if (x < y) { r = -x; k = 4.0f + rdiv(y, x); }
            else { r = -y; k = (y != 0.0f) ? 6.0f - rdiv(x, y) : 0.0f; }
        }
        float sk, ck;
        sincos_2pi(k * 0.125f, sk, ck);
        lx = r * ck;
        ly = r * sk;
        scale = width[0];
    }
    // light point = R * (lx, ly, 0); light normal = R * (0, 0, 1), normalised
    v3f lp = mk3(m[0] * lx + m[1] * ly, m[3] * lx + m[4] * ly, m[6] * lx + m[7] * ly);
    v3f ln = nrm3(mk3(m[2], m[5], m[8]));
    v3f rd = type == 3u ? sub3(add3(lp, p), o) : sub3(add3(mul3(lp, scale), p), o);
    if (dot3(rd, ln) > 0.0f) return false;
    float mag = rsqrt_cr(dot3(rd, rd));
    dir = nrm3(rd);
    float g = rdiv((inv_n * fabsf(dot3(n, dir))) * fabsf(dot3(ln, dir)), mag * mag);
    strength = type == 3u ? rdiv(g * (width[0] * width[1]), RT_PI_F) : (g * scale) * scale;
    return true;
}

// ---- cameras (camera_raygen.comp:204-310, pixel_label_raygen.comp:80-115) ----
// camera_rotate_point (transforms.glsl:66-89): Rz(phi) * Ry(theta) * p
RT_HD v3f camera_rotate(v3f q, float st, float ct, float sp, float cp) {
    return mk3((cp * ct) * q.x - sp * q.y + (cp * st) * q.z,
               (sp * ct) * q.x + cp * q.y + (sp * st) * q.z,
               -st * q.x + ct * q.z);
}

// Camera ray through pixel (ii, jj) of the full image at sub-pixel offset (rx, ry); (lx, ly) is the lens
// point in units of the lens radius (0, 0 for a pinhole or the pixel-label rays).
RT_HD void camera_ray(float px, float py, float pz, float st, float ct, float sp, float cp, float focal, float lens,
                      float aspect, float viewplane, uint32_t ii, uint32_t jj, uint32_t full_x, uint32_t full_y,
                      float rx, float ry, float lx, float ly, v3f &o, v3f &d) {
    const float mult = rdiv(1.0f, aspect);
    const float spy = -0.5f + rdiv((float)ii + rx, (float)full_x);
    const float spz = (0.5f - rdiv((float)jj + ry, (float)full_y)) * mult;
    const v3f p = mk3(focal, rdiv(spy, viewplane) * focal, rdiv(spz, viewplane) * focal);
    const v3f org = mk3(0.0f, (0.5f * lx) * lens, (0.5f * ly) * lens);
    const v3f dir = sub3(p, org);
    o = add3(camera_rotate(org, st, ct, sp, cp), mk3(px, py, pz));
    d = nrm3(camera_rotate(dir, st, ct, sp, cp));
}

// Concentric square-to-disk map of two uniform numbers (direct_raygen.comp:227-266): a point of the unit disk.
RT_HD void unit_disk(float Rx, float Ry, float &dx, float &dy) {
    float x = -1.0f + 2.0f * Rx, y = -1.0f + 2.0f * Ry;
    float r, k;
    if (x > -y) {
        if (x > y) { r = x; k = rdiv(y, x); }
        else { r = y; k = 2.0f - rdiv(x, y); }
    } else {
        if (x < y) { r = -x; k = 4.0f + rdiv(y, x); }
        else { r = -y; k = (y != 0.0f) ? 6.0f - rdiv(x, y) : 0.0f; }
    }
    float sk, ck;
    sincos_2pi(k * 0.125f, sk, ck);
    dx = r * ck;
    dy = r * sk;
}

// sphere_source_visible (camera_raygen.comp:131-143)
RT_HD bool cam_sphere_visible(v3f o, v3f d, v3f c, float radius, float tmax) {
    const v3f oc = sub3(o, c);
    const float b = dot3(oc, d);
    const float cc = dot3(oc, oc) - radius * radius;
    const float disc = b * b - cc;
    if (!(disc >= 0.0f)) return false;
    const float ts = -b - rsqrt_cr(disc);
    return ts > 0.0f && ts < tmax;
}

// rectangle_source_visible / disk_source_visible (camera_raygen.comp:146-203); the source normal is the
// third row of the rotation matrix (m[6..8]) as those shaders read it.  disk: width[0] = radius.
RT_HD bool cam_planar_visible(bool disk, v3f o, v3f d, v3f c, const float *width, const float *m, float tmax,
                              float &cos_angle) {
    const v3f nrm = mk3(m[6], m[7], m[8]);
    const float denom = dot3(d, nrm);
    if (denom >= -1e-6f) return false;
    const v3f oc = sub3(c, o);
    const float t = rdiv(dot3(oc, nrm), denom);
    if (t <= 0.0f || t >= tmax) return false;
    const v3f h = add3(o, mul3(d, t));
    const v3f off = sub3(h, c);
    bool in;
    if (disk) {
        in = dot3(off, off) <= width[0] * width[0];
    } else {
        const float lx = m[0] * off.x + m[3] * off.y + m[6] * off.z;
        const float ly = m[1] * off.x + m[4] * off.y + m[7] * off.z;
        in = fabsf(lx) <= width[0] * 0.5f && fabsf(ly) <= width[1] * 0.5f;
    }
    if (in) cos_angle = -denom;
    return in;
}

// Origin rejection against a texture mask (direct_raygen.comp:431-473, diffuse_raygen.comp:411-455),
// continuing the ray's RNG stream after its two sample draws.  Patch: while the sample's texel is
// transparent, up to 10 redraws of a uniform point of the whole patch (one subdivision); triangle: up to 50
// draws of a point of the canonical triangle, looked up with floor() and no abs().  Returns false when the
// ray is dropped: a triangle whose 50 draws all fell on transparent texels, or (diffuse rays only) a patch
// whose last draw did; direct patch rays keep their last draw.
RT_HD bool origin_mask(const float4 *base, uint32_t didx, bool patch, bool diffuse, uint32_t &st, float &spx,
                       float &spy) {
    const float4 *q = base + didx;
    const float4 q0 = q[0], q1 = q[1], q2 = q[2];
    if (patch) {
        bool ok = mask_patch_uv(base, q0, q1, q2, spx + 0.5f, spy + 0.5f);
        for (int k = 0; k < 10 && !ok; k++) {
            spx = -0.5f + rnd24(st);
            spy = -0.5f + rnd24(st);
            ok = mask_patch_uv(base, q0, q1, q2, spx + 0.5f, spy + 0.5f);
        }
        return ok || !diffuse;
    }
    const float fx = (float)((int)__float_as_uint(q0.y) - 1), fy = (float)((int)__float_as_uint(q0.z) - 1);
    bool masked = true;
    for (int a = 0; a < 50 && masked; a++) {
        float r1 = rnd24(st), r2 = rnd24(st);
        if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
        spx = r2;
        spy = r1 + r2;
        const float tu = (q1.x + r1 * (q1.z - q1.x)) + r2 * (q2.x - q1.x);
        const float tv = 1.0f - ((q1.y + r1 * (q1.w - q1.y)) + r2 * (q2.y - q1.y));
        masked = !mask_texel(base, q0, mask_int(floorf(fx * tu)), mask_int(floorf(fy * tv)));
    }
    return !masked;
}

// ---- sky radiance distributions (R/shaders/common/sky_distribution.glsl:40-116) ----
// Prague pattern (Perez-type circumsolar and horizon brightening) times its normalisation and pi
RT_HD float sky_prague(v3f d, float px, float py, float pz, const float *prm) {
    const float c = fminf(1.0f, fmaxf(-1.0f, d.x * px + d.y * py + d.z * pz));
    const float gamma = acosf(c) * (180.0f / RT_PI_F);
    const float cos_theta = fmaxf(0.0f, d.z);
    const float circ = 1.0f + prm[0] * expf(-gamma / prm[1]);
    const float horizon = 1.0f + (prm[2] - 1.0f) * (1.0f - cos_theta);
    return circ * horizon * prm[3] * RT_PI_F;
}

// evaluate_sky_distribution: p = {power-law K, power-law normalisation, peak direction xyz, Prague
// circumsolar strength, width (degrees), horizon brightness, normalisation}.  Power law when K > 0
// (angular distance from the peak clamped to 1 degree), else Prague when its normalisation is positive,
// else isotropic (1).
RT_HD float sky_weight(v3f d, const float *p) {
    if (p[0] > 0.0f) {
        const float c = fminf(1.0f, fmaxf(-1.0f, p[2] * d.x + p[3] * d.y + p[4] * d.z));
        const float psi = fmaxf(acosf(c), RT_PI_F / 180.0f);
        return powf(psi, -p[0]) * p[1];
    }
    if (p[8] > 0.0f) return sky_prague(d, p[2], p[3], p[4], p + 5);
    return 1.0f;
}

// Diffuse-pass sample point and local hemisphere direction (diffuse_raygen.comp:375-482)
RT_HD void diffuse_sample(uint32_t prim, uint32_t lx, uint32_t ly, uint32_t n, uint32_t seed, bool patch, float &spx,
                          float &spy, v3f &dl) {
    uint32_t st = tea16(n * ly + lx + n * n * prim, seed);
    float Rx = rnd24(st), Ry = rnd24(st);
    if (patch) { spx = -0.5f + Rx; spy = -0.5f + Ry; }
    else { if (Rx < Ry) { spx = Rx; spy = Ry; } else { spx = Ry; spy = Rx; } }
    float fn = (float)n;
    float Rt = rdiv((float)lx + rnd24(st), fn);
    float Rp = rdiv((float)ly + rnd24(st), fn);
    float sth = rsqrt_cr(Rt);
    float cth = rsqrt_cr(1.0f - Rt);
    float sp, cp;
    sincos_2pi(Rp, sp, cp);
    dl = mk3(sth * cp, sth * sp, cth);
}

// diffuse_sample() for a primitive that may carry a texture mask (descriptor didx, RT_MISS = none): the
// origin rejection draws come between the origin and direction draws.  Returns false when the ray is dropped.
RT_HD bool diffuse_sample_m(uint32_t prim, uint32_t lx, uint32_t ly, uint32_t n, uint32_t seed, bool patch,
                            const float4 *base, uint32_t didx, float &spx, float &spy, v3f &dl) {
    uint32_t st = tea16(n * ly + lx + n * n * prim, seed);
    float Rx = rnd24(st), Ry = rnd24(st);
    if (patch) { spx = -0.5f + Rx; spy = -0.5f + Ry; }
    else { if (Rx < Ry) { spx = Rx; spy = Ry; } else { spx = Ry; spy = Rx; } }
    if (didx != RT_MISS && !origin_mask(base, didx, patch, true, st, spx, spy)) return false;
    float fn = (float)n;
    float Rt = rdiv((float)lx + rnd24(st), fn);
    float Rp = rdiv((float)ly + rnd24(st), fn);
    float sth = rsqrt_cr(Rt);
    float cth = rsqrt_cr(1.0f - Rt);
    float sp, cp;
    sincos_2pi(Rp, sp, cp);
    dl = mk3(sth * cp, sth * sp, cth);
    return true;
}

// rotate_to_normal with the per-primitive frame (ce, se, ca, sa) = cos/sin of elevation and azimuth
RT_HD v3f frame_rotate(v3f dl, float ce, float se, float ca, float sa) {
    v3f dw;
    dw.x = dl.x * (ce * ca) - dl.y * sa + dl.z * (se * ca);
    dw.y = dl.x * (ce * sa) + dl.y * ca + dl.z * (se * sa);
    dw.z = -dl.x * se + dl.z * ce;
    return nrm3(dw);
}

// Neumaier-compensated FP32 accumulator
// Angular transmittance, reflectance and absorptance of a glass sheet (R/shaders/common/glass_cover.glsl:15-53):
// Snell refraction, Fresnel reflectance of each polarisation (sin^2(tr - t) / sin^2(tr + t) and
// tan^2(tr - t) / tan^2(tr + t), the normal-incidence limit below 1e-3 rad), Bouguer attenuation
// exp(-KL / cos tr) and the closed-form sums over the slab's internal reflections, averaged over the two
// polarisations.  tra = (tau, rho, alpha), tau + rho + alpha = 1.
RT_HD void glass_tra(float cos_theta, float n, float KL, float &tau, float &rho, float &alpha) {
    cos_theta = fmaxf(1e-4f, fminf(1.0f, cos_theta));
    const float theta = acosf(fmaxf(-1.0f, fminf(1.0f, cos_theta)));
    const float sin_tr = sinf(theta) / n;
    const float cos_tr = sqrtf(fmaxf(0.0f, 1.0f - sin_tr * sin_tr));
    const float theta_r = asinf(fmaxf(-1.0f, fminf(1.0f, sin_tr)));
    float r_par, r_per;
    if (theta < 1e-3f) {
        const float r0 = ((n - 1.0f) / (n + 1.0f)) * ((n - 1.0f) / (n + 1.0f));
        r_par = r0;
        r_per = r0;
    } else {
        const float sm = sinf(theta_r - theta), sp = sinf(theta_r + theta);
        const float tm = tanf(theta_r - theta), tp = tanf(theta_r + theta);
        r_per = (sm * sm) / fmaxf(1e-12f, sp * sp);
        r_par = (tm * tm) / fmaxf(1e-12f, tp * tp);
    }
    const float tau_a = KL > 0.0f ? expf(-KL / fmaxf(1e-4f, cos_tr)) : 1.0f;
    float t = 0.0f, r_sum = 0.0f;
    for (int pol = 0; pol < 2; pol++) {
        const float r = pol == 0 ? r_per : r_par;
        const float den = fmaxf(1e-6f, 1.0f - (r * tau_a) * (r * tau_a));
        const float ti = tau_a * (1.0f - r) * (1.0f - r) / den;
        t += 0.5f * ti;
        r_sum += 0.5f * (r * (1.0f + tau_a * ti));
    }
    tau = fmaxf(0.0f, fminf(1.0f, t));
    rho = fmaxf(0.0f, fminf(1.0f, r_sum));
    alpha = fmaxf(0.0f, 1.0f - tau - rho);
}

struct ksum {
    float s, c;
    RT_HD void init() { s = 0.f; c = 0.f; }
    RT_HD void add(float x) {
        float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    RT_HD void merge(float os, float oc) { add(os); c += oc; }
    RT_HD float value() const { return s + c; }
};
