/*
 * rt_oracle.c -- CPU restatement of the Helios radiation-transfer hot path (test infrastructure only;
 * the product never links it).  See rt_oracle.h for what pins each part.
 *
 * Arithmetic contract: every per-ray quantity (sample point, direction, intersection predicate) is FP32
 * with the operation order written out below and no contraction (built with -ffp-contract=off), so it
 * reproduces the HIP kernels' rays bit for bit.  Energy accumulation is FP64 ("FP64 host oracle").
 */
#include "rt_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MISS 0xFFFFFFFFu

/* ------------------------------------------------------------------------------------------------ */
/* RNG: R/shaders/common/random.glsl:10-45 (== R/include/RayTracing.cuh:237-262)                     */
/* ------------------------------------------------------------------------------------------------ */

/* Generic TEA encryption (Wheeler & Needham 1994); the Helios variant is 16 rounds with the key
 * below (random.glsl:10-23).  The general form exists so the published 32-round vectors pin it. */
uint32_t orc_tea(uint32_t v0, uint32_t v1, uint32_t rounds, const uint32_t key[4], uint32_t *v1_out) {
    uint32_t s0 = 0;
    for (uint32_t n = 0; n < rounds; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + key[0]) ^ (v1 + s0) ^ ((v1 >> 5) + key[1]);
        v1 += ((v0 << 4) + key[2]) ^ (v0 + s0) ^ ((v0 >> 5) + key[3]);
    }
    if (v1_out) *v1_out = v1;
    return v0;
}

static const uint32_t HELIOS_TEA_KEY[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};

uint32_t orc_tea16(uint32_t val0, uint32_t val1) { return orc_tea(val0, val1, 16, HELIOS_TEA_KEY, NULL); }

/* random.glsl:36-39: Numerical Recipes quick LCG */
uint32_t orc_lcg_next(uint32_t state) { return 1664525u * state + 1013904223u; }

/* random.glsl:42-44: 24-bit mantissa uniform in [0,1) (exact) */
float orc_rnd(uint32_t *state) {
    *state = orc_lcg_next(*state);
    return (float)(*state & 0x00FFFFFFu) / 16777216.0f;
}

/* sin and cos of 2*pi*r.  Quadrant reduction u = 4r (exact), q = nearest integer, x = u - q (exact),
 * a = x*pi/2 in [-pi/4, pi/4]; Taylor series to a^9 (sin) and a^10 (cos), whose truncation error at
 * pi/4 is below 2e-9.  The reference evaluates sin(2*pi*Rp)/cos(...) with the device libm; this
 * restatement replaces that with a function that is identical on every platform. */
void orc_sincos_2pi(float r, float *s, float *c) {
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
    switch (qi) {
        case 0: *s = ps; *c = pc; break;
        case 1: *s = pc; *c = -ps; break;
        case 2: *s = -ps; *c = -pc; break;
        default: *s = -pc; *c = ps; break;
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* small FP32 vector helpers (left-to-right evaluation, no contraction)                              */
/* ------------------------------------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
/* core/include/helios_vector_types.h:891-900 (normalize multiplies by 1/sqrt(m2)) */
static inline v3 nrm(v3 a) {
    float m2 = a.x * a.x + a.y * a.y + a.z * a.z;
    float inv = 1.0f / sqrtf(m2);
    return scl(a, inv);
}
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }

/* Row-major affine transform of a point (R/src/backends/VulkanComputeBackend.cpp:341-352) */
static inline v3 xform(const float *t, float px, float py, float pz) {
    return mk(t[0] * px + t[1] * py + t[2] * pz + t[3],
              t[4] * px + t[5] * py + t[6] * pz + t[7],
              t[8] * px + t[9] * py + t[10] * pz + t[11]);
}

/* ------------------------------------------------------------------------------------------------ */
/* scene                                                                                             */
/* ------------------------------------------------------------------------------------------------ */
typedef struct {
    float lo[3], hi[3];
    uint32_t left, right; /* right == MISS => leaf */
    uint32_t start, count;
} onode;

struct orc_scene {
    uint32_t P;
    float *T;          /* [P][16] */
    uint32_t *type;    /* [P] */
    uint32_t *twos;    /* [P] */
    uint32_t *uuid;    /* [P] */
    v3 *verts;         /* [P][4] world vertices */
    v3 *normal;        /* [P] precomputed world normal */
    float *frame;      /* [P][4] ce, se, ca, sa */
    v3 *pn;            /* [P] patch plane normal (intersections.glsl:33) */
    float *aa, *bb;    /* [P] |a|^2, |b|^2 */
    float *blo, *bhi;  /* [P][3] inflated AABBs */
    /* periodic */
    uint32_t nbbox;
    v3 bbox[4][4];
    int px, py;
    float dom[4];
    /* SAH BVH */
    onode *nodes;
    uint32_t nnodes, cap;
    uint32_t *pidx;
    /* texture masks (R/shaders/common/texture_mask.glsl) */
    uint32_t nmasks;
    uint8_t *tex;      /* texels, one byte each, masks concatenated */
    uint32_t *msize;   /* [M][2] width, height */
    uint64_t *moff;    /* [M] first texel */
    int32_t *mid;      /* [P] mask of each primitive, -1 = none (NULL: no masks) */
    float *muv;        /* [P][8] vertex UVs */
    int32_t *muvid;    /* [P] custom-UV flag of a patch (>= 0) */
};

static int is_patch(uint32_t t) { return t == 0 || t == 3; }

/* World vertices: canonical quad (VulkanComputeBackend.cpp:328) / triangle (:329) */
static void prep_prim(orc_scene *s, uint32_t i) {
    const float *t = s->T + (size_t)i * 16;
    v3 *v = s->verts + (size_t)i * 4;
    uint32_t ty = s->type[i];
    if (is_patch(ty)) {
        v[0] = xform(t, -0.5f, -0.5f, 0.f);
        v[1] = xform(t, 0.5f, -0.5f, 0.f);
        v[2] = xform(t, 0.5f, 0.5f, 0.f);
        v[3] = xform(t, -0.5f, 0.5f, 0.f);
    } else if (ty == 1) {
        v[0] = xform(t, 0.f, 0.f, 0.f);
        v[1] = xform(t, 0.f, 1.f, 0.f);
        v[2] = xform(t, 1.f, 1.f, 0.f);
        v[3] = v[2];
    } else if (ty == 4) {
        /* a voxel's box: running min/max over the Context's eight corners in their order
         * (core/src/Context_primitive.cpp:917-929), as castRays forms it
         * (CD/src/CollisionDetection_RayTracing.cpp:1627-1637); v[0] = lower, v[1] = upper corner */
        static const float Y[8][3] = {{-0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, -0.5f}, {0.5f, 0.5f, -0.5f},
                                      {-0.5f, 0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f},
                                      {0.5f, 0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f}};
        v3 lo = xform(t, Y[0][0], Y[0][1], Y[0][2]), hi = lo;
        for (int k = 1; k < 8; k++) {
            v3 c = xform(t, Y[k][0], Y[k][1], Y[k][2]);
            if (c.x < lo.x) lo.x = c.x; if (c.y < lo.y) lo.y = c.y; if (c.z < lo.z) lo.z = c.z;
            if (c.x > hi.x) hi.x = c.x; if (c.y > hi.y) hi.y = c.y; if (c.z > hi.z) hi.z = c.z;
        }
        v[0] = lo; v[1] = hi; v[2] = v[3] = lo;
    } else {
        v[0] = v[1] = v[2] = v[3] = xform(t, 0.f, 0.f, 0.f);
    }
    /* Precomputed world normal (VulkanComputeBackend.cpp:377-435) */
    v3 w0, w1, w2, n;
    if (is_patch(ty) || ty == 2) {
        w0 = xform(t, 0.f, 0.f, 0.f); w1 = xform(t, 1.f, 0.f, 0.f); w2 = xform(t, 0.f, 1.f, 0.f);
    } else {
        w0 = xform(t, 0.f, 0.f, 0.f); w1 = xform(t, 0.f, 1.f, 0.f); w2 = xform(t, 1.f, 1.f, 0.f);
    }
    if (ty == 0 || ty == 1 || ty == 2 || ty == 3) {
        n = cross(sub(w1, w0), sub(w2, w0));
        float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
        if (len > 1e-8f) { float inv = 1.0f / len; n = scl(n, inv); }
        else n = mk(0.f, 0.f, 1.f);
    } else {
        n = mk(0.f, 0.f, 1.f);
    }
    s->normal[i] = n;
    /* diffuse frame: cos/sin of elevation acos(n.z) and azimuth atan2(n.y, n.x)
     * (diffuse_raygen.comp:303-327), evaluated with exact identities. */
    float ce = n.z < -1.f ? -1.f : (n.z > 1.f ? 1.f : n.z);
    float se = sqrtf((1.0f - ce) * (1.0f + ce));
    float r = sqrtf(n.x * n.x + n.y * n.y);
    float ca, sa;
    if (r > 0.f) { ca = n.x / r; sa = n.y / r; }
    else { ca = signbit(n.x) ? -1.f : 1.f; sa = 0.f; }
    s->frame[i * 4 + 0] = ce; s->frame[i * 4 + 1] = se; s->frame[i * 4 + 2] = ca; s->frame[i * 4 + 3] = sa;
    /* patch predicate data (intersections.glsl:31-53) */
    if (is_patch(ty)) {
        s->pn[i] = nrm(cross(sub(v[1], v[0]), sub(v[2], v[0])));
        v3 a = sub(v[1], v[0]), b = sub(v[3], v[0]);
        s->aa[i] = dot(a, a);
        s->bb[i] = dot(b, b);
    } else {
        s->pn[i] = mk(0, 0, 1); s->aa[i] = 0; s->bb[i] = 0;
    }
    /* conservative AABB: exact vertex bounds widened by 2^-18 of the extent plus 2^-20 of the
     * magnitude, so a slab test can never reject a ray that a predicate (incl. the 1e-5 barycentric
     * slack of the CD triangle test) would accept. */
    int nv = is_patch(ty) ? 4 : (ty == 4 ? 2 : 3);
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int k = 0; k < nv; k++) {
        float c[3] = {v[k].x, v[k].y, v[k].z};
        for (int a = 0; a < 3; a++) { if (c[a] < lo[a]) lo[a] = c[a]; if (c[a] > hi[a]) hi[a] = c[a]; }
    }
    float ext = 0.f;
    for (int a = 0; a < 3; a++) if (hi[a] - lo[a] > ext) ext = hi[a] - lo[a];
    for (int a = 0; a < 3; a++) {
        float m = fmaxf(fabsf(lo[a]), fabsf(hi[a]));
        float pad = ext * 3.0e-5f + m * 1.0e-6f + 1e-30f;
        s->blo[i * 3 + a] = lo[a] - pad;
        s->bhi[i * 3 + a] = hi[a] + pad;
    }
}

/* ---- binned SAH BVH2 (CD/src/CollisionDetection.cpp:940-1216: 16 bins, <=8 prims/leaf, depth<=64,
 *      median fallback).  The tree shape is not a parity target; it only has to be exact. ---- */
static uint32_t new_node(orc_scene *s) {
    if (s->nnodes == s->cap) {
        s->cap = s->cap ? s->cap * 2 : 1024;
        s->nodes = (onode *)realloc(s->nodes, sizeof(onode) * s->cap);
    }
    return s->nnodes++;
}

static void centroid(const orc_scene *s, uint32_t p, float c[3]) {
    for (int a = 0; a < 3; a++) c[a] = (s->blo[p * 3 + a] + s->bhi[p * 3 + a]) * 0.5f;
}

static float sarea(const float lo[3], const float hi[3]) {
    float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
    if (d0 < 0 || d1 < 0 || d2 < 0) return 0.f;
    return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
}

static int cmp_axis;
static const orc_scene *cmp_scene;
static int cmp_centroid(const void *A, const void *B) {
    uint32_t a = *(const uint32_t *)A, b = *(const uint32_t *)B;
    float ca[3], cb[3];
    centroid(cmp_scene, a, ca); centroid(cmp_scene, b, cb);
    if (ca[cmp_axis] < cb[cmp_axis]) return -1;
    if (ca[cmp_axis] > cb[cmp_axis]) return 1;
    return (a < b) ? -1 : (a > b);
}

static void build_rec(orc_scene *s, uint32_t ni, uint32_t start, uint32_t count, int depth) {
    onode nd;
    for (int a = 0; a < 3; a++) { nd.lo[a] = FLT_MAX; nd.hi[a] = -FLT_MAX; }
    for (uint32_t k = 0; k < count; k++) {
        uint32_t p = s->pidx[start + k];
        for (int a = 0; a < 3; a++) {
            if (s->blo[p * 3 + a] < nd.lo[a]) nd.lo[a] = s->blo[p * 3 + a];
            if (s->bhi[p * 3 + a] > nd.hi[a]) nd.hi[a] = s->bhi[p * 3 + a];
        }
    }
    nd.start = start; nd.count = count; nd.left = MISS; nd.right = MISS;
    if (count <= 8 || depth >= 64) { s->nodes[ni] = nd; return; }

    float cmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t k = 0; k < count; k++) {
        float c[3]; centroid(s, s->pidx[start + k], c);
        for (int a = 0; a < 3; a++) { if (c[a] < cmin[a]) cmin[a] = c[a]; if (c[a] > cmax[a]) cmax[a] = c[a]; }
    }
    enum { NB = 16 };
    int best_axis = -1, best_b = -1;
    float best_cost = FLT_MAX;
    for (int a = 0; a < 3; a++) {
        float ext = cmax[a] - cmin[a];
        if (ext <= 0.f) continue;
        float scale = (float)NB / ext;
        uint32_t cnt[NB] = {0};
        float blo[NB][3], bhi[NB][3];
        for (int b = 0; b < NB; b++) for (int q = 0; q < 3; q++) { blo[b][q] = FLT_MAX; bhi[b][q] = -FLT_MAX; }
        for (uint32_t k = 0; k < count; k++) {
            uint32_t p = s->pidx[start + k];
            float c[3]; centroid(s, p, c);
            int b = (int)((c[a] - cmin[a]) * scale);
            if (b >= NB) b = NB - 1;
            if (b < 0) b = 0;
            cnt[b]++;
            for (int q = 0; q < 3; q++) {
                if (s->blo[p * 3 + q] < blo[b][q]) blo[b][q] = s->blo[p * 3 + q];
                if (s->bhi[p * 3 + q] > bhi[b][q]) bhi[b][q] = s->bhi[p * 3 + q];
            }
        }
        for (int split = 1; split < NB; split++) {
            float llo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            float rlo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, rhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            uint32_t nl = 0, nr = 0;
            for (int b = 0; b < NB; b++) {
                if (!cnt[b]) continue;
                float *lo = b < split ? llo : rlo, *hi = b < split ? lhi : rhi;
                for (int q = 0; q < 3; q++) { if (blo[b][q] < lo[q]) lo[q] = blo[b][q]; if (bhi[b][q] > hi[q]) hi[q] = bhi[b][q]; }
                if (b < split) nl += cnt[b]; else nr += cnt[b];
            }
            if (!nl || !nr) continue;
            float cost = sarea(llo, lhi) * (float)nl + sarea(rlo, rhi) * (float)nr;
            if (cost < best_cost) { best_cost = cost; best_axis = a; best_b = split; }
        }
    }
    uint32_t mid;
    if (best_axis < 0) {
        /* median fallback on the longest node axis */
        int ax = 0;
        float e0 = nd.hi[0] - nd.lo[0], e1 = nd.hi[1] - nd.lo[1], e2 = nd.hi[2] - nd.lo[2];
        if (e1 > e0 && e1 >= e2) ax = 1; else if (e2 > e0 && e2 > e1) ax = 2;
        cmp_axis = ax; cmp_scene = s;
        qsort(s->pidx + start, count, sizeof(uint32_t), cmp_centroid);
        mid = count / 2;
    } else {
        float scale = (float)NB / (cmax[best_axis] - cmin[best_axis]);
        uint32_t i = start, j = start + count;
        while (i < j) {
            float c[3]; centroid(s, s->pidx[i], c);
            int b = (int)((c[best_axis] - cmin[best_axis]) * scale);
            if (b >= NB) b = NB - 1;
            if (b < 0) b = 0;
            if (b < best_b) i++;
            else { uint32_t t = s->pidx[i]; s->pidx[i] = s->pidx[--j]; s->pidx[j] = t; }
        }
        mid = i - start;
        if (mid == 0 || mid == count) mid = count / 2;
    }
    uint32_t l = new_node(s), r = new_node(s);
    nd.left = l; nd.right = r;
    s->nodes[ni] = nd;
    build_rec(s, l, start, mid, depth + 1);
    build_rec(s, r, start + mid, count - mid, depth + 1);
}

/* Periodic boundary faces (R/src/RadiationModel.cpp:6023-6105) and domain bounds
 * (R/src/backends/VulkanComputeBackend.cpp:544-582). */
static void build_periodic(orc_scene *s) {
    s->nbbox = 0;
    s->dom[0] = s->dom[1] = s->dom[2] = s->dom[3] = 0.f;
    if (!s->px && !s->py) return;
    float xb[2] = {FLT_MAX, -FLT_MAX}, yb[2] = {FLT_MAX, -FLT_MAX}, zb[2] = {FLT_MAX, -FLT_MAX};
    for (uint32_t i = 0; i < s->P; i++) {
        /* Context::getDomainBoundingBox: every vertex, a voxel's eight corners (its box) included */
        int nv = is_patch(s->type[i]) ? 4 : (s->type[i] == 4 ? 2 : 3);
        for (int k = 0; k < nv; k++) {
            v3 v = s->verts[(size_t)i * 4 + k];
            if (v.x < xb[0]) xb[0] = v.x; if (v.x > xb[1]) xb[1] = v.x;
            if (v.y < yb[0]) yb[0] = v.y; if (v.y > yb[1]) yb[1] = v.y;
            if (v.z < zb[0]) zb[0] = v.z; if (v.z > zb[1]) zb[1] = v.z;
        }
    }
    xb[0] = (float)((double)xb[0] - 1e-5); xb[1] = (float)((double)xb[1] + 1e-5);
    yb[0] = (float)((double)yb[0] - 1e-5); yb[1] = (float)((double)yb[1] + 1e-5);
    zb[0] = (float)((double)zb[0] - 1e-5); zb[1] = (float)((double)zb[1] + 1e-5);
    uint32_t k = 0;
    if (s->px) {
        s->bbox[k][0] = mk(xb[0], yb[0], zb[0]); s->bbox[k][1] = mk(xb[0], yb[1], zb[0]);
        s->bbox[k][2] = mk(xb[0], yb[1], zb[1]); s->bbox[k][3] = mk(xb[0], yb[0], zb[1]); k++;
        s->bbox[k][0] = mk(xb[1], yb[0], zb[0]); s->bbox[k][1] = mk(xb[1], yb[1], zb[0]);
        s->bbox[k][2] = mk(xb[1], yb[1], zb[1]); s->bbox[k][3] = mk(xb[1], yb[0], zb[1]); k++;
    }
    if (s->py) {
        s->bbox[k][0] = mk(xb[0], yb[0], zb[0]); s->bbox[k][1] = mk(xb[1], yb[0], zb[0]);
        s->bbox[k][2] = mk(xb[1], yb[0], zb[1]); s->bbox[k][3] = mk(xb[0], yb[0], zb[1]); k++;
        s->bbox[k][0] = mk(xb[0], yb[1], zb[0]); s->bbox[k][1] = mk(xb[1], yb[1], zb[0]);
        s->bbox[k][2] = mk(xb[1], yb[1], zb[1]); s->bbox[k][3] = mk(xb[0], yb[1], zb[1]); k++;
    }
    s->nbbox = k;
    float x0 = 1e30f, x1 = -1e30f, y0 = 1e30f, y1 = -1e30f;
    for (uint32_t f = 0; f < k; f++)
        for (int q = 0; q < 4; q++) {
            v3 v = s->bbox[f][q];
            if (v.x < x0) x0 = v.x; if (v.x > x1) x1 = v.x;
            if (v.y < y0) y0 = v.y; if (v.y > y1) y1 = v.y;
        }
    s->dom[0] = x0; s->dom[1] = x1; s->dom[2] = y0; s->dom[3] = y1;
}

orc_scene *orc_scene_create(uint32_t P, const float *transforms, const uint32_t *types, const uint32_t *twosided,
                            const uint32_t *uuids, int periodic_x, int periodic_y) {
    orc_scene *s = (orc_scene *)calloc(1, sizeof(orc_scene));
    s->P = P;
    s->T = (float *)malloc(sizeof(float) * 16 * (P ? P : 1));
    memcpy(s->T, transforms, sizeof(float) * 16 * P);
    s->type = (uint32_t *)malloc(sizeof(uint32_t) * (P ? P : 1));
    memcpy(s->type, types, sizeof(uint32_t) * P);
    s->twos = (uint32_t *)malloc(sizeof(uint32_t) * (P ? P : 1));
    memcpy(s->twos, twosided, sizeof(uint32_t) * P);
    s->uuid = (uint32_t *)malloc(sizeof(uint32_t) * (P ? P : 1));
    for (uint32_t i = 0; i < P; i++) s->uuid[i] = uuids ? uuids[i] : i;
    s->verts = (v3 *)malloc(sizeof(v3) * 4 * (P ? P : 1));
    s->normal = (v3 *)malloc(sizeof(v3) * (P ? P : 1));
    s->frame = (float *)malloc(sizeof(float) * 4 * (P ? P : 1));
    s->pn = (v3 *)malloc(sizeof(v3) * (P ? P : 1));
    s->aa = (float *)malloc(sizeof(float) * (P ? P : 1));
    s->bb = (float *)malloc(sizeof(float) * (P ? P : 1));
    s->blo = (float *)malloc(sizeof(float) * 3 * (P ? P : 1));
    s->bhi = (float *)malloc(sizeof(float) * 3 * (P ? P : 1));
    s->px = periodic_x; s->py = periodic_y;
    for (uint32_t i = 0; i < P; i++) prep_prim(s, i);
    build_periodic(s);
    s->pidx = (uint32_t *)malloc(sizeof(uint32_t) * (P ? P : 1));
    uint32_t np = 0;
    for (uint32_t i = 0; i < P; i++)
        if (s->type[i] <= 1 || s->type[i] == 3 || s->type[i] == 4) s->pidx[np++] = i;
    if (np) {
        uint32_t root = new_node(s);
        build_rec(s, root, 0, np, 0);
    }
    return s;
}

void orc_scene_destroy(orc_scene *s) {
    if (!s) return;
    free(s->T); free(s->type); free(s->twos); free(s->uuid); free(s->verts); free(s->normal); free(s->frame);
    free(s->pn); free(s->aa); free(s->bb); free(s->blo); free(s->bhi); free(s->nodes); free(s->pidx);
    free(s->tex); free(s->msize); free(s->moff); free(s->mid); free(s->muv); free(s->muvid);
    free(s);
}

void orc_scene_geometry(const orc_scene *s, float *verts, float *normals, float *frame, float *domain, uint32_t *nbbox) {
    for (uint32_t i = 0; i < s->P; i++) {
        if (verts) for (int k = 0; k < 4; k++) {
            verts[i * 12 + k * 3 + 0] = s->verts[i * 4 + k].x;
            verts[i * 12 + k * 3 + 1] = s->verts[i * 4 + k].y;
            verts[i * 12 + k * 3 + 2] = s->verts[i * 4 + k].z;
        }
        if (normals) { normals[i * 3] = s->normal[i].x; normals[i * 3 + 1] = s->normal[i].y; normals[i * 3 + 2] = s->normal[i].z; }
        if (frame) for (int k = 0; k < 4; k++) frame[i * 4 + k] = s->frame[i * 4 + k];
    }
    if (domain) for (int k = 0; k < 4; k++) domain[k] = s->dom[k];
    if (nbbox) *nbbox = s->nbbox;
}

void orc_areas(const orc_scene *s, double *out) {
    for (uint32_t i = 0; i < s->P; i++) {
        const v3 *v = s->verts + (size_t)i * 4;
        double e1[3] = {(double)v[1].x - v[0].x, (double)v[1].y - v[0].y, (double)v[1].z - v[0].z};
        double e2[3], c[3];
        if (is_patch(s->type[i])) { e2[0] = (double)v[3].x - v[0].x; e2[1] = (double)v[3].y - v[0].y; e2[2] = (double)v[3].z - v[0].z; }
        else { e2[0] = (double)v[2].x - v[0].x; e2[1] = (double)v[2].y - v[0].y; e2[2] = (double)v[2].z - v[0].z; }
        c[0] = e1[1] * e2[2] - e1[2] * e2[1]; c[1] = e1[2] * e2[0] - e1[0] * e2[2]; c[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double a = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        out[i] = is_patch(s->type[i]) ? a : (s->type[i] == 1 ? 0.5 * a : 0.0);
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* intersection predicates                                                                           */
/* ------------------------------------------------------------------------------------------------ */

/* ---- watertight predicates (orc_set_predicates) ---- */
static int g_watertight = 0;
void orc_set_predicates(int mode) { g_watertight = mode != 0; }

static float comp3(v3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

/* Woop, Benthin and Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013: shear the vertices into
 * the ray's frame (largest direction component as z, x and y swapped to keep the winding), decide with the 2-D
 * edge functions (recomputed in double when one is exactly 0; on an edge counts as inside), both faces. */
static int wt_tri_orc(v3 o, v3 d, v3 a, v3 b, v3 c, float *t, float *ub, float *uc) {
    float ax_ = fabsf(d.x), ay_ = fabsf(d.y), az_ = fabsf(d.z);
    int kz = ax_ > ay_ ? (ax_ > az_ ? 0 : 2) : (ay_ > az_ ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    float dz = comp3(d, kz);
    if (dz < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
    float Sx = comp3(d, kx) / dz, Sy = comp3(d, ky) / dz, Sz = 1.0f / dz;
    v3 A = sub(a, o), B = sub(b, o), C = sub(c, o);
    float Ax = comp3(A, kx) - Sx * comp3(A, kz), Ay = comp3(A, ky) - Sy * comp3(A, kz);
    float Bx = comp3(B, kx) - Sx * comp3(B, kz), By = comp3(B, ky) - Sy * comp3(B, kz);
    float Cx = comp3(C, kx) - Sx * comp3(C, kz), Cy = comp3(C, ky) - Sy * comp3(C, kz);
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    float det = U + V + W;
    if (det == 0.0f) return 0;
    float T = U * (Sz * comp3(A, kz)) + V * (Sz * comp3(B, kz)) + W * (Sz * comp3(C, kz));
    float inv = 1.0f / det;
    *t = T * inv;
    *ub = V * inv;
    *uc = W * inv;
    return 1;
}

/* the record's vertices as the GPU rebuilds them: v0, v0 + e1, v0 + e2 (triangle); v0, v0 + a, v0 + a + b,
 * v0 + b (patch, two triangles); the distance window of the mode (castRays: t > 1e-5; radiation: (1e-8, 1e8)) */
static int isect_wt_orc(const orc_scene *s, uint32_t i, v3 o, v3 d, int cd, float *t_out, float *ua, float *ub) {
    v3 v0 = s->verts[(size_t)i * 4];
    v3 e1 = sub(s->verts[i * 4 + 1], v0);
    int patch = is_patch(s->type[i]);
    v3 e2 = sub(s->verts[i * 4 + (patch ? 3 : 2)], v0);
    float t = 0.f, u = 0.f, v = 0.f;
    int hit;
    if (patch) {
        v3 p1 = add(v0, e1), p2 = add(p1, e2), p3 = add(v0, e2);
        hit = wt_tri_orc(o, d, v0, p1, p2, &t, &u, &v);
        if (!hit) hit = wt_tri_orc(o, d, v0, p2, p3, &t, &u, &v);
        v3 dd = sub(add(o, scl(d, t)), v0);
        u = dot(dd, e1);
        v = dot(dd, e2);
    } else {
        hit = wt_tri_orc(o, d, v0, add(v0, e1), add(v0, e2), &t, &u, &v);
    }
    hit = hit && (cd ? (t > 1e-5f) : (t > 1e-8f && t < 1e8f));
    if (hit) { *t_out = t; *ua = u; *ub = v; }
    return hit;
}

/* R/shaders/common/intersections.glsl:26-70 */
static int isect_patch_rad_uv(const orc_scene *s, uint32_t i, v3 o, v3 d, float *t_out, float *ua, float *ub) {
    if (g_watertight) return isect_wt_orc(s, i, o, d, 0, t_out, ua, ub);
    v3 n = s->pn[i];
    float denom = dot(d, n);
    if (fabsf(denom) < 1e-8f) return 0;
    v3 anchor = s->verts[(size_t)i * 4];
    float t = dot(sub(anchor, o), n) / denom;
    if (t < 1e-8f || t > 1e8f) return 0;
    v3 p = add(o, scl(d, t));
    v3 dd = sub(p, anchor);
    v3 a = sub(s->verts[i * 4 + 1], anchor), b = sub(s->verts[i * 4 + 3], anchor);
    float da = dot(dd, a), db = dot(dd, b);
    if (da > 0.0f && da < s->aa[i] && db > 0.0f && db < s->bb[i]) { *t_out = t; *ua = da; *ub = db; return 1; }
    return 0;
}

/* R/shaders/common/intersections.glsl:73-112 (Moller-Trumbore) */
static int isect_tri_rad_uv(const orc_scene *s, uint32_t i, v3 o, v3 d, float *t_out, float *ua, float *ub) {
    if (g_watertight) return isect_wt_orc(s, i, o, d, 0, t_out, ua, ub);
    v3 v0 = s->verts[(size_t)i * 4];
    v3 e1 = sub(s->verts[i * 4 + 1], v0), e2 = sub(s->verts[i * 4 + 2], v0);
    v3 h = cross(d, e2);
    float a = dot(e1, h);
    if (fabsf(a) < 1e-8f) return 0;
    float f = 1.0f / a;
    v3 sv = sub(o, v0);
    float u = f * dot(sv, h);
    if (u < 0.0f || u > 1.0f) return 0;
    v3 q = cross(sv, e1);
    float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return 0;
    float t = f * dot(e2, q);
    if (t > 1e-8f && t < 1e8f) { *t_out = t; *ua = u; *ub = v; return 1; }
    return 0;
}

/* ---- texture masks (R/shaders/common/texture_mask.glsl:20-66) ---- */
static int tex_int(float x) { return x >= 2147483520.0f ? 2147483520 : (x > -2147483520.0f ? (int)x : -2147483520); }

static int mask_texel_orc(const orc_scene *s, int m, int tx, int ty) {
    int sx = (int)s->msize[2 * m], sy = (int)s->msize[2 * m + 1];
    if (tx < 0) tx = 0;
    if (tx > sx - 1) tx = sx - 1;
    if (ty < 0) ty = 0;
    if (ty > sy - 1) ty = sy - 1;
    return s->tex[s->moff[m] + (uint64_t)ty * (uint64_t)sx + (uint64_t)tx] != 0;
}

/* check_texture_mask: patch at parametric (u, v), triangle at barycentric (u, v) */
static int mask_opaque_orc(const orc_scene *s, uint32_t i, float u, float v) {
    int m = s->mid[i];
    float fx = (float)((int)s->msize[2 * m] - 1), fy = (float)((int)s->msize[2 * m + 1] - 1);
    const float *uv = s->muv + (size_t)i * 8;
    int tx, ty;
    if (s->type[i] == 1) {
        float tu = (uv[0] + u * (uv[2] - uv[0])) + v * (uv[4] - uv[0]);
        float tv = 1.0f - ((uv[1] + u * (uv[3] - uv[1])) + v * (uv[5] - uv[1]));
        tx = tex_int(roundf(fx * fabsf(tu)));
        ty = tex_int(roundf(fy * fabsf(tv)));
    } else if (s->muvid[i] < 0) {
        tx = tex_int(floorf(fx * u));
        ty = tex_int(floorf(fy * (1.0f - v)));
    } else {
        float du = uv[2] - uv[0], dv = uv[5] - uv[3];
        tx = tex_int(floorf(fx * (uv[0] + u * du)));
        ty = tex_int(floorf(fy * ((1.0f - uv[1]) - v * dv)));
    }
    return mask_texel_orc(s, m, tx, ty);
}

static int masked_prim(const orc_scene *s, uint32_t i) { return s->mid && s->mid[i] >= 0; }

/* radiation predicates with the mask applied at the hit (bvh_traversal.glsl:143-148) */
static int isect_patch_rad(const orc_scene *s, uint32_t i, v3 o, v3 d, float *t_out) {
    float ua, ub;
    if (!isect_patch_rad_uv(s, i, o, d, t_out, &ua, &ub)) return 0;
    if (!masked_prim(s, i)) return 1;
    /* u = d.a / |a|^2 with |a| the rounded length (intersections.glsl:64-66) */
    float am = sqrtf(s->aa[i]), bm = sqrtf(s->bb[i]);
    return mask_opaque_orc(s, i, ua / (am * am), ub / (bm * bm));
}

static int isect_tri_rad(const orc_scene *s, uint32_t i, v3 o, v3 d, float *t_out) {
    float ua, ub;
    if (!isect_tri_rad_uv(s, i, o, d, t_out, &ua, &ub)) return 0;
    if (!masked_prim(s, i)) return 1;
    return mask_opaque_orc(s, i, ua, ub);
}

/* Origin rejection (direct_raygen.comp:431-473, diffuse_raygen.comp:411-455): patch, up to 10 redraws of
 * a uniform point of the whole patch; triangle, up to 50 draws of a canonical-triangle point looked up
 * with floor() and no abs().  Returns 0 when the ray is dropped (triangle: every draw transparent;
 * diffuse patch: the last draw transparent). */
static int origin_mask_orc(const orc_scene *s, uint32_t i, int diffuse, uint32_t *st, float *spx, float *spy) {
    if (!masked_prim(s, i)) return 1;
    if (is_patch(s->type[i])) {
        int ok = mask_opaque_orc(s, i, *spx + 0.5f, *spy + 0.5f);
        for (int k = 0; k < 10 && !ok; k++) {
            *spx = -0.5f + orc_rnd(st);
            *spy = -0.5f + orc_rnd(st);
            ok = mask_opaque_orc(s, i, *spx + 0.5f, *spy + 0.5f);
        }
        return ok || !diffuse;
    }
    int m = s->mid[i];
    float fx = (float)((int)s->msize[2 * m] - 1), fy = (float)((int)s->msize[2 * m + 1] - 1);
    const float *uv = s->muv + (size_t)i * 8;
    int masked = 1;
    for (int a = 0; a < 50 && masked; a++) {
        float r1 = orc_rnd(st), r2 = orc_rnd(st);
        if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
        *spx = r2;
        *spy = r1 + r2;
        float tu = (uv[0] + r1 * (uv[2] - uv[0])) + r2 * (uv[4] - uv[0]);
        float tv = 1.0f - ((uv[1] + r1 * (uv[3] - uv[1])) + r2 * (uv[5] - uv[1]));
        masked = !mask_texel_orc(s, m, tex_int(floorf(fx * tu)), tex_int(floorf(fy * tv)));
    }
    return !masked;
}

int orc_scene_set_masks(orc_scene *s, uint32_t nmasks, const uint8_t *texels, const uint32_t *sizes,
                        const int32_t *mask_ids, const float *uv, const int32_t *uv_ids) {
    free(s->tex); free(s->msize); free(s->moff); free(s->mid); free(s->muv); free(s->muvid);
    s->tex = NULL; s->msize = NULL; s->moff = NULL; s->mid = NULL; s->muv = NULL; s->muvid = NULL;
    s->nmasks = nmasks;
    if (!nmasks) return 0;
    s->msize = (uint32_t *)malloc(sizeof(uint32_t) * 2 * nmasks);
    s->moff = (uint64_t *)malloc(sizeof(uint64_t) * nmasks);
    uint64_t tot = 0;
    for (uint32_t m = 0; m < nmasks; m++) {
        s->msize[2 * m] = sizes[2 * m];
        s->msize[2 * m + 1] = sizes[2 * m + 1];
        s->moff[m] = tot;
        tot += (uint64_t)sizes[2 * m] * sizes[2 * m + 1];
    }
    s->tex = (uint8_t *)malloc(tot ? tot : 1);
    memcpy(s->tex, texels, tot);
    s->mid = (int32_t *)malloc(sizeof(int32_t) * (s->P ? s->P : 1));
    s->muv = (float *)calloc((size_t)(s->P ? s->P : 1) * 8, sizeof(float));
    s->muvid = (int32_t *)malloc(sizeof(int32_t) * (s->P ? s->P : 1));
    for (uint32_t i = 0; i < s->P; i++) {
        s->mid[i] = mask_ids[i];
        if (mask_ids[i] >= (int32_t)nmasks) return -1;
        s->muvid[i] = uv_ids ? uv_ids[i] : -1;
        if (uv) memcpy(s->muv + (size_t)i * 8, uv + (size_t)i * 8, 8 * sizeof(float));
    }
    return 0;
}

/* CD/src/CollisionDetection_RayTracing.cpp:1793-1847 */
static int isect_tri_cd(const orc_scene *s, uint32_t i, v3 o, v3 d, float *t_out) {
    if (g_watertight) { float ua, ub; return isect_wt_orc(s, i, o, d, 1, t_out, &ua, &ub); }
    const float EPS = 1e-5f;
    v3 v0 = s->verts[(size_t)i * 4];
    v3 e1 = sub(s->verts[i * 4 + 1], v0), e2 = sub(s->verts[i * 4 + 2], v0);
    v3 h = cross(d, e2);
    float a = e1.x * h.x + e1.y * h.y + e1.z * h.z;
    if (a > -EPS && a < EPS) return 0;
    float f = 1.0f / a;
    v3 sv = sub(o, v0);
    float u = f * (sv.x * h.x + sv.y * h.y + sv.z * h.z);
    if (u < -EPS || u > 1.0f + EPS) return 0;
    v3 q = cross(sv, e1);
    float v = f * (d.x * q.x + d.y * q.y + d.z * q.z);
    if (v < -EPS || u + v > 1.0f + EPS) return 0;
    float t = f * (e2.x * q.x + e2.y * q.y + e2.z * q.z);
    if (t > EPS) { *t_out = t; return 1; }
    return 0;
}

/* CD/src/CollisionDetection_RayTracing.cpp:1849-1884 */
static int isect_patch_cd(const orc_scene *s, uint32_t i, v3 o, v3 d, float *t_out) {
    if (g_watertight) { float ua, ub; return isect_wt_orc(s, i, o, d, 1, t_out, &ua, &ub); }
    const float EPS = 1e-8f;
    v3 v0 = s->verts[(size_t)i * 4];
    v3 n = nrm(cross(sub(s->verts[i * 4 + 1], v0), sub(s->verts[i * 4 + 2], v0)));
    v3 a = sub(s->verts[i * 4 + 1], v0), b = sub(s->verts[i * 4 + 3], v0);
    float denom = n.x * d.x + n.y * d.y + n.z * d.z;
    if (fabsf(denom) > EPS) {
        v3 ao = sub(v0, o);
        float t = (n.x * ao.x + n.y * ao.y + n.z * ao.z) / denom;
        if (t > EPS && t < 1e8f) {
            v3 p = add(o, scl(d, t));
            v3 dd = sub(p, v0);
            float da = a.x * dd.x + a.y * dd.y + a.z * dd.z;
            float db = b.x * dd.x + b.y * dd.y + b.z * dd.z;
            float a2 = a.x * a.x + a.y * a.y + a.z * a.z;
            float b2 = b.x * b.x + b.y * b.y + b.z * b.z;
            if (da >= -EPS && da <= a2 + EPS && db >= -EPS && db <= b2 + EPS) { *t_out = t; return 1; }
        }
    }
    return 0;
}

/* castRays' voxel test, CD/src/CollisionDetection_RayTracing.cpp:1640-1682: slab method over the box in
 * FP32 with IEEE division; a component below 1e-8 in magnitude is parallel; the distance is the entry if it
 * is at least 1e-6, else the exit (an origin inside the box) */
static int isect_voxel_cd(const orc_scene *s, uint32_t i, v3 o, v3 d, float *t_out) {
    const v3 lo = s->verts[(size_t)i * 4], hi = s->verts[(size_t)i * 4 + 1];
    const float dc[3] = {d.x, d.y, d.z}, oc[3] = {o.x, o.y, o.z};
    const float lc[3] = {lo.x, lo.y, lo.z}, hc[3] = {hi.x, hi.y, hi.z};
    float tn = -FLT_MAX, tf = FLT_MAX;
    for (int a = 0; a < 3; a++) {
        if (fabsf(dc[a]) < 1e-8f) {
            if (oc[a] < lc[a] || oc[a] > hc[a]) return 0;
        } else {
            float t1 = (lc[a] - oc[a]) / dc[a], t2 = (hc[a] - oc[a]) / dc[a];
            if (t1 > t2) { float x = t1; t1 = t2; t2 = x; }
            if (t1 > tn) tn = t1;
            if (t2 < tf) tf = t2;
            if (tn > tf) return 0;
        }
    }
    if (!(tf >= 0.0f)) return 0;
    float dist = tn >= 1e-6f ? tn : tf;
    if (!(dist >= 1e-6f)) return 0;
    *t_out = dist;
    return 1;
}

/* outward normal of the voxel face hit, CD/src/CollisionDetection_RayTracing.cpp:1689-1711 */
static v3 voxel_normal_cd(const orc_scene *s, uint32_t i, v3 p) {
    const v3 lo = s->verts[(size_t)i * 4], hi = s->verts[(size_t)i * 4 + 1];
    v3 c = scl(add(lo, hi), 0.5f), e = scl(sub(hi, lo), 0.5f), l = sub(p, c);
    float rx = fabsf(l.x) / e.x, ry = fabsf(l.y) / e.y, rz = fabsf(l.z) / e.z;
    if (rx >= ry && rx >= rz) return mk(l.x > 0 ? 1.0f : -1.0f, 0.f, 0.f);
    if (ry >= rz) return mk(0.f, l.y > 0 ? 1.0f : -1.0f, 0.f);
    return mk(0.f, 0.f, l.z > 0 ? 1.0f : -1.0f);
}

/* One primitive's predicate: castRays' in cd_mode (voxels included), else the radiation traversal's,
 * which tests patches, tiles and triangles only (R/shaders/common/bvh_traversal.glsl:118-140). */
static int isect_prim(const orc_scene *s, uint32_t i, v3 o, v3 d, int cd_mode, float *t) {
    const uint32_t ty = s->type[i];
    if (ty == 4) return cd_mode ? isect_voxel_cd(s, i, o, d, t) : 0;
    if (cd_mode) return is_patch(ty) ? isect_patch_cd(s, i, o, d, t) : isect_tri_cd(s, i, o, d, t);
    return is_patch(ty) ? isect_patch_rad(s, i, o, d, t) : isect_tri_rad(s, i, o, d, t);
}

/* conservative slab test in FP64 (boxes are inflated; culling decisions never change results) */
static int slab(const float *lo, const float *hi, const double o[3], const double inv[3], double tmax) {
    double tn = -1e300, tf = 1e300;
    for (int a = 0; a < 3; a++) {
        double t1 = (lo[a] - o[a]) * inv[a], t2 = (hi[a] - o[a]) * inv[a];
        if (isnan(t1) || isnan(t2)) { if (o[a] < lo[a] || o[a] > hi[a]) return 0; continue; }
        if (t1 > t2) { double t = t1; t1 = t2; t2 = t; }
        if (t1 > tn) tn = t1;
        if (t2 < tf) tf = t2;
    }
    return tn <= tf && tf >= 0.0 && tn <= tmax;
}

typedef int (*isect_fn)(const orc_scene *, uint32_t, v3, v3, float *);

/* Closest accepted hit; ties in t resolve to the lower primitive index (deterministic canonical rule
 * replacing the reference's traversal-order rule, R/shaders/common/bvh_traversal.glsl:328-334). */
static uint32_t closest(const orc_scene *s, v3 o, v3 d, float tmin, uint32_t skip, int brute, int cd_mode,
                        float tstart, float *t_out) {
    uint32_t best = MISS;
    float bt = tstart;
    double od[3] = {o.x, o.y, o.z}, inv[3] = {1.0 / (double)d.x, 1.0 / (double)d.y, 1.0 / (double)d.z};
    uint32_t stack[256];
    int sp = 0;
    uint32_t np = s->nnodes ? s->nodes[0].count : 0;
    (void)np;
    if (!s->nnodes) { *t_out = bt; return MISS; }
    if (brute) {
        for (uint32_t i = 0; i < s->P; i++) {
            uint32_t ty = s->type[i];
            if (i == skip || !(ty <= 1 || ty == 3 || ty == 4)) continue;
            float t;
            if (!isect_prim(s, i, o, d, cd_mode, &t) || !(t > tmin || (cd_mode && ty == 4 && t == tmin))) continue;
            if (t < bt || (t == bt && best != MISS && i < best)) { bt = t; best = i; }
        }
        *t_out = bt;
        return best;
    }
    stack[sp++] = 0;
    while (sp) {
        const onode *nd = &s->nodes[stack[--sp]];
        if (!slab(nd->lo, nd->hi, od, inv, (double)bt)) continue;
        if (nd->right == MISS) {
            for (uint32_t k = 0; k < nd->count; k++) {
                uint32_t i = s->pidx[nd->start + k];
                if (i == skip) continue;
                float t;
                /* castRays takes a patch or triangle beyond 1e-6 and a voxel at 1e-6 or beyond
                 * (CD/src/CollisionDetection_RayTracing.cpp:1562, 1595, 1681) */
                if (!isect_prim(s, i, o, d, cd_mode, &t) ||
                    !(t > tmin || (cd_mode && s->type[i] == 4 && t == tmin)))
                    continue;
                if (t < bt || (t == bt && best != MISS && i < best)) { bt = t; best = i; }
            }
        } else {
            if (sp + 2 > 256) abort();
            stack[sp++] = nd->right;
            stack[sp++] = nd->left;
        }
    }
    *t_out = bt;
    return best;
}

uint32_t orc_trace(const orc_scene *s, const float o[3], const float d[3], float tmin, uint32_t skip, int brute, float *t_out) {
    float t;
    uint32_t h = closest(s, ld3(o), ld3(d), tmin, skip, brute, 0, 1e30f, &t);
    if (t_out) *t_out = t;
    return h;
}

/* R/shaders/common/periodic_boundary.glsl:28-59 */
static float isect_bbox_face(const orc_scene *s, uint32_t f, v3 o, v3 d) {
    v3 v0 = s->bbox[f][0], v1 = s->bbox[f][1], v2 = s->bbox[f][2], v3_ = s->bbox[f][3];
    v3 n = nrm(cross(sub(v1, v0), sub(v2, v0)));
    float denom = dot(d, n);
    if (fabsf(denom) < 1e-8f) return -1.0f;
    float t = dot(sub(v0, o), n) / denom;
    if (t <= 1e-5f || t > 1e8f) return -1.0f;
    v3 p = add(o, scl(d, t));
    v3 dd = sub(p, v0);
    v3 a = sub(v1, v0), b = sub(v3_, v0);
    float da = dot(dd, a), db = dot(dd, b);
    if (da > 0.0f && da < dot(a, a) && db > 0.0f && db < dot(b, b)) return t;
    return -1.0f;
}

/* R/shaders/common/periodic_boundary.glsl:73-112; returns 1 and the wrapped origin if a face is hit */
static int periodic_wrap(const orc_scene *s, v3 o, v3 d, float bvh_t, v3 *wrapped) {
    float bt = bvh_t;
    int cf = -1;
    for (uint32_t f = 0; f < s->nbbox; f++) {
        float t = isect_bbox_face(s, f, o, d);
        if (t > 0.0f && t < bt) { bt = t; cf = (int)f; }
    }
    if (cf < 0) return 0;
    v3 hp = add(o, scl(d, bt));
    const float eps = 1e-5f;
    float wx = s->dom[1] - s->dom[0], wy = s->dom[3] - s->dom[2];
    *wrapped = hp;
    if (s->px && fabsf(hp.x - s->dom[0]) <= eps) wrapped->x += wx - eps;
    else if (s->px && fabsf(hp.x - s->dom[1]) <= eps) wrapped->x -= wx - eps;
    else if (s->py && fabsf(hp.y - s->dom[2]) <= eps) wrapped->y += wy - eps;
    else if (s->py && fabsf(hp.y - s->dom[3]) <= eps) wrapped->y -= wy - eps;
    return 1;
}

/* direct_raygen.comp:603-668 / diffuse_raygen.comp:498-560 without translucent covers: up to 32 traces,
 * each followed by a periodic-face check; the last traversal result is returned. */
static uint32_t trace_loop(const orc_scene *s, v3 o, v3 d, float tmin, uint32_t skip, int *exhausted) {
    uint32_t hit = MISS;
    v3 cur = o;
    *exhausted = 1;
    for (int step = 0; step < 32; ++step) {
        float t;
        hit = closest(s, cur, d, tmin, skip, 0, 0, 1e30f, &t);
        if (s->nbbox) {
            v3 w;
            if (periodic_wrap(s, cur, d, t, &w)) { cur = w; continue; }
        }
        *exhausted = 0;
        break;
    }
    return hit;
}

uint32_t orc_trace_periodic(const orc_scene *s, const float o[3], const float d[3], float tmin, uint32_t skip) {
    int ex;
    return trace_loop(s, ld3(o), ld3(d), tmin, skip, &ex);
}

/* ------------------------------------------------------------------------------------------------ */
/* ray generation                                                                                    */
/* ------------------------------------------------------------------------------------------------ */

/* direct_raygen.comp:360-480 for patch/tile (stratified, subdivisions (1,1)) and triangle samples */
/* *st receives the RNG state after the origin's draws (including mask redraws); returns 0 when the ray
 * is dropped by its texture mask */
static int direct_origin(const orc_scene *s, uint32_t i, uint32_t lx, uint32_t ly, uint32_t d, uint32_t seed, v3 *o,
                         uint32_t *st_out) {
    uint32_t Nrays = d * d;
    uint32_t ray_index = d * ly + lx;
    uint32_t st = orc_tea16(ray_index + Nrays * i, seed);
    float spx, spy;
    uint32_t ty = s->type[i];
    float Rx = orc_rnd(&st), Ry = orc_rnd(&st);
    if (is_patch(ty)) {
        spx = (-0.5f + (float)lx / (float)d) + Rx / (float)d;
        spy = (-0.5f + (float)ly / (float)d) + Ry / (float)d;
    } else {
        if (Rx < Ry) { spx = Rx; spy = Ry; } else { spx = Ry; spy = Rx; }
    }
    int ok = origin_mask_orc(s, i, 0, &st, &spx, &spy);
    *o = xform(s->T + (size_t)i * 16, spx, spy, 0.0f);
    if (st_out) *st_out = st;
    return ok;
}

void orc_direct_ray(const orc_scene *s, uint32_t prim, uint32_t lx, uint32_t ly, uint32_t d, uint32_t seed, float origin[3]) {
    v3 o;
    if (!direct_origin(s, prim, lx, ly, d, seed, &o, NULL)) o = mk(NAN, NAN, NAN);
    origin[0] = o.x; origin[1] = o.y; origin[2] = o.z;
}

/* diffuse_raygen.comp:375-482 */
static int diffuse_ray(const orc_scene *s, uint32_t i, uint32_t lx, uint32_t ly, uint32_t n, uint32_t face,
                       uint32_t seed, v3 *o, v3 *dir) {
    uint32_t dimx = n * n;
    uint32_t ray_index = n * ly + lx;
    uint32_t st = orc_tea16(ray_index + dimx * i, seed);
    uint32_t ty = s->type[i];
    float spx, spy;
    float Rx = orc_rnd(&st), Ry = orc_rnd(&st);
    if (is_patch(ty)) { spx = -0.5f + Rx; spy = -0.5f + Ry; }
    else { if (Rx < Ry) { spx = Rx; spy = Ry; } else { spx = Ry; spy = Rx; } }
    if (!origin_mask_orc(s, i, 1, &st, &spx, &spy)) return 0;
    *o = xform(s->T + (size_t)i * 16, spx, spy, 0.0f);
    float Rt = ((float)lx + orc_rnd(&st)) / (float)n;
    float Rp = ((float)ly + orc_rnd(&st)) / (float)n;
    float st_ = sqrtf(Rt);           /* sin(asin(sqrt(Rt))) */
    float ct = sqrtf(1.0f - Rt);     /* cos(asin(sqrt(Rt))) */
    float sp, cp;
    orc_sincos_2pi(Rp, &sp, &cp);
    v3 dl = mk(st_ * cp, st_ * sp, ct);
    const float *fr = s->frame + (size_t)i * 4;
    float ce = fr[0], se = fr[1], ca = fr[2], sa = fr[3];
    v3 dw;
    dw.x = dl.x * (ce * ca) - dl.y * sa + dl.z * (se * ca);
    dw.y = dl.x * (ce * sa) + dl.y * ca + dl.z * (se * sa);
    dw.z = -dl.x * se + dl.z * ce;
    v3 r = nrm(dw);
    if (face == 0 && s->twos[i] == 1) r = mk(-r.x, -r.y, -r.z);
    *dir = r;
    return 1;
}

void orc_diffuse_ray(const orc_scene *s, uint32_t prim, uint32_t lx, uint32_t ly, uint32_t n, uint32_t face,
                     uint32_t seed, float origin[3], float dir[3]) {
    v3 o, d;
    if (!diffuse_ray(s, prim, lx, ly, n, face, seed, &o, &d)) { o = mk(NAN, NAN, NAN); d = o; }
    origin[0] = o.x; origin[1] = o.y; origin[2] = o.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

static int traced_type(uint32_t ty) { return ty == 0 || ty == 1 || ty == 3; }

void orc_diffuse_hits_range(const orc_scene *s, uint32_t n, uint32_t seed, uint32_t p0, uint32_t p1, uint32_t *out) {
    uint32_t nn = n * n;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long ii = (long long)p0; ii < (long long)p1; ii++) {
        uint32_t i = (uint32_t)ii;
        for (uint32_t face = 0; face < 2; face++) {
            uint32_t *w = out + ((size_t)(i - p0) * 2 + (1 - face)) * nn; /* slot 0 = top (face 1), slot 1 = bottom */
            int launched = traced_type(s->type[i]) && !(face == 0 && s->twos[i] == 0);
            for (uint32_t r = 0; r < nn; r++) {
                if (!launched) { w[r] = 0xFFFFFFFEu; continue; }
                uint32_t lx = r % n, ly = r / n;
                v3 o, d;
                if (!diffuse_ray(s, i, lx, ly, n, face, seed, &o, &d)) { w[r] = 0xFFFFFFFEu; continue; }
                int ex;
                uint32_t h = trace_loop(s, o, d, 1e-5f, i, &ex);
                if (h == MISS) { w[r] = MISS; continue; }
                int top = dot(s->normal[h], d) < 0.0f;
                w[r] = h | (top ? 0u : 0x80000000u);
            }
        }
    }
}

void orc_diffuse_hits(const orc_scene *s, uint32_t n, uint32_t seed, uint32_t *out) {
    orc_diffuse_hits_range(s, n, seed, 0, s->P, out);
}

/* direct_raygen.comp:413-670 for collimated sources: per (prim, source) sums of ray strength over
 * unoccluded rays, split by the lit face. */
void orc_direct_sums_range(const orc_scene *s, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_dir,
                           uint32_t p0, uint32_t p1, double *out) {
    uint32_t Nrays = d * d;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long ii = (long long)p0; ii < (long long)p1; ii++) {
        uint32_t i = (uint32_t)ii;
        double *o_i = out + (size_t)(i - p0) * nsrc * 2;
        for (uint32_t k = 0; k < nsrc * 2; k++) o_i[k] = 0.0;
        if (!traced_type(s->type[i])) continue;
        v3 n = s->normal[i];
        for (uint32_t r = 0; r < Nrays; r++) {
            uint32_t lx = r % d, ly = r / d;
            v3 o;
            if (!direct_origin(s, i, lx, ly, d, seed, &o, NULL)) continue;
            for (uint32_t src = 0; src < nsrc; src++) {
                v3 dir = nrm(ld3(src_dir + src * 3));
                float strength = fabsf(dot(n, dir)) / (float)Nrays;
                int top = dot(dir, n) > 0.0f;
                uint32_t tw = s->twos[i];
                if (tw == 3 || (!top && tw == 0)) continue;
                int ex;
                uint32_t h = trace_loop(s, o, dir, 1e-4f, i, &ex);
                if (h != MISS && !ex) continue; /* opaque occluder */
                o_i[src * 2 + (top ? 0 : 1)] += (double)strength;
            }
        }
    }
}

void orc_direct_sums(const orc_scene *s, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_dir, double *out) {
    orc_direct_sums_range(s, d, seed, nsrc, src_dir, 0, s->P, out);
}

/* ------------------------------------------------------------------------------------------------ */
/* non-collimated sources (direct_raygen.comp:482-628)                                               */
/* ------------------------------------------------------------------------------------------------ */
/* make_rotation_matrix (transforms.glsl:49-64), ZYX Euler angles; rows of the column-major GLSL matrix,
 * evaluated in double and rounded to float */
void orc_rotation_matrix(const float rot[3], float m[9]) {
    double cx = cos((double)rot[0]), sx = sin((double)rot[0]);
    double cy = cos((double)rot[1]), sy = sin((double)rot[1]);
    double cz = cos((double)rot[2]), sz = sin((double)rot[2]);
    m[0] = (float)(cz * cy); m[1] = (float)(cz * sy * sx - sz * cx); m[2] = (float)(cz * sy * cx + sz * sx);
    m[3] = (float)(sz * cy); m[4] = (float)(sz * sy * sx + cz * cx); m[5] = (float)(sz * sy * cx - cz * sx);
    m[6] = (float)(-sy);     m[7] = (float)(cy * sx);                m[8] = (float)(cy * cx);
}

/* calc_sphere_strength (direct_raygen.comp:305-320); sin(acos(c)) = sqrt(1 - c^2) */
static float sphere_strength_orc(v3 n, v3 rd, float mag, float diam, uint32_t Nrays) {
    float sum = 0.0f;
    for (int j = 0; j < 10; j++) {
        float sp, cp;
        orc_sincos_2pi(((float)j + 0.5f) / 10.0f, &sp, &cp);
        for (int i = 0; i < 10; i++) {
            float c = 1.0f - (2.0f * ((float)i + 0.5f)) / 10.0f;
            float sn = sqrtf(fmaxf(0.0f, 1.0f - c * c));
            v3 L = mk(sn * cp, sn * sp, c);
            float ldr = dot(L, rd);
            if (ldr < 0.0f)
                sum += ((((1.0f / (float)Nrays) * fabsf(dot(n, rd))) * fabsf(ldr)) / (mag * mag)) / 100.0f * diam * diam;
        }
    }
    return sum;
}

/* direction and strength of one ray from o towards source (type, pos, width, m); returns 0 when the
 * rectangle/disk back-face test rejects it.  Draws from *st. */
static int source_sample_orc(uint32_t type, v3 pos, const float *width, const float *m, v3 o, v3 n, uint32_t Nrays,
                             uint32_t *st, v3 *dir, float *strength) {
    if (type == 1 || type == 2) {
        float u1 = orc_rnd(st), u2 = orc_rnd(st);
        float c = 1.0f - 2.0f * u1;
        float sn = sqrtf(fmaxf(0.0f, 1.0f - c * c));
        float sp, cp;
        orc_sincos_2pi(u2, &sp, &cp);
        float r = width[0] * 0.5f;
        v3 q = mk(r * (sn * cp), r * (sn * sp), r * c);
        v3 rd = sub(add(q, pos), o);
        float mag = sqrtf(dot(rd, rd));
        *dir = nrm(rd);
        *strength = sphere_strength_orc(n, *dir, mag, width[0], Nrays);
        return 1;
    }
    float px, py, scale = 1.0f;
    if (type == 3) {
        float Rx = orc_rnd(st), Ry = orc_rnd(st);
        px = (-0.5f + Rx) * width[0];
        py = (-0.5f + Ry) * width[1];
    } else { /* disk: concentric map (direct_raygen.comp:227-266) */
        float Rx = orc_rnd(st), Ry = orc_rnd(st);
        float x = -1.0f + 2.0f * Rx, y = -1.0f + 2.0f * Ry, r, phi8;
        if (x > -y) {
            if (x > y) { r = x; phi8 = y / x; } else { r = y; phi8 = 2.0f - x / y; }
        } else {
            if (x < y) { r = -x; phi8 = 4.0f + y / x; } else { r = -y; phi8 = (y != 0.0f) ? 6.0f - x / y : 0.0f; }
        }
        float sp, cp;
        orc_sincos_2pi(phi8 * 0.125f, &sp, &cp); /* phi = phi8 * pi/4 */
        px = r * cp;
        py = r * sp;
        scale = width[0];
    }
    v3 lp = mk(m[0] * px + m[1] * py, m[3] * px + m[4] * py, m[6] * px + m[7] * py);
    v3 ln = nrm(mk(m[2], m[5], m[8]));
    v3 rd = (type == 3) ? sub(add(lp, pos), o) : sub(add(scl(lp, scale), pos), o);
    if (dot(rd, ln) > 0.0f) return 0;
    float mag = sqrtf(dot(rd, rd));
    *dir = nrm(rd);
    float g = (((1.0f / (float)Nrays) * fabsf(dot(n, *dir))) * fabsf(dot(ln, *dir))) / (mag * mag);
    *strength = (type == 3) ? (g * (width[0] * width[1])) / 3.14159265359f : (g * scale) * scale;
    return 1;
}

/* occlusion towards a non-collimated source: the reference loop with the "hit beyond the source" rule */
static int source_blocked_orc(const orc_scene *s, v3 o, v3 d, uint32_t skip, float dist0) {
    v3 cur = o;
    for (int step = 0; step < 32; ++step) {
        float t;
        uint32_t h = closest(s, cur, d, 1e-4f, skip, 0, 0, 1e30f, &t);
        if (s->nbbox) {
            v3 w;
            if (periodic_wrap(s, cur, d, t, &w)) { cur = w; continue; }
        }
        if (h == MISS) return 0;
        return t < dist0;
    }
    return 0;
}

void orc_direct_sums_src_range(const orc_scene *s, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_pos,
                               const uint32_t *types, const float *widths, const float *rotations, uint32_t p0,
                               uint32_t p1, double *out) {
    uint32_t Nrays = d * d;
    float *mats = (float *)calloc((size_t)nsrc * 9 + 1, sizeof(float));
    for (uint32_t q = 0; q < nsrc; q++) {
        uint32_t ty = types ? types[q] : 0;
        if (ty == 3 || ty == 4) orc_rotation_matrix(rotations + 3 * q, mats + 9 * q);
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (long long ii = (long long)p0; ii < (long long)p1; ii++) {
        uint32_t i = (uint32_t)ii;
        double *o_i = out + (size_t)(i - p0) * nsrc * 2;
        for (uint32_t k = 0; k < nsrc * 2; k++) o_i[k] = 0.0;
        if (!traced_type(s->type[i])) continue;
        v3 n = s->normal[i];
        uint32_t tw = s->twos[i];
        for (uint32_t r = 0; r < Nrays; r++) {
            uint32_t lx = r % d, ly = r / d;
            /* origin sample (two draws, then any mask redraws), then each source in order draws from the
             * same state */
            uint32_t st;
            v3 o;
            if (!direct_origin(s, i, lx, ly, d, seed, &o, &st)) continue;
            for (uint32_t src = 0; src < nsrc; src++) {
                uint32_t ty = types ? types[src] : 0;
                v3 dir;
                float strength;
                if (ty == 0) {
                    dir = nrm(ld3(src_pos + src * 3));
                    strength = fabsf(dot(n, dir)) / (float)Nrays;
                } else if (!source_sample_orc(ty, ld3(src_pos + src * 3), widths + 2 * src, mats + 9 * src, o, n, Nrays,
                                              &st, &dir, &strength)) {
                    continue;
                }
                int top = dot(dir, n) > 0.0f;
                if (tw == 3 || (!top && tw == 0)) continue;
                if (ty == 0) {
                    int ex;
                    uint32_t h = trace_loop(s, o, dir, 1e-4f, i, &ex);
                    if (h != MISS && !ex) continue;
                } else {
                    v3 dv = sub(ld3(src_pos + src * 3), o);
                    if (source_blocked_orc(s, o, dir, i, sqrtf(dot(dv, dv)))) continue;
                }
                o_i[src * 2 + (top ? 0 : 1)] += (double)strength;
            }
        }
    }
    free(mats);
}

void orc_direct_sums_src(const orc_scene *s, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_pos,
                         const uint32_t *types, const float *widths, const float *rotations, double *out) {
    orc_direct_sums_src_range(s, d, seed, nsrc, src_pos, types, widths, rotations, 0, s->P, out);
}

/* ------------------------------------------------------------------------------------------------ */
/* castRaysSoA (CD/src/CollisionDetection_RayTracing.cpp:551-680, 980-1080, 1555-1620)               */
/* ------------------------------------------------------------------------------------------------ */
void orc_cast_rays(const orc_scene *s, uint64_t n, const float *origins, const float *dirs, float max_dist,
                   float *out_t, float *out_normal, uint32_t *out_uuid, int brute) {
#pragma omp parallel for schedule(dynamic, 256)
    for (long long r = 0; r < (long long)n; r++) {
        v3 o = ld3(origins + r * 3), d = ld3(dirs + r * 3);
        float start = max_dist > 0 ? max_dist : FLT_MAX;
        float t;
        /* the CD acceptance is distance > 1e-6 and distance < closest */
        uint32_t h = closest(s, o, d, 1e-6f, MISS, brute, 1, start, &t);
        if (h == MISS) { out_uuid[r] = MISS; continue; }
        out_uuid[r] = s->uuid[h];
        out_t[r] = t;
        if (s->type[h] == 4) {
            v3 nn = voxel_normal_cd(s, h, add(o, scl(d, t)));
            out_normal[r * 3] = nn.x; out_normal[r * 3 + 1] = nn.y; out_normal[r * 3 + 2] = nn.z;
            continue;
        }
        const v3 *v = s->verts + (size_t)h * 4;
        v3 nn = cross(sub(v[1], v[0]), sub(v[2], v[0]));
        float mag = sqrtf(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
        if (mag > 1e-8f) {
            float inv = 1.0f / mag;
            nn = scl(nn, inv);
            v3 ip = add(o, scl(d, t));
            v3 to = sub(o, ip);
            if (nn.x * to.x + nn.y * to.y + nn.z * to.z < 0) nn = scl(nn, -1.0f);
        } else {
            nn = mk(-d.x, -d.y, -d.z);
            float m2 = sqrtf(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
            float inv = 1.0f / m2;
            nn = scl(nn, inv);
        }
        out_normal[r * 3] = nn.x; out_normal[r * 3 + 1] = nn.y; out_normal[r * 3 + 2] = nn.z;
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* runBand pass schedule (R/src/RadiationModel.cpp:3470-4672)                                        */
/* ------------------------------------------------------------------------------------------------ */
/* cameras                                                                                           */
/* ------------------------------------------------------------------------------------------------ */

/* camera_rotate_point (R/shaders/common/transforms.glsl): Rz(phi) * Ry(theta) * q */
static v3 cam_rotate(v3 q, float st, float ct, float sp, float cp) {
    return mk((cp * ct) * q.x - sp * q.y + (cp * st) * q.z, (sp * ct) * q.x + cp * q.y + (sp * st) * q.z,
              -st * q.x + ct * q.z);
}

/* theta = elevation - pi/2, phi = pi/2 - azimuth (camera_raygen.comp:258-259); sines and cosines are
 * taken in double from the FP32 angles and rounded. */
static void cam_angles(const orc_camera *c, float *st, float *ct, float *sp, float *cp) {
    const float pi = 3.14159265359f;
    double th = (double)(-0.5f * pi + c->direction[0]), ph = (double)(0.5f * pi - c->direction[1]);
    *st = (float)sin(th); *ct = (float)cos(th); *sp = (float)sin(ph); *cp = (float)cos(ph);
}

/* Ray through pixel (ii, jj) at sub-pixel offset (rx, ry) and lens point (lx, ly) in lens radii
 * (camera_raygen.comp:228-261, pixel_label_raygen.comp:93-115). */
static void cam_ray(const orc_camera *c, float st, float ct, float sp, float cp, uint32_t ii, uint32_t jj, float rx,
                    float ry, float lx, float ly, v3 *o, v3 *d) {
    float mult = 1.0f / c->fov_aspect;
    float spy = -0.5f + ((float)ii + rx) / (float)c->resolution_full[0];
    float spz = (0.5f - ((float)jj + ry) / (float)c->resolution_full[1]) * mult;
    v3 p = mk(c->focal_length, spy / c->viewplane_length * c->focal_length, spz / c->viewplane_length * c->focal_length);
    v3 org = mk(0.0f, (0.5f * lx) * c->lens_diameter, (0.5f * ly) * c->lens_diameter);
    v3 dir = sub(p, org);
    *o = add(cam_rotate(org, st, ct, sp, cp), ld3(c->position));
    *d = nrm(cam_rotate(dir, st, ct, sp, cp));
}

/* Concentric square-to-disk map (Shirley & Chiu 1997; camera_raygen.comp:96-128) */
static void cam_disk(float Rx, float Ry, float *dx, float *dy) {
    float x = -1.0f + 2.0f * Rx, y = -1.0f + 2.0f * Ry, r, k;
    if (x > -y) {
        if (x > y) { r = x; k = y / x; }
        else { r = y; k = 2.0f - x / y; }
    } else {
        if (x < y) { r = -x; k = 4.0f + y / x; }
        else { r = -y; k = (y != 0.0f) ? 6.0f - x / y : 0.0f; }
    }
    float sk, ck;
    orc_sincos_2pi(k * 0.125f, &sk, &ck);
    *dx = r * ck;
    *dy = r * sk;
}

/* Closest hit with up to 10 periodic wraps (camera_raygen.comp:266-285); *depth sums the distances to
 * the faces crossed. */
static uint32_t cam_trace(const orc_scene *s, v3 *o, v3 d, float *t, float *depth) {
    uint32_t hit = MISS;
    *depth = 0.0f;
    for (int wrap = 0; wrap < 10; ++wrap) {
        hit = closest(s, *o, d, 1e-5f, MISS, 0, 0, 1e30f, t);
        if (!s->nbbox) break;
        float bt = *t;
        int cf = -1;
        for (uint32_t f = 0; f < s->nbbox; f++) {
            float tf = isect_bbox_face(s, f, *o, d);
            if (tf > 0.0f && tf < bt) { bt = tf; cf = (int)f; }
        }
        if (cf < 0) break;
        v3 w;
        periodic_wrap(s, *o, d, *t, &w);
        *depth += bt;
        *o = w;
        hit = MISS;
    }
    return hit;
}

/* sphere_source_visible (camera_raygen.comp:131-143) */
static int cam_sphere_vis(v3 o, v3 d, v3 c, float radius, float tmax) {
    v3 oc = sub(o, c);
    float b = dot(oc, d);
    float cc = dot(oc, oc) - radius * radius;
    float disc = b * b - cc;
    if (!(disc >= 0.0f)) return 0;
    float ts = -b - sqrtf(disc);
    return ts > 0.0f && ts < tmax;
}

/* rectangle_source_visible / disk_source_visible (camera_raygen.comp:146-203): normal = third row of
 * the rotation matrix; disk width[0] = radius. */
static int cam_planar_vis(int disk, v3 o, v3 d, v3 c, const float *w, const float *m, float tmax, float *cosang) {
    v3 n = mk(m[6], m[7], m[8]);
    float denom = dot(d, n);
    if (denom >= -1e-6f) return 0;
    float t = dot(sub(c, o), n) / denom;
    if (t <= 0.0f || t >= tmax) return 0;
    v3 off = sub(add(o, scl(d, t)), c);
    int in;
    if (disk) in = dot(off, off) <= w[0] * w[0];
    else {
        float lx = m[0] * off.x + m[3] * off.y + m[6] * off.z;
        float ly = m[1] * off.x + m[4] * off.y + m[7] * off.z;
        in = fabsf(lx) <= w[0] * 0.5f && fabsf(ly) <= w[1] * 0.5f;
    }
    if (in) *cosang = -denom;
    return in;
}

/* Radiance of the sources met before tmax for band b, in source order (camera_raygen.comp:331-365 on a
 * hit; 414-446 on a miss, where collimated and sun-sphere sources add the solar disk). */
static float cam_sources(const orc_camera_light *L, const float *mats, v3 o, v3 d, uint32_t b, float tmax, int miss) {
    const float pi = 3.14159265359f;
    float rad = 0.0f;
    for (uint32_t s = 0; s < L->nsources; s++) {
        float flux = L->source_flux[(size_t)s * L->nbands + b];
        if (flux <= 0.0f) continue;
        uint32_t ty = L->source_type ? L->source_type[s] : 0u;
        v3 sp = ld3(L->source_pos + 3 * s);
        const float *w = L->source_width ? L->source_width + 2 * s : NULL;
        float ca;
        if (ty == 0u || ty == 2u) {
            if (miss && d.x * L->sun[0] + d.y * L->sun[1] + d.z * L->sun[2] >= L->solar_cos && L->solar_radiance[b] > 0.0f)
                rad += L->solar_radiance[b];
        } else if (ty == 1u) {
            float r = w[0] * 0.5f;
            if (cam_sphere_vis(o, d, sp, r, tmax)) rad += (flux / (4.0f * pi * r * r)) / pi;
        } else if (ty == 3u) {
            if (cam_planar_vis(0, o, d, sp, w, mats + 9 * s, tmax, &ca)) rad += (flux / (w[0] * w[1])) * ca / pi;
        } else if (ty == 4u) {
            float r = w[0];
            if (cam_planar_vis(1, o, d, sp, w, mats + 9 * s, tmax, &ca)) rad += (flux / (pi * r * r)) * ca / pi;
        }
    }
    return rad;
}

/* Prague sky pattern (R/shaders/common/sky_distribution.glsl), 1 when the normalisation is not positive */
static float cam_sky(v3 d, const float *sun, const float *prm) {
    const float pi = 3.14159265359f;
    if (!(prm[3] > 0.0f)) return 1.0f;
    float c = fminf(1.0f, fmaxf(-1.0f, d.x * sun[0] + d.y * sun[1] + d.z * sun[2]));
    float gamma = acosf(c) * (180.0f / pi);
    float cos_theta = fmaxf(0.0f, d.z);
    float circ = 1.0f + prm[0] * expf(-gamma / prm[1]);
    float horizon = 1.0f + (prm[2] - 1.0f) * (1.0f - cos_theta);
    return circ * horizon * prm[3] * pi;
}

/* Blinn-Phong specular term of a camera hit on primitive h for band b (camera_raygen.comp:362-409): per
 * source, radiation_specular x the camera weight x 0.25 (Fresnel reflectance); when positive, x the scale
 * (mode 2: specular_scale, else 1) x max(0, n.h)^e x (e + 2) / (samples x 2 pi), with the half vector h =
 * normalize(l - d), l the direction of a collimated or sun-sphere source or towards the centre of any
 * other source from the hit point o + d t, and n the primitive's normal.  FP32 throughout. */
static float cam_specular(const orc_scene *s, const orc_camera_light *L, v3 o, v3 d, float t, uint32_t h,
                          uint32_t b, uint32_t spp) {
    const float pi = 3.14159265359f;
    const float e = L->spec_exp[h];
    float acc = 0.0f;
    if (!(e > 0.0f)) return 0.0f;
    for (uint32_t k = 0; k < L->nsources; k++) {
        float spec = L->spec[((size_t)k * s->P + h) * L->nbands + b];
        spec = spec * (L->spec_weight[(size_t)k * L->nbands + b] * 0.25f);
        if (!(spec > 0.0f)) continue;
        uint32_t ty = L->source_type ? L->source_type[k] : 0u;
        v3 l = ld3(L->source_pos + 3 * k);
        if (!(ty == 0u || ty == 2u)) l = sub(l, add(o, scl(d, t)));
        l = nrm(l);
        v3 hv = nrm(sub(l, d));
        float scale = L->spec_mode == 2u ? L->spec_scale[h] : 1.0f;
        float ch = fmaxf(0.0f, hv.x * s->normal[h].x + hv.y * s->normal[h].y + hv.z * s->normal[h].z);
        acc += spec * scale * powf(ch, e) * (e + 2.0f) / ((float)spp * 2.0f * pi);
    }
    return acc;
}

void orc_camera_pixels(const orc_scene *s, const orc_camera *c, const orc_camera_light *L, double *pixels) {
    const float pi = 3.14159265359f;
    float st, ct, sp, cp;
    cam_angles(c, &st, &ct, &sp, &cp);
    const uint32_t B = L->nbands, spp = c->antialiasing_samples ? c->antialiasing_samples : 1u;
    const uint32_t rx = c->resolution[0], ry = c->resolution[1];
    const size_t npix = (size_t)c->resolution_full[0] * c->resolution_full[1];
    float *mats = (float *)calloc(9 * (size_t)(L->nsources + 1), sizeof(float));
    for (uint32_t k = 0; k < L->nsources; k++)
        if (L->source_rotation) orc_rotation_matrix(L->source_rotation + 3 * k, mats + 9 * k);
    const float strength = 1.0f / (float)spp;
    const uint32_t dimx = rx * ry;
    for (uint32_t py = 0; py < ry; py++)
        for (uint32_t px = 0; px < rx; px++) {
            const uint32_t ii = c->pixel_offset[0] + px, jj = c->pixel_offset[1] + py;
            const size_t pix = (size_t)jj * c->resolution_full[0] + ii;
            const uint32_t indx = rx * py + px;
            for (uint32_t smp = 0; smp < spp; smp++) {
                uint32_t state = orc_tea16(indx + dimx * smp, c->random_seed);
                float Rx = orc_rnd(&state), Ry = orc_rnd(&state), lx = 0.0f, ly = 0.0f;
                if (c->lens_diameter > 0.0f) {
                    float a = orc_rnd(&state), bq = orc_rnd(&state);
                    cam_disk(a, bq, &lx, &ly);
                }
                v3 o, d;
                cam_ray(c, st, ct, sp, cp, ii, jj, Rx, Ry, lx, ly, &o, &d);
                float t, depth;
                uint32_t h = cam_trace(s, &o, d, &t, &depth);
                if (h != MISS) {
                    int top = dot(s->normal[h], d) < 0.0f || s->type[h] == 4u;
                    for (uint32_t b = 0; b < B; b++) {
                        float srad = cam_sources(L, mats, o, d, b, t, 0);
                        float surf = (top ? L->out_top : L->out_bottom)[(size_t)h * B + b] * strength;
                        if (srad > 0.0f) surf += srad * strength;
                        if (L->spec) surf += cam_specular(s, L, o, d, t, h, b, spp);
                        pixels[(size_t)b * npix + pix] += (double)(surf / pi);
                    }
                } else {
                    for (uint32_t b = 0; b < B; b++) {
                        float rad = cam_sources(L, mats, o, d, b, 1e30f, 1);
                        if (rad <= 0.0f && L->sky_radiance[b] > 0.0f) rad = L->sky_radiance[b] * cam_sky(d, L->sun, L->sky_params + 4 * b);
                        float df = L->camera_diffuse_flux ? L->camera_diffuse_flux[b] : 0.0f;
                        if (L->emission && L->emission[b] && df > 0.0f) rad += df / pi;
                        if (rad > 0.0f) pixels[(size_t)b * npix + pix] += (double)(rad * strength);
                    }
                }
            }
        }
    free(mats);
}

void orc_pixel_labels(const orc_scene *s, const orc_camera *c, uint32_t *labels, float *depths) {
    float st, ct, sp, cp;
    cam_angles(c, &st, &ct, &sp, &cp);
    const v3 fwd = cam_rotate(mk(1.0f, 0.0f, 0.0f), st, ct, sp, cp);
    for (uint32_t py = 0; py < c->resolution[1]; py++)
        for (uint32_t px = 0; px < c->resolution[0]; px++) {
            const uint32_t ii = c->pixel_offset[0] + px, jj = c->pixel_offset[1] + py;
            const size_t pix = (size_t)jj * c->resolution_full[0] + ii;
            v3 o, d;
            cam_ray(c, st, ct, sp, cp, ii, jj, 0.5f, 0.5f, 0.0f, 0.0f, &o, &d);
            float t, depth;
            uint32_t h = cam_trace(s, &o, d, &t, &depth);
            if (h != MISS) {
                labels[pix] = s->uuid[h] + 1u;
                depths[pix] = fabsf(dot(fwd, d)) * (depth + t);
            } else {
                depths[pix] = -1.0f;
            }
        }
}

/* ------------------------------------------------------------------------------------------------ */

/* One diffuse pass, faces 1 then 0 (RadiationModel.cpp:4204-4209, diffuse_raygen.comp:330-656). */
/* evaluate_sky_distribution (R/shaders/common/sky_distribution.glsl:40-116) for band b: power law
 * psi^-K * norm about the peak direction (psi clamped to 1 degree) when K > 0, else the Prague pattern
 * (times its normalisation and pi) when that normalisation is positive, else 1 */
static float sky_weight_orc(const orc_band_params *p, uint32_t b, v3 d) {
    const float pi = 3.14159265359f;
    float K = p->diffuse_extinction ? p->diffuse_extinction[b] : 0.0f;
    const float *pk = p->diffuse_peak_dir ? p->diffuse_peak_dir + 3 * b : NULL;
    if (K > 0.0f) {
        float c = fminf(1.0f, fmaxf(-1.0f, pk[0] * d.x + pk[1] * d.y + pk[2] * d.z));
        float psi = fmaxf(acosf(c), pi / 180.0f);
        return powf(psi, -K) * p->diffuse_dist_norm[b];
    }
    const float *prm = p->sky_params ? p->sky_params + 4 * b : NULL;
    if (prm && prm[3] > 0.0f) {
        float c = fminf(1.0f, fmaxf(-1.0f, d.x * pk[0] + d.y * pk[1] + d.z * pk[2]));
        float gamma = acosf(c) * (180.0f / pi);
        float cos_theta = fmaxf(0.0f, d.z);
        float circ = 1.0f + prm[0] * expf(-gamma / prm[1]);
        float horizon = 1.0f + (prm[2] - 1.0f) * (1.0f - cos_theta);
        return circ * horizon * prm[3] * pi;
    }
    return 1.0f;
}

/* ------------------------------------------------------------------------------------------------ */
/* translucent covers (R/shaders/common/glass_cover.glsl, direct_raygen.comp:587-664,                 */
/* diffuse_raygen.comp:484-560)                                                                       */
/* ------------------------------------------------------------------------------------------------ */
/* glass_cover.glsl:15-53: Snell refraction, Fresnel reflectance of each polarisation
 * (sin^2(tr - t) / sin^2(tr + t), tan^2(tr - t) / tan^2(tr + t); the normal-incidence limit below 1e-3 rad),
 * Bouguer attenuation exp(-KL / cos tr) inside the sheet and the sums over internal reflections of a slab,
 * averaged over the two polarisations. */
void orc_glass_tra(float cos_theta, float n, float KL, float out[3]) {
    cos_theta = fmaxf(1e-4f, fminf(1.0f, cos_theta));
    float theta = acosf(fmaxf(-1.0f, fminf(1.0f, cos_theta)));
    float sin_t = sinf(theta);
    float sin_tr = sin_t / n;
    float cos_tr = sqrtf(fmaxf(0.0f, 1.0f - sin_tr * sin_tr));
    float theta_r = asinf(fmaxf(-1.0f, fminf(1.0f, sin_tr)));
    float r_par, r_per;
    if (theta < 1e-3f) {
        float r0 = ((n - 1.0f) / (n + 1.0f)) * ((n - 1.0f) / (n + 1.0f));
        r_par = r0;
        r_per = r0;
    } else {
        float sm = sinf(theta_r - theta), sp = sinf(theta_r + theta);
        float tm = tanf(theta_r - theta), tp = tanf(theta_r + theta);
        r_per = (sm * sm) / fmaxf(1e-12f, sp * sp);
        r_par = (tm * tm) / fmaxf(1e-12f, tp * tp);
    }
    float tau_a = KL > 0.0f ? expf(-KL / fmaxf(1e-4f, cos_tr)) : 1.0f;
    float tau = 0.0f, rho = 0.0f;
    for (int pol = 0; pol < 2; pol++) {
        float r = pol == 0 ? r_per : r_par;
        float den = fmaxf(1e-6f, 1.0f - (r * tau_a) * (r * tau_a));
        float ti = tau_a * (1.0f - r) * (1.0f - r) / den;
        tau += 0.5f * ti;
        rho += 0.5f * (r * (1.0f + tau_a * ti));
    }
    tau = fmaxf(0.0f, fminf(1.0f, tau));
    rho = fmaxf(0.0f, fminf(1.0f, rho));
    out[0] = tau;
    out[1] = rho;
    out[2] = fmaxf(0.0f, 1.0f - tau - rho);
}

static void atomic_addd(double *a, double v) {
#pragma omp atomic
    *a += v;
}

/* The reference's trace loop with cover pass-through: up to 32 closest-hit traces, a periodic face wraps
 * the ray, a miss (or, with dist0 >= 0, a hit at or beyond dist0) reaches the end, a hit on a non-cover
 * stops it, and a hit on a cover multiplies ctau[b] by the cover's transmittance, pushes
 * incoming_b = base[b] * ctau[b] (before the crossing) x alpha into the cover's radiation_in and x rho into
 * its scatter buffer of the face hit (top when the ray meets the cover's normal head-on; nothing when
 * incoming_b is 0), and continues from just past the cover.  status: 0 reached, 1 stopped by the returned
 * primitive, 2 the traces ran out (the last trace's hit is returned). */
static uint32_t trace_covers_orc(const orc_scene *s, const orc_band_params *p, const uint8_t *cover, v3 o, v3 d,
                                 float tmin, uint32_t skip, float dist0, const float *base, float *ctau, double *in,
                                 double *st, double *sb, int *status) {
    uint32_t B = p->nbands;
    v3 cur = o;
    uint32_t h = MISS;
    for (int step = 0; step < 32; ++step) {
        float t;
        h = closest(s, cur, d, tmin, skip, 0, 0, 1e30f, &t);
        if (s->nbbox) {
            v3 w;
            if (periodic_wrap(s, cur, d, t, &w)) { cur = w; continue; }
        }
        if (h == MISS) { *status = 0; return MISS; }
        if (dist0 >= 0.0f && t >= dist0) { *status = 0; return h; }
        if (!cover[h]) { *status = 1; return h; }
        v3 nh = s->normal[h];
        float c = fabsf(dot(d, nh));
        int top = dot(nh, d) < 0.0f;
        for (uint32_t b = 0; b < B; b++) {
            float tra[3];
            orc_glass_tra(c, p->glass_n[(size_t)h * B + b], p->glass_KL[(size_t)h * B + b], tra);
            float inc = base[b] * ctau[b];
            if (inc > 0.0f) {
                size_t k = (size_t)h * B + b;
                atomic_addd(&in[k], (double)(inc * tra[2]));
                atomic_addd(top ? &st[k] : &sb[k], (double)(inc * tra[1]));
            }
            ctau[b] *= tra[0];
        }
        cur = add(cur, scl(d, t + 1e-4f));
    }
    *status = 2;
    return h;
}

/* Direct pass with covers (direct_raygen.comp:413-731): per primitive, ray and source, the strength of an
 * unblocked ray times the source flux and the transmittance of the covers crossed, deposited per band on
 * the lit face; covers take their pushed shares.  Deposits are atomic (covers receive from every primitive). */
static void direct_glass(const orc_scene *s, const orc_band_params *p, const uint8_t *cover, double *in, double *st,
                         double *sb, const double *areas, double *diag) {
    const uint32_t B = p->nbands, S = p->nsources, d = (uint32_t)ceil(sqrt((double)p->direct_rays)), Nrays = d * d;
    float *mats = (float *)calloc((size_t)S * 9 + 1, sizeof(float));
    for (uint32_t q = 0; q < S; q++) {
        uint32_t ty = p->source_type ? p->source_type[q] : 0;
        if (ty == 3 || ty == 4) orc_rotation_matrix(p->source_rotation + 3 * q, mats + 9 * q);
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (long long ii = 0; ii < (long long)s->P; ii++) {
        uint32_t i = (uint32_t)ii;
        if (!traced_type(s->type[i])) continue;
        v3 n = s->normal[i];
        uint32_t tw = s->twos[i];
        float base[256], ctau[256];
        for (uint32_t r = 0; r < Nrays; r++) {
            uint32_t stt;
            v3 o;
            if (!direct_origin(s, i, r % d, r / d, d, p->direct_seed, &o, &stt)) continue;
            for (uint32_t src = 0; src < S; src++) {
                uint32_t ty = p->source_type ? p->source_type[src] : 0;
                v3 dir;
                float strength;
                if (ty == 0) {
                    dir = nrm(ld3(p->source_dir + src * 3));
                    strength = fabsf(dot(n, dir)) / (float)Nrays;
                } else if (!source_sample_orc(ty, ld3(p->source_dir + src * 3), p->source_width + 2 * src,
                                              mats + 9 * src, o, n, Nrays, &stt, &dir, &strength)) {
                    continue;
                }
                int top = dot(dir, n) > 0.0f;
                if (tw == 3 || (!top && tw == 0)) continue;
                float dist0 = -1.0f;
                if (ty != 0) {
                    v3 dv = sub(ld3(p->source_dir + src * 3), o);
                    dist0 = sqrtf(dot(dv, dv));
                }
                for (uint32_t b = 0; b < B; b++) { base[b] = p->source_flux[src * B + b] * strength; ctau[b] = 1.0f; }
                int status;
                trace_covers_orc(s, p, cover, o, dir, 1e-4f, i, dist0, base, ctau, in, st, sb, &status);
                if (status == 1) continue;
                for (uint32_t b = 0; b < B; b++) {
                    double E = (double)(p->source_flux[src * B + b] * ctau[b]) * (double)strength;
                    if (E == 0.0) continue;
                    double rho = p->rho[(size_t)i * B + b], tau = p->tau[(size_t)i * B + b];
                    size_t k = (size_t)i * B + b;
                    atomic_addd(&in[k], E * (1.0 - rho - tau));
                    atomic_addd(top ? &st[k] : &sb[k], E * rho);
                    atomic_addd(top ? &sb[k] : &st[k], E * tau);
                    if (diag) atomic_addd(&diag[2 * B + b], E * areas[i]);
                }
            }
        }
    }
    free(mats);
}

static void diffuse_pass(const orc_scene *s, const orc_band_params *p, uint32_t n, uint32_t seed,
                         const double *out_top, const double *out_bot, const float *dflux, const uint8_t *active,
                         double *in, double *st, double *sb, const uint8_t *cover) {
    uint32_t B = p->nbands, nn = n * n;
    float strength = 1.0f / (float)nn;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long ii = 0; ii < (long long)s->P; ii++) {
        uint32_t i = (uint32_t)ii;
        if (!traced_type(s->type[i])) continue;
        for (int fi = 0; fi < 2; fi++) {
            uint32_t face = fi == 0 ? 1u : 0u;
            if (face == 0 && s->twos[i] == 0) continue;
            float base[256], ctau[256];
            for (uint32_t r = 0; r < nn; r++) {
                v3 o, d;
                if (!diffuse_ray(s, i, r % n, r / n, n, face, seed, &o, &d)) continue;
                uint32_t h;
                if (cover) {
                    /* covers: the sky flux (0 in scattering passes) times the transmittance so far is what
                     * a crossed cover takes its shares of (diffuse_raygen.comp:536-557) */
                    for (uint32_t b = 0; b < B; b++) {
                        base[b] = active[b] ? dflux[b] * strength : 0.0f;
                        ctau[b] = 1.0f;
                    }
                    int status;
                    h = trace_covers_orc(s, p, cover, o, d, 1e-5f, i, -1.0f, base, ctau, in, st, sb, &status);
                } else {
                    int ex;
                    h = trace_loop(s, o, d, 1e-5f, i, &ex);
                }
                int top = 0;
                if (h != MISS) top = dot(s->normal[h], d) < 0.0f;
                for (uint32_t b = 0; b < B; b++) {
                    if (!active[b]) continue;
                    double E = 0.0;
                    if (h != MISS) {
                        /* a surface hit is not attenuated by the covers in front of it (diffuse_raygen.comp:593-598) */
                        double inc = top ? out_top[(size_t)h * B + b] : out_bot[(size_t)h * B + b];
                        E = inc * (double)strength;
                    } else if (dflux[b] > 0.f) {
                        /* fd = 1 for an isotropic sky, else the distribution in the ray's direction */
                        float fd = sky_weight_orc(p, b, d);
                        float sky = fd * dflux[b];
                        if (cover) sky = sky * ctau[b];
                        E = (double)sky * (double)strength;
                    }
                    if (E == 0.0) continue;
                    double rho = p->rho[(size_t)i * B + b], tau = p->tau[(size_t)i * B + b];
                    size_t k = (size_t)i * B + b;
                    if (cover) {
                        atomic_addd(&in[k], E * (1.0 - rho - tau));
                        atomic_addd(face == 1 ? &st[k] : &sb[k], E * rho);
                        atomic_addd(face == 1 ? &sb[k] : &st[k], E * tau);
                        continue;
                    }
                    in[k] += E * (1.0 - rho - tau);
                    if (face == 1) { st[k] += E * rho; sb[k] += E * tau; }
                    else { sb[k] += E * rho; st[k] += E * tau; }
                }
            }
        }
    }
}

int orc_run_band(const orc_scene *s, const orc_band_params *p, double *R, double *diag) {
    const uint32_t P = s->P, B = p->nbands, S = p->nsources;
    const double sigma = 5.6703744E-8; /* R/include/RadiationModel.h:2210 */
    size_t PB = (size_t)P * B;
    double *in = (double *)calloc(PB, sizeof(double));
    double *st = (double *)calloc(PB, sizeof(double));
    double *sb = (double *)calloc(PB, sizeof(double));
    double *ot = (double *)calloc(PB, sizeof(double));
    double *ob = (double *)calloc(PB, sizeof(double));
    double *ft = (double *)calloc(PB, sizeof(double));
    double *fb = (double *)calloc(PB, sizeof(double));
    float *dflux = (float *)calloc(B ? B : 1, sizeof(float));
    uint8_t *active = (uint8_t *)calloc(B ? B : 1, 1);
    double *areas = (double *)calloc(P ? P : 1, sizeof(double));
    orc_areas(s, areas);
    /* covers: glass in every launched band (direct_raygen.comp:629-637) */
    uint8_t *cover = NULL;
    if (p->is_glass && B <= 256) {
        int any = 0;
        cover = (uint8_t *)calloc(P ? P : 1, 1);
        for (uint32_t i = 0; i < P; i++) {
            int all = B > 0;
            for (uint32_t b = 0; b < B; b++) all = all && p->is_glass[(size_t)i * B + b];
            cover[i] = (uint8_t)all;
            any |= all;
        }
        if (!any) { free(cover); cover = NULL; }
    }
    if (diag) for (uint32_t k = 0; k < 3 * B; k++) diag[k] = 0.0;

    uint32_t maxdepth = 0;
    int scattering = 0, emission = 0, diffuse = 0, rundirect = 0;
    for (uint32_t b = 0; b < B; b++) {
        if (p->band_depth[b] > maxdepth) maxdepth = p->band_depth[b];
        if (p->band_depth[b] > 0) scattering = 1;                      /* :3686-3692 */
        if (p->emission[b]) emission = 1;                              /* :3772-3777 */
        if (p->diffuse_flux[b] > 0.f) diffuse = 1;                     /* :3714-3720 */
        for (uint32_t k = 0; k < S; k++) if (p->source_flux[k * B + b] > 0.f) rundirect = 1; /* :3818-3826 */
    }
    uint32_t d = (uint32_t)ceil(sqrt((double)p->direct_rays));
    uint32_t n = (uint32_t)ceil(sqrt((double)p->diffuse_rays));

    /* DIRECT (:3847-3982) */
    if (cover && S > 0 && rundirect && d > 0) {
        direct_glass(s, p, cover, in, st, sb, areas, diag);
    } else if (S > 0 && rundirect && d > 0) {
        double *sums = (double *)calloc((size_t)P * S * 2, sizeof(double));
        orc_direct_sums_src(s, d, p->direct_seed, S, p->source_dir, p->source_type, p->source_width, p->source_rotation,
                            sums);
        for (uint32_t i = 0; i < P; i++)
            for (uint32_t k = 0; k < S; k++) {
                double stp = sums[((size_t)i * S + k) * 2], sbt = sums[((size_t)i * S + k) * 2 + 1];
                for (uint32_t b = 0; b < B; b++) {
                    double F = p->source_flux[k * B + b];
                    double rho = p->rho[(size_t)i * B + b], tau = p->tau[(size_t)i * B + b];
                    size_t q = (size_t)i * B + b;
                    in[q] += F * (stp + sbt) * (1.0 - rho - tau);
                    st[q] += F * (stp * rho + sbt * tau);
                    sb[q] += F * (stp * tau + sbt * rho);
                    if (diag) diag[2 * B + b] += F * (stp + sbt) * areas[i];
                }
            }
        free(sums);
    }
    /* scattered direct energy -> radiation_out; one-sided primitives combine faces (:3999-4036) */
    if (scattering && rundirect) {
        for (uint32_t i = 0; i < P; i++)
            for (uint32_t b = 0; b < B; b++) {
                size_t q = (size_t)i * B + b;
                ft[q] = st[q]; fb[q] = sb[q];
                if (s->twos[i] == 0) { double tot = ft[q] + fb[q]; ft[q] = tot; fb[q] = tot; }
            }
        memcpy(ot, ft, PB * sizeof(double)); memcpy(ob, fb, PB * sizeof(double));
        memset(st, 0, PB * sizeof(double)); memset(sb, 0, PB * sizeof(double));
    }
    int primary = 0;
    if (emission || diffuse) {
        if (emission) { /* :4041-4137 */
            for (uint32_t b = 0; b < B; b++) {
                if (!p->emission[b]) continue;
                for (uint32_t i = 0; i < P; i++) {
                    float eps = p->eps ? p->eps[(size_t)i * B + b] : 1.0f;
                    if (p->band_depth[b] == 0 && eps != 1.f) eps = 1.f;
                    float T = p->temperature ? p->temperature[i] : 300.f;
                    if (T < 0) T = 300.f;
                    double out = sigma * (double)eps * pow((double)T, 4.0);
                    size_t q = (size_t)i * B + b;
                    ft[q] += out;
                    if (s->twos[i] != 0) fb[q] += out;
                }
            }
        }
        memcpy(ot, ft, PB * sizeof(double)); memcpy(ob, fb, PB * sizeof(double));
        for (uint32_t b = 0; b < B; b++) { dflux[b] = p->diffuse_flux[b]; active[b] = 1; }
        diffuse_pass(s, p, n, p->diffuse_seed[0], ot, ob, dflux, active, in, st, sb, cover);
        primary = 1;
        if (diag) for (uint32_t b = 0; b < B; b++) diag[1 * B + b] = 0.0;
    }
    if (scattering && (emission || diffuse) && !rundirect) { /* :4225-4234 */
        memcpy(ot, st, PB * sizeof(double)); memcpy(ob, sb, PB * sizeof(double));
        memset(st, 0, PB * sizeof(double)); memset(sb, 0, PB * sizeof(double));
    }
    if (scattering && (emission || diffuse || rundirect)) { /* :4236-4356 */
        for (uint32_t b = 0; b < B; b++) { dflux[b] = 0.f; active[b] = 1; }
        for (uint32_t sidx = 0; sidx < maxdepth; sidx++) {
            for (uint32_t b = 0; b < B; b++) {
                if (active[b] && sidx + 1 > p->band_depth[b]) {
                    active[b] = 0;
                    if (p->schedule == 1) { /* conserving: a retiring band deposits its leftover scatter */
                        for (uint32_t i = 0; i < P; i++) {
                            size_t q = (size_t)i * B + b;
                            in[q] += st[q] + sb[q]; st[q] = 0; sb[q] = 0;
                        }
                    }
                }
            }
            int copy = (sidx > 0) || (emission && rundirect);
            /* conserving: also hand the primary pass's scatter on when direct ran and emission is off
             * (without direct it was already moved to radiation_out after the primary pass) */
            if (p->schedule == 1) copy = (sidx > 0) || (primary && rundirect);
            if (copy) { memcpy(ot, st, PB * sizeof(double)); memcpy(ob, sb, PB * sizeof(double)); }
            memset(st, 0, PB * sizeof(double)); memset(sb, 0, PB * sizeof(double));
            diffuse_pass(s, p, n, p->diffuse_seed[1 + sidx], ot, ob, dflux, active, in, st, sb, cover);
        }
    }
    /* leftover scatter deposited as absorbed (:4631-4654) */
    for (size_t q = 0; q < PB; q++) R[q] = in[q] + st[q] + sb[q];
    if (diag) for (uint32_t i = 0; i < P; i++) for (uint32_t b = 0; b < B; b++) diag[b] += R[(size_t)i * B + b] * areas[i];
    free(in); free(st); free(sb); free(ot); free(ob); free(ft); free(fb); free(dflux); free(active); free(areas);
    free(cover);
    return 0;
}
