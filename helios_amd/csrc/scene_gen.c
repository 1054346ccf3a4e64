/*
 * scene_gen.c -- seeded synthetic canopies for the five workload configurations.
 *
 * Output is what the backend consumes: [P][16] row-major transforms, [P] types (0 patch,
 * 1 triangle), [P] raw twosided flags.  Every generator is integer-seeded and uses only IEEE
 * + - * / sqrt and a local sin/cos of 2*pi*r, so a scene is bit-identical on every x86-64 host.
 *
 * Transform conventions follow the reference's primitive construction:
 *   patch:    scale(sx, sy, 1) -> rotate(-elevation about x) -> rotate(-azimuth about z) -> translate
 *             (core/src/Context_primitive.cpp:36-76), canonical square (-0.5..0.5)^2;
 *   triangle: affine map of (0,0,0),(0,1,0),(1,1,0) onto (v0,v1,v2) (core/src/Context_primitive.cpp:1327).
 * Leaf normals follow a spherical (uniform) leaf-angle distribution: cos(elevation) = 1 - u,
 * azimuth = 2*pi*u' (R/tests/selfTest.cpp:1393).
 *
 * Every function returns the primitive count; with NULL output pointers it only counts.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { uint64_t s; } sg_rng;

static uint64_t sm64(sg_rng *r) {
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* uniform in [0,1) with 24 random bits (exact in binary32) */
static float uni(sg_rng *r) { return (float)(sm64(r) >> 40) * (1.0f / 16777216.0f); }

/* sin/cos of 2*pi*r: exact quadrant reduction + Taylor polynomials on [-pi/4, pi/4] */
static void sc2pi(float r, float *s, float *c) {
    float u = r * 4.0f;
    float q = floorf(u + 0.5f);
    float x = u - q;
    float a = x * 1.57079632679489662f;
    float a2 = a * a;
    float ps = a + a * a2 * (-0.166666666666666667f + a2 * (0.00833333333333333333f +
                                a2 * (-0.000198412698412698413f + a2 * 2.75573192239858907e-6f)));
    float pc = 1.0f + a2 * (-0.5f + a2 * (0.0416666666666666667f + a2 * (-0.00138888888888888889f +
                                a2 * (2.48015873015873016e-5f + a2 * -2.75573192239858907e-7f))));
    switch (((int)q) & 3) {
        case 0: *s = ps; *c = pc; break;
        case 1: *s = pc; *c = -ps; break;
        case 2: *s = -ps; *c = -pc; break;
        default: *s = -pc; *c = ps; break;
    }
}

typedef struct { float *T; uint32_t *type; uint32_t *twos; uint32_t n; } sg_out;

/* rotation R = Rz(-az) * Rx(-el) given cos/sin; normal = R*(0,0,1) */
static void rot(float ce, float se, float ca, float sa, float R[9]) {
    R[0] = ca; R[1] = sa * ce; R[2] = sa * se;
    R[3] = -sa; R[4] = ca * ce; R[5] = ca * se;
    R[6] = 0.f; R[7] = -se; R[8] = ce;
}

static void put_patch(sg_out *o, float cx, float cy, float cz, float sx, float sy, const float R[9], uint32_t twos) {
    if (o->T) {
        float *t = o->T + (size_t)o->n * 16;
        t[0] = R[0] * sx; t[1] = R[1] * sy; t[2] = R[2]; t[3] = cx;
        t[4] = R[3] * sx; t[5] = R[4] * sy; t[6] = R[5]; t[7] = cy;
        t[8] = R[6] * sx; t[9] = R[7] * sy; t[10] = R[8]; t[11] = cz;
        t[12] = 0.f; t[13] = 0.f; t[14] = 0.f; t[15] = 1.f;
        o->type[o->n] = 0;
        o->twos[o->n] = twos;
    }
    o->n++;
}

static void put_tri(sg_out *o, const float v0[3], const float v1[3], const float v2[3], uint32_t twos) {
    if (o->T) {
        float *t = o->T + (size_t)o->n * 16;
        float c0[3], c1[3], nz[3];
        for (int a = 0; a < 3; a++) { c0[a] = v2[a] - v1[a]; c1[a] = v1[a] - v0[a]; }
        nz[0] = c0[1] * c1[2] - c0[2] * c1[1];
        nz[1] = c0[2] * c1[0] - c0[0] * c1[2];
        nz[2] = c0[0] * c1[1] - c0[1] * c1[0];
        float m = sqrtf(nz[0] * nz[0] + nz[1] * nz[1] + nz[2] * nz[2]);
        if (m > 0.f) { nz[0] /= m; nz[1] /= m; nz[2] /= m; } else { nz[0] = 0.f; nz[1] = 0.f; nz[2] = 1.f; }
        t[0] = c0[0]; t[1] = c1[0]; t[2] = nz[0]; t[3] = v0[0];
        t[4] = c0[1]; t[5] = c1[1]; t[6] = nz[1]; t[7] = v0[1];
        t[8] = c0[2]; t[9] = c1[2]; t[10] = nz[2]; t[11] = v0[2];
        t[12] = 0.f; t[13] = 0.f; t[14] = 0.f; t[15] = 1.f;
        o->type[o->n] = 1;
        o->twos[o->n] = twos;
    }
    o->n++;
}

static void spherical_leaf(sg_rng *r, float R[9]) {
    float ce = 1.0f - uni(r);               /* cos(elevation) = 1 - u */
    float se = sqrtf(fmaxf(0.f, (1.0f - ce) * (1.0f + ce)));
    float sa, ca;
    sc2pi(uni(r), &sa, &ca);
    rot(ce, se, ca, sa, R);
}

static const float IDENT[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

/* isosceles leaf triangle of area A (base w, height w), centred on its centroid */
static void leaf_tri(sg_out *o, sg_rng *r, float cx, float cy, float cz, float A, uint32_t twos) {
    float R[9];
    spherical_leaf(r, R);
    float w = sqrtf(2.0f * A), h = w;
    float lv[3][3] = {{-0.5f * w, -h / 3.0f, 0.f}, {0.5f * w, -h / 3.0f, 0.f}, {0.f, 2.0f * h / 3.0f, 0.f}};
    float v[3][3];
    for (int k = 0; k < 3; k++) {
        v[k][0] = cx + (R[0] * lv[k][0] + R[1] * lv[k][1]);
        v[k][1] = cy + (R[3] * lv[k][0] + R[4] * lv[k][1]);
        v[k][2] = cz + (R[6] * lv[k][0] + R[7] * lv[k][1]);
    }
    put_tri(o, v[0], v[1], v[2], twos);
}

static void ground(sg_out *o, float x0, float x1, float y0, float y1, uint32_t nx, uint32_t ny) {
    float dx = (x1 - x0) / (float)nx, dy = (y1 - y0) / (float)ny;
    for (uint32_t j = 0; j < ny; j++)
        for (uint32_t i = 0; i < nx; i++)
            put_patch(o, x0 + ((float)i + 0.5f) * dx, y0 + ((float)j + 0.5f) * dy, 0.f, dx, dy, IDENT, 0);
}

/* Config 1: 1,600 square leaves of side 0.1 m, centres uniform in [-1,1]x[-1,1]x[0,1], spherical
 * leaf angles, two-sided; one 2x2 m one-sided ground patch at z = 0.  P = 1601. */
uint32_t sg_config1(uint32_t seed, float *T, uint32_t *type, uint32_t *twos) {
    sg_out o = {T, type, twos, 0};
    sg_rng r = {0x1234567ull ^ ((uint64_t)seed << 20)};
    for (int i = 0; i < 1600; i++) {
        float cx = -1.0f + 2.0f * uni(&r), cy = -1.0f + 2.0f * uni(&r), cz = uni(&r);
        float R[9];
        spherical_leaf(&r, R);
        put_patch(&o, cx, cy, cz, 0.1f, 0.1f, R, 1);
    }
    put_patch(&o, 0.f, 0.f, 0.f, 2.0f, 2.0f, IDENT, 0);
    return o.n;
}

/* Generic homogeneous canopy of square leaves (used by the Beer's-law checks):
 * nleaves leaves of side w in [-D/2,D/2]^2 x [w/2, w/2 + h], ground D x D split ng x ng.
 * inset != 0 keeps every leaf's horizontal extent inside [-D/2, D/2]^2, so that with periodic
 * boundaries the ground alone defines the periodic cell (no uncovered seam strip). */
uint32_t sg_patch_canopy(uint32_t seed, uint32_t nleaves, float w, float D, float h, uint32_t ng, uint32_t inset,
                         float *T, uint32_t *type, uint32_t *twos) {
    sg_out o = {T, type, twos, 0};
    sg_rng r = {0xBEEFull ^ ((uint64_t)seed << 24)};
    for (uint32_t i = 0; i < nleaves; i++) {
        float ux = uni(&r), uy = uni(&r), cz = 0.5f * w + uni(&r) * h;
        float R[9];
        spherical_leaf(&r, R);
        float ex = 0.f, ey = 0.f;
        if (inset) { /* half extents of the rotated square in x and y */
            ex = 0.5f * w * (fabsf(R[0]) + fabsf(R[1]));
            ey = 0.5f * w * (fabsf(R[3]) + fabsf(R[4]));
        }
        float cx = -0.5f * D + ex + ux * (D - 2.0f * ex), cy = -0.5f * D + ey + uy * (D - 2.0f * ey);
        put_patch(&o, cx, cy, cz, w, w, R, 1);
    }
    if (ng) ground(&o, -0.5f * D, 0.5f * D, -0.5f * D, 0.5f * D, ng, ng);
    return o.n;
}

/* Config 2: 200,000 isosceles leaf triangles, area log-uniform in [4, 16] cm^2, centres uniform in
 * [-5,5]^2 x [0,2] m, spherical leaf angles, two-sided; 10 x 10 m ground of 20 x 20 one-sided
 * patches.  Periodic in x and y.  P = 200,400.  (Realised LAI = sum(area)/100 m^2 ~ 1.73.) */
uint32_t sg_config2(uint32_t seed, uint32_t nleaves, float *T, uint32_t *type, uint32_t *twos) {
    sg_out o = {T, type, twos, 0};
    sg_rng r = {0xC0FFEEull ^ ((uint64_t)seed << 16)};
    for (uint32_t i = 0; i < nleaves; i++) {
        float cx = -5.0f + 10.0f * uni(&r), cy = -5.0f + 10.0f * uni(&r), cz = 2.0f * uni(&r);
        /* log-uniform area in [4e-4, 16e-4] m^2: 4e-4 * 4^u = 4e-4 * 2^(2u), 2^(2u) by 24 square roots */
        float u = uni(&r);
        float f = 2.0f * u;            /* exponent in [0, 2) */
        int ip = (int)f;               /* integer part */
        float fr = f - (float)ip;
        float m = 1.0f, root = 2.0f;
        for (int b = 0; b < 24; b++) { /* 2^fr via binary expansion of fr and repeated square roots */
            root = sqrtf(root);
            fr *= 2.0f;
            if (fr >= 1.0f) { m *= root; fr -= 1.0f; }
        }
        float e2 = (float)(1 << ip) * m;
        leaf_tri(&o, &r, cx, cy, cz, 4e-4f * e2, 1);
    }
    ground(&o, -5.f, 5.f, -5.f, 5.f, 20, 20);
    return o.n;
}

/* 16-gon tube segment between rings (c0, r0) and (c1, r1) along axis a (unit), one-sided outward. */
static const float RING_C[16] = {1.0f, 0.923879532511f, 0.707106781187f, 0.382683432365f, 0.0f, -0.382683432365f,
                                 -0.707106781187f, -0.923879532511f, -1.0f, -0.923879532511f, -0.707106781187f,
                                 -0.382683432365f, 0.0f, 0.382683432365f, 0.707106781187f, 0.923879532511f};
static const float RING_S[16] = {0.0f, 0.382683432365f, 0.707106781187f, 0.923879532511f, 1.0f, 0.923879532511f,
                                 0.707106781187f, 0.382683432365f, 0.0f, -0.382683432365f, -0.707106781187f,
                                 -0.923879532511f, -1.0f, -0.923879532511f, -0.707106781187f, -0.382683432365f};

static void tube_n(sg_out *o, const float c0[3], const float c1[3], float r0, float r1, int ntri);

static void tube(sg_out *o, const float c0[3], const float c1[3], float r0, float r1) { tube_n(o, c0, c1, r0, r1, 32); }

/* first ntri (<= 32) triangles of a 16-gon tube segment */
static void tube_n(sg_out *o, const float c0[3], const float c1[3], float r0, float r1, int ntri) {
    float ax[3] = {c1[0] - c0[0], c1[1] - c0[1], c1[2] - c0[2]};
    float L = sqrtf(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    for (int a = 0; a < 3; a++) ax[a] /= L;
    /* perpendicular basis u, v */
    float h[3] = {fabsf(ax[2]) < 0.9f ? 0.f : 1.f, 0.f, fabsf(ax[2]) < 0.9f ? 1.f : 0.f};
    float u[3] = {h[1] * ax[2] - h[2] * ax[1], h[2] * ax[0] - h[0] * ax[2], h[0] * ax[1] - h[1] * ax[0]};
    float um = sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int a = 0; a < 3; a++) u[a] /= um;
    float v[3] = {ax[1] * u[2] - ax[2] * u[1], ax[2] * u[0] - ax[0] * u[2], ax[0] * u[1] - ax[1] * u[0]};
    for (int k = 0; k < 16; k++) {
        int k2 = (k + 1) & 15;
        float p00[3], p01[3], p10[3], p11[3];
        for (int a = 0; a < 3; a++) {
            p00[a] = c0[a] + r0 * (RING_C[k] * u[a] + RING_S[k] * v[a]);
            p01[a] = c0[a] + r0 * (RING_C[k2] * u[a] + RING_S[k2] * v[a]);
            p10[a] = c1[a] + r1 * (RING_C[k] * u[a] + RING_S[k] * v[a]);
            p11[a] = c1[a] + r1 * (RING_C[k2] * u[a] + RING_S[k2] * v[a]);
        }
        /* outward normal = cross(v1 - v0, v2 - v0) */
        if (2 * k < ntri) put_tri(o, p00, p01, p11, 0);
        if (2 * k + 1 < ntri) put_tri(o, p00, p11, p10, 0);
    }
}

/* One tree at (tx, ty): trunk + branches as 16-gon tubes (radii 0.15 -> 0.02 m, exactly ntube_tris
 * tube triangles: whole 32-triangle ring segments plus one partial segment) and nleaf leaf
 * triangles (area uniform 6-14 cm^2) in an ellipsoidal crown. */
static void tree(sg_out *o, sg_rng *r, float tx, float ty, uint32_t nleaf, uint32_t ntube_tris, float jitter) {
    uint32_t segs = ntube_tris / 32;
    uint32_t rest = ntube_tris % 32;
    uint32_t trunk = segs < 12 ? segs : 12;
    float base[3] = {tx, ty, 0.f};
    float top = 1.6f;
    for (uint32_t s = 0; s < trunk; s++) {
        float z0 = top * (float)s / (float)trunk, z1 = top * (float)(s + 1) / (float)trunk;
        float c0[3] = {base[0], base[1], z0}, c1[3] = {base[0], base[1], z1};
        float r0 = 0.15f - 0.07f * (float)s / (float)trunk, r1 = 0.15f - 0.07f * (float)(s + 1) / (float)trunk;
        tube(o, c0, c1, r0, r1);
    }
    uint32_t left = segs - trunk;
    uint32_t per_branch = 5;
    uint32_t b = 0;
    while (left > 0) {
        uint32_t ns = left < per_branch ? left : per_branch;
        float az_s, az_c;
        sc2pi(((float)b + 0.5f * uni(r)) / 10.0f, &az_s, &az_c);
        float el = 0.35f + 0.4f * uni(r); /* tan of branch elevation-ish */
        float start[3] = {base[0], base[1], top * (0.7f + 0.3f * uni(r))};
        float len = 1.1f + 0.4f * uni(r);
        float dir[3] = {az_c, az_s, el};
        float dm = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        for (int a = 0; a < 3; a++) dir[a] /= dm;
        for (uint32_t s = 0; s < ns; s++) {
            float f0 = (float)s / (float)ns, f1 = (float)(s + 1) / (float)ns;
            float c0[3], c1[3];
            for (int a = 0; a < 3; a++) { c0[a] = start[a] + dir[a] * len * f0; c1[a] = start[a] + dir[a] * len * f1; }
            tube(o, c0, c1, 0.07f - 0.05f * f0, 0.07f - 0.05f * f1);
        }
        left -= ns;
        b++;
    }
    if (rest) { /* partial collar segment at the trunk top */
        float c0[3] = {base[0], base[1], top}, c1[3] = {base[0], base[1], top + 0.05f};
        tube_n(o, c0, c1, 0.08f, 0.07f, (int)rest);
    }
    /* crown: ellipsoid centred at (tx, ty, 2.4), radii (1.5, 1.5, 1.1) */
    for (uint32_t i = 0; i < nleaf; i++) {
        float x, y, z;
        do {
            x = -1.0f + 2.0f * uni(r); y = -1.0f + 2.0f * uni(r); z = -1.0f + 2.0f * uni(r);
        } while (x * x + y * y + z * z > 1.0f);
        float jx = jitter * (uni(r) - 0.5f), jy = jitter * (uni(r) - 0.5f);
        float u = uni(r);
        float A = 6e-4f + 8e-4f * u;
        leaf_tri(o, r, tx + 1.5f * x + jx, ty + 1.5f * y + jy, 2.4f + 1.1f * z, A, 1);
    }
}

/* Config 3: orchard of nx*ny trees on a 4 m grid, each 2,000 tube triangles + leaves_per_tree leaf
 * triangles (38,000 at full size), over a one-sided ground of 40 x 40 patches (20 x 20 m).
 * Full size: 25 trees -> 1,000,000 triangles + 1,600 ground patches. */
uint32_t sg_orchard(uint32_t seed, uint32_t nx, uint32_t ny, uint32_t leaves_per_tree, uint32_t tube_tris,
                    uint32_t ntiles, float jitter, float *T, uint32_t *type, uint32_t *twos) {
    sg_out o = {T, type, twos, 0};
    float W = 4.0f * (float)nx, H = 4.0f * (float)ny;
    for (uint32_t tile = 0; tile < (ntiles ? ntiles : 1); tile++) {
        sg_rng r = {0x0C4A7Dull ^ ((uint64_t)seed << 12) ^ ((uint64_t)tile << 40)};
        float ox = (float)tile * W;
        for (uint32_t j = 0; j < ny; j++)
            for (uint32_t i = 0; i < nx; i++) {
                float tx = ox - 0.5f * W + 2.0f + 4.0f * (float)i, ty = -0.5f * H + 2.0f + 4.0f * (float)j;
                tree(&o, &r, tx, ty, leaves_per_tree, tube_tris, tile ? jitter : 0.f);
            }
        ground(&o, ox - 0.5f * W, ox + 0.5f * W, -0.5f * H, 0.5f * H, 8 * nx, 8 * ny);
    }
    return o.n;
}

/* Config 4 request scenes: P primitives, mixed patches / triangles (every other), leaf area
 * ~ 10 cm^2, domain sized for LAI ~ 2, spherical angles, plus a 10 x 10 one-sided ground. */
uint32_t sg_random_canopy(uint32_t seed, uint32_t P, float *T, uint32_t *type, uint32_t *twos) {
    sg_out o = {T, type, twos, 0};
    sg_rng r = {0x5EEDull ^ ((uint64_t)seed << 8)};
    uint32_t nleaf = P > 100 ? P - 100 : P;
    float A = 1e-3f;
    float D = sqrtf((float)nleaf * A / 2.0f);
    for (uint32_t i = 0; i < nleaf; i++) {
        float cx = (-0.5f + uni(&r)) * D, cy = (-0.5f + uni(&r)) * D, cz = 0.05f + 2.0f * uni(&r);
        if (i & 1) leaf_tri(&o, &r, cx, cy, cz, A, 1);
        else {
            float R[9];
            spherical_leaf(&r, R);
            float s = sqrtf(A);
            put_patch(&o, cx, cy, cz, s, s, R, 1);
        }
    }
    if (P > 100) ground(&o, -0.5f * D, 0.5f * D, -0.5f * D, 0.5f * D, 10, 10);
    return o.n;
}

/* Disk of radius rad as the Context builds one (core/src/Context.cpp:2503-2537, one ring): ndiv triangles
 * (centre, p_t, p_t+1) with p_t = rad (cos, sin)(2 pi t / ndiv) in the disk's plane, oriented by a
 * spherical leaf rotation and translated to the centre. */
static void disk_fan(sg_out *o, sg_rng *r, float cx, float cy, float cz, float rad, uint32_t ndiv, uint32_t twos) {
    float R[9];
    spherical_leaf(r, R);
    float c[3] = {cx, cy, cz};
    for (uint32_t t = 0; t < ndiv; t++) {
        float s0, c0, s1, c1;
        sc2pi((float)t / (float)ndiv, &s0, &c0);
        sc2pi((float)(t + 1) / (float)ndiv, &s1, &c1);
        float l[2][2] = {{rad * c0, rad * s0}, {rad * c1, rad * s1}};
        float v[2][3];
        for (int k = 0; k < 2; k++) {
            v[k][0] = cx + (R[0] * l[k][0] + R[1] * l[k][1]);
            v[k][1] = cy + (R[3] * l[k][0] + R[4] * l[k][1]);
            v[k][2] = cz + (R[6] * l[k][0] + R[7] * l[k][1]);
        }
        put_tri(o, c, v[0], v[1], twos);
    }
}

/* Mixed canopy of about P primitives (config 4): leaves of area 1e-3 m^2 cycling through square patches,
 * isosceles triangles and 6-triangle disks, centres uniform over a square of LAI ~2 between 0.05 and 2.05 m,
 * spherical leaf angles, over a 10 x 10 ground of one-sided patches. */
uint32_t sg_mixed_canopy(uint32_t seed, uint32_t P, float *T, uint32_t *type, uint32_t *twos) {
    sg_out o = {T, type, twos, 0};
    sg_rng r = {0xD15Cull ^ ((uint64_t)seed << 8)};
    const uint32_t target = P > 100 ? P - 100 : P;
    float A = 1e-3f;
    /* three leaves (patch, triangle, disk) take 1 + 1 + 6 primitives */
    uint32_t nleaf = 3 * (target / 8);
    if (nleaf == 0) nleaf = 1;
    float D = sqrtf((float)nleaf * A / 2.0f);
    float rad = sqrtf(A / 3.14159265f);
    for (uint32_t i = 0; i < nleaf; i++) {
        float cx = (-0.5f + uni(&r)) * D, cy = (-0.5f + uni(&r)) * D, cz = 0.05f + 2.0f * uni(&r);
        if (i % 3 == 0) {
            float R[9];
            spherical_leaf(&r, R);
            float sd = sqrtf(A);
            put_patch(&o, cx, cy, cz, sd, sd, R, 1);
        } else if (i % 3 == 1) {
            leaf_tri(&o, &r, cx, cy, cz, A, 1);
        } else {
            disk_fan(&o, &r, cx, cy, cz, rad, 6, 1);
        }
    }
    if (P > 100) ground(&o, -0.5f * D, 0.5f * D, -0.5f * D, 0.5f * D, 10, 10);
    return o.n;
}

/* Random rays: origins uniform in the box [lo, hi], directions uniform on the sphere
 * (z = 1 - 2u, phi = 2*pi*u'). */
void sg_rays(uint32_t seed, uint64_t n, const float lo[3], const float hi[3], float *O, float *D) {
    sg_rng r = {0xA5A5ull ^ ((uint64_t)seed << 32)};
    for (uint64_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) O[i * 3 + a] = lo[a] + (hi[a] - lo[a]) * uni(&r);
        float z = 1.0f - 2.0f * uni(&r);
        float s, c;
        sc2pi(uni(&r), &s, &c);
        float rr = sqrtf(fmaxf(0.f, 1.0f - z * z));
        D[i * 3] = rr * c; D[i * 3 + 1] = rr * s; D[i * 3 + 2] = z;
    }
}

/* sin/cos of 2*pi*r for callers that need the same function (sun positions). */
void sg_sincos_2pi(float r, float *s, float *c) { sc2pi(r, s, c); }
