// rt_scene.hip -- geometry preparation and Morton-code LBVH build on the GPU (piece a).
//
// Replaces the reference's single-threaded CPU SAH builders (R/src/backends/BVHBuilder.cpp:24-766,
// CD/src/CollisionDetection.cpp:340-513, 940-1216) with a data-parallel build:
//   k_prep      world vertices, normals, diffuse frame, predicate records, raw AABBs
//               (R/src/backends/VulkanComputeBackend.cpp:327-435)
//   k_bounds    centroid / vertex bounds (exact float min/max through order-preserving integer atomics)
//   k_morton    63-bit Morton codes of the centroids (21 bits per axis)
//   radix sort  hipcub DeviceRadixSort on (code, index)
//   k_karras    Karras (HPG 2012) binary radix tree over the sorted codes
//   k_refit     bottom-up AABB refit; the second child to arrive at a node computes its box
//   k_emit      64-byte BVH2 nodes holding both child boxes, primitive records in leaf order
// The tree shape is not a parity target: the kernels use conservative boxes and a canonical
// closest-hit tie rule, so any valid tree gives the same hits.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rt_internal.h"

namespace rt {

__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float o2f(uint32_t u) {
    uint32_t v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &v, 4);
    return f;
}

__device__ __forceinline__ bool is_patch_t(uint32_t t) { return t == 0 || t == 3; }
__device__ __forceinline__ bool traced_t(uint32_t t) { return t == 0 || t == 1 || t == 3; }

// bounds layout (uint, order-preserving): [0..2] centroid min, [3..5] centroid max,
// [6..8] vertex min, [9..11] vertex max, [12] traced count
__global__ void k_bounds_init(uint32_t *b) {
    int i = threadIdx.x;
    if (i < 3) { b[i] = 0xFFFFFFFFu; b[6 + i] = 0xFFFFFFFFu; b[3 + i] = 0u; b[9 + i] = 0u; }
    if (i == 0) b[12] = 0u;
}

__global__ void k_prep(uint32_t P, const float *__restrict__ T, const uint32_t *__restrict__ type,
                       const uint32_t *__restrict__ twos, float4 *__restrict__ rec_u, float4 *__restrict__ nrm,
                       float4 *__restrict__ frame, float *__restrict__ box /*[P][6]*/, uint32_t *__restrict__ bounds) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float *t = T + (size_t)i * 16;
    uint32_t ty = type[i];
    bool patch = is_patch_t(ty);
    v3f v[4];
    if (patch) {
        v[0] = xform3(t, -0.5f, -0.5f, 0.f);
        v[1] = xform3(t, 0.5f, -0.5f, 0.f);
        v[2] = xform3(t, 0.5f, 0.5f, 0.f);
        v[3] = xform3(t, -0.5f, 0.5f, 0.f);
    } else if (ty == 1) {
        v[0] = xform3(t, 0.f, 0.f, 0.f);
        v[1] = xform3(t, 0.f, 1.f, 0.f);
        v[2] = xform3(t, 1.f, 1.f, 0.f);
        v[3] = v[2];
    } else {
        v[0] = v[1] = v[2] = v[3] = xform3(t, 0.f, 0.f, 0.f);
    }
    // precomputed world normal (VulkanComputeBackend.cpp:377-435)
    v3f w0, w1, w2, n;
    if (patch || ty == 2) { w0 = xform3(t, 0.f, 0.f, 0.f); w1 = xform3(t, 1.f, 0.f, 0.f); w2 = xform3(t, 0.f, 1.f, 0.f); }
    else { w0 = xform3(t, 0.f, 0.f, 0.f); w1 = xform3(t, 0.f, 1.f, 0.f); w2 = xform3(t, 1.f, 1.f, 0.f); }
    if (ty <= 3) {
        n = cross3(sub3(w1, w0), sub3(w2, w0));
        float len = rsqrt_cr(n.x * n.x + n.y * n.y + n.z * n.z);
        if (len > 1e-8f) { float inv = rdiv(1.0f, len); n = mul3(n, inv); }
        else n = mk3(0.f, 0.f, 1.f);
    } else {
        n = mk3(0.f, 0.f, 1.f);
    }
    nrm[i] = make_float4(n.x, n.y, n.z, __uint_as_float(twos[i]));
    // diffuse frame: cos/sin of acos(n.z) and atan2(n.y, n.x) from exact identities
    float ce = fminf(fmaxf(n.z, -1.f), 1.f);
    float se = rsqrt_cr((1.0f - ce) * (1.0f + ce));
    float r = rsqrt_cr(n.x * n.x + n.y * n.y);
    float ca, sa;
    if (r > 0.f) { ca = rdiv(n.x, r); sa = rdiv(n.y, r); }
    else { ca = signbit(n.x) ? -1.f : 1.f; sa = 0.f; }
    frame[i] = make_float4(ce, se, ca, sa);
    // predicate record
    v3f e1 = sub3(v[1], v[0]);
    v3f e2 = patch ? sub3(v[3], v[0]) : sub3(v[2], v[0]);
    v3f pn = patch ? nrm3(cross3(sub3(v[1], v[0]), sub3(v[2], v[0]))) : mk3(0.f, 0.f, 0.f);
    float aa = patch ? dot3(e1, e1) : 0.f, bb = patch ? dot3(e2, e2) : 0.f;
    uint32_t tag = i | (patch ? RT_PATCH_BIT : 0u);
    float4 *rr = rec_u + (size_t)i * 4;
    rr[0] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(tag));
    rr[1] = make_float4(e1.x, e1.y, e1.z, aa);
    rr[2] = make_float4(e2.x, e2.y, e2.z, bb);
    rr[3] = make_float4(pn.x, pn.y, pn.z, 0.f);
    // raw AABB (inflated later) and bounds
    int nv = patch ? 4 : 3;
    float lo[3] = {v[0].x, v[0].y, v[0].z}, hi[3] = {v[0].x, v[0].y, v[0].z};
    for (int k = 1; k < nv; k++) {
        float c[3] = {v[k].x, v[k].y, v[k].z};
        for (int a = 0; a < 3; a++) { lo[a] = fminf(lo[a], c[a]); hi[a] = fmaxf(hi[a], c[a]); }
    }
    float *bx = box + (size_t)i * 6;
    for (int a = 0; a < 3; a++) { bx[a] = lo[a]; bx[3 + a] = hi[a]; }
    for (int a = 0; a < 3; a++) {
        atomicMin(&bounds[6 + a], f2o(lo[a]));
        atomicMax(&bounds[9 + a], f2o(hi[a]));
    }
    if (traced_t(ty)) {
        for (int a = 0; a < 3; a++) {
            float c = (lo[a] + hi[a]) * 0.5f;
            atomicMin(&bounds[a], f2o(c));
            atomicMax(&bounds[3 + a], f2o(c));
        }
        atomicAdd(&bounds[12], 1u);
    }
}

__device__ __forceinline__ uint64_t expand21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(uint32_t P, const uint32_t *__restrict__ type, const float *__restrict__ box,
                         const uint32_t *__restrict__ bounds, uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    vals[i] = i;
    if (!traced_t(type[i])) { keys[i] = ~0ull; return; }
    const float *bx = box + (size_t)i * 6;
    uint64_t code = 0;
    for (int a = 0; a < 3; a++) {
        float cmin = o2f(bounds[a]), cmax = o2f(bounds[3 + a]);
        float c = (bx[a] + bx[3 + a]) * 0.5f;
        float e = cmax - cmin;
        float q = e > 0.f ? (c - cmin) / e * 2097151.0f : 0.f;
        q = fminf(fmaxf(q, 0.f), 2097151.0f);
        code |= expand21((uint64_t)q) << (2 - a);
    }
    keys[i] = code;
}

// sorted leaves: records in leaf order and inflated leaf boxes
__global__ void k_leaves(uint32_t nleaf, const uint32_t *__restrict__ vals, const float4 *__restrict__ rec_u,
                         const float *__restrict__ box, const uint32_t *__restrict__ bounds, float4 *__restrict__ rec,
                         float *__restrict__ lbox) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nleaf) return;
    uint32_t i = vals[k];
    for (int q = 0; q < 4; q++) rec[(size_t)k * 4 + q] = rec_u[(size_t)i * 4 + q];
    const float *bx = box + (size_t)i * 6;
    float S = 0.f;
    for (int a = 0; a < 3; a++) S = fmaxf(S, fmaxf(fabsf(o2f(bounds[6 + a])), fabsf(o2f(bounds[9 + a]))));
    float ext = fmaxf(fmaxf(bx[3] - bx[0], bx[4] - bx[1]), bx[5] - bx[2]);
    float pad = ext * 1e-4f + S * 4e-6f + 1e-30f;
    float *lb = lbox + (size_t)k * 6;
    for (int a = 0; a < 3; a++) { lb[a] = bx[a] - pad; lb[3 + a] = bx[3 + a] + pad; }
}

__device__ __forceinline__ int kdelta(const uint64_t *keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll((long long)(a ^ b));
}

// Karras 2012: internal node i covers a sorted key range; children are leaves (RT_LEAF_BIT | k) or internals.
__global__ void k_karras(int n, const uint64_t *__restrict__ keys, uint32_t *__restrict__ child /*[n-1][2]*/,
                         uint32_t *__restrict__ parent /*[n-1]*/, uint32_t *__restrict__ lparent /*[n]*/) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(keys, n, i, i - d);
    int lmax = 2;
    while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        int ns = s + t;
        if (ns < l && kdelta(keys, n, i, i + ns * d) > dnode) s = ns;
    } while (t > 1);
    int g = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    uint32_t c0, c1;
    if (lo == g) { c0 = RT_LEAF_BIT | (uint32_t)g; lparent[g] = i; }
    else { c0 = (uint32_t)g; parent[g] = i; }
    if (hi == g + 1) { c1 = RT_LEAF_BIT | (uint32_t)(g + 1); lparent[g + 1] = i; }
    else { c1 = (uint32_t)(g + 1); parent[g + 1] = i; }
    child[2 * i] = c0;
    child[2 * i + 1] = c1;
    if (i == 0) parent[0] = 0xFFFFFFFFu;
}

__device__ __forceinline__ void child_box(uint32_t c, const float *lbox, const float *ibox, float b[6]) {
    const float *src = (c & RT_LEAF_BIT) ? lbox + (size_t)(c & ~RT_LEAF_BIT) * 6 : ibox + (size_t)c * 6;
    for (int q = 0; q < 6; q++) b[q] = __builtin_nontemporal_load(src + q);
}

__global__ void k_refit(int n, const uint32_t *__restrict__ child, const uint32_t *__restrict__ parent,
                        const uint32_t *__restrict__ lparent, const float *__restrict__ lbox, float *ibox,
                        uint32_t *visit, uint32_t *depth_max) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t p = lparent[k];
    if (p >= (uint32_t)(n - 1)) { atomicMax(depth_max, 0x7fffffffu); return; }
    uint32_t depth = 1;
    // depth of this leaf (bounded walk: a valid tree over n leaves is shallower than n)
    for (uint32_t q = p; parent[q] != 0xFFFFFFFFu; q = parent[q]) {
        depth++;
        if (parent[q] >= (uint32_t)(n - 1) || depth > 4096) { atomicMax(depth_max, 0x7fffffffu); return; }
    }
    atomicMax(depth_max, depth);
    if (depth > RT_STACK) return;  // the build is rejected on the host; skip the refit
    while (true) {
        __threadfence();
        uint32_t prev = atomicAdd(&visit[p], 1u);
        __threadfence();
        if (prev == 0) return;  // sibling subtree not finished; it will continue upwards
        float b0[6], b1[6];
        child_box(child[2 * p], lbox, ibox, b0);
        child_box(child[2 * p + 1], lbox, ibox, b1);
        float *dst = ibox + (size_t)p * 6;
        for (int a = 0; a < 3; a++) {
            __builtin_nontemporal_store(fminf(b0[a], b1[a]), dst + a);
            __builtin_nontemporal_store(fmaxf(b0[3 + a], b1[3 + a]), dst + 3 + a);
        }
        if (p == 0) return;
        p = parent[p];
    }
}

// node layout: (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y), (c0.lo.z, c0.hi.z, c1.lo.x, c1.hi.x),
//              (c1.lo.y, c1.hi.y, c1.lo.z, c1.hi.z), (c0 link, c1 link, 0, 0)
__global__ void k_emit(int n, const uint32_t *__restrict__ child, const float *__restrict__ lbox,
                       const float *__restrict__ ibox, float4 *__restrict__ node) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    uint32_t c0 = child[2 * i], c1 = child[2 * i + 1];
    float a[6], b[6];
    child_box(c0, lbox, ibox, a);
    child_box(c1, lbox, ibox, b);
    float4 *o = node + (size_t)i * 4;
    o[0] = make_float4(a[0], a[3], a[1], a[4]);
    o[1] = make_float4(a[2], a[5], b[0], b[3]);
    o[2] = make_float4(b[1], b[4], b[2], b[5]);
    o[3] = make_float4(__uint_as_float(c0), __uint_as_float(c1), 0.f, 0.f);
}

#define CK(x)                                                                    \
    do {                                                                         \
        hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess) {                                                  \
            set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #x); \
            return -1;                                                           \
        }                                                                        \
    } while (0)

static inline uint32_t blocks(size_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// Periodic faces from the domain vertex bounds (R/src/RadiationModel.cpp:6046-6105) and the
// domain extents used for wrapping (R/src/backends/VulkanComputeBackend.cpp:565-577).
static void make_periodic(DevScene &h, const float vmin[3], const float vmax[3]) {
    h.nbbox = 0;
    h.dom[0] = h.dom[1] = h.dom[2] = h.dom[3] = 0.f;
    if (!h.px && !h.py) return;
    float xb[2] = {(float)((double)vmin[0] - 1e-5), (float)((double)vmax[0] + 1e-5)};
    float yb[2] = {(float)((double)vmin[1] - 1e-5), (float)((double)vmax[1] + 1e-5)};
    float zb[2] = {(float)((double)vmin[2] - 1e-5), (float)((double)vmax[2] + 1e-5)};
    v3f f[4][4];
    uint32_t k = 0;
    if (h.px) {
        f[k][0] = mk3(xb[0], yb[0], zb[0]); f[k][1] = mk3(xb[0], yb[1], zb[0]);
        f[k][2] = mk3(xb[0], yb[1], zb[1]); f[k][3] = mk3(xb[0], yb[0], zb[1]); k++;
        f[k][0] = mk3(xb[1], yb[0], zb[0]); f[k][1] = mk3(xb[1], yb[1], zb[0]);
        f[k][2] = mk3(xb[1], yb[1], zb[1]); f[k][3] = mk3(xb[1], yb[0], zb[1]); k++;
    }
    if (h.py) {
        f[k][0] = mk3(xb[0], yb[0], zb[0]); f[k][1] = mk3(xb[1], yb[0], zb[0]);
        f[k][2] = mk3(xb[1], yb[0], zb[1]); f[k][3] = mk3(xb[0], yb[0], zb[1]); k++;
        f[k][0] = mk3(xb[0], yb[1], zb[0]); f[k][1] = mk3(xb[1], yb[1], zb[0]);
        f[k][2] = mk3(xb[1], yb[1], zb[1]); f[k][3] = mk3(xb[0], yb[1], zb[1]); k++;
    }
    h.nbbox = k;
    float x0 = 1e30f, x1 = -1e30f, y0 = 1e30f, y1 = -1e30f;
    for (uint32_t q = 0; q < k; q++) {
        for (int j = 0; j < 4; j++) {
            x0 = std::min(x0, f[q][j].x); x1 = std::max(x1, f[q][j].x);
            y0 = std::min(y0, f[q][j].y); y1 = std::max(y1, f[q][j].y);
        }
        v3f a = sub3(f[q][1], f[q][0]), b = sub3(f[q][3], f[q][0]);
        v3f nn = nrm3(cross3(sub3(f[q][1], f[q][0]), sub3(f[q][2], f[q][0])));
        BBoxFace &bf = h.bbox[q];
        bf.v0[0] = f[q][0].x; bf.v0[1] = f[q][0].y; bf.v0[2] = f[q][0].z;
        bf.a[0] = a.x; bf.a[1] = a.y; bf.a[2] = a.z;
        bf.b[0] = b.x; bf.b[1] = b.y; bf.b[2] = b.z;
        bf.n[0] = nn.x; bf.n[1] = nn.y; bf.n[2] = nn.z;
        bf.aa = dot3(a, a);
        bf.bb = dot3(b, b);
    }
    h.dom[0] = x0; h.dom[1] = x1; h.dom[2] = y0; h.dom[3] = y1;
}

int scene_build(Scene &S, uint32_t P, const float *T, const uint32_t *type, const uint32_t *uuid,
                const uint32_t *twos, int px, int py, hipStream_t st) {
    scene_free(S);
    memset(&S.h, 0, sizeof(S.h));
    S.h.P = P;
    S.h.px = px;
    S.h.py = py;
    if (P == 0) {
        CK(hipMalloc(&S.d, sizeof(DevScene)));
        CK(hipMemcpyAsync(S.d, &S.h, sizeof(DevScene), hipMemcpyHostToDevice, st));
        CK(hipStreamSynchronize(st));
        return 0;
    }
    float *dT;
    uint32_t *dtype, *duuid, *dtwos;
    float4 *nrm, *frame;
    CK(hipMalloc(&dT, sizeof(float) * 16 * (size_t)P));
    S.own.push_back(dT);
    CK(hipMalloc(&dtype, sizeof(uint32_t) * P));
    S.own.push_back(dtype);
    CK(hipMalloc(&duuid, sizeof(uint32_t) * P));
    S.own.push_back(duuid);
    CK(hipMalloc(&dtwos, sizeof(uint32_t) * P));
    S.own.push_back(dtwos);
    CK(hipMalloc(&nrm, sizeof(float4) * P));
    S.own.push_back(nrm);
    CK(hipMalloc(&frame, sizeof(float4) * P));
    S.own.push_back(frame);
    CK(hipMemcpyAsync(dT, T, sizeof(float) * 16 * (size_t)P, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(dtype, type, sizeof(uint32_t) * P, hipMemcpyHostToDevice, st));
    std::vector<uint32_t> uu;
    if (!uuid) { uu.resize(P); for (uint32_t i = 0; i < P; i++) uu[i] = i; uuid = uu.data(); }
    CK(hipMemcpyAsync(duuid, uuid, sizeof(uint32_t) * P, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(dtwos, twos, sizeof(uint32_t) * P, hipMemcpyHostToDevice, st));

    // temporaries
    float4 *rec_u;
    float *box;
    uint32_t *bounds;
    uint64_t *keys, *keys2;
    uint32_t *vals, *vals2;
    CK(hipMalloc(&rec_u, sizeof(float4) * 4 * (size_t)P));
    CK(hipMalloc(&box, sizeof(float) * 6 * (size_t)P));
    CK(hipMalloc(&bounds, sizeof(uint32_t) * 16));
    CK(hipMalloc(&keys, sizeof(uint64_t) * P));
    CK(hipMalloc(&keys2, sizeof(uint64_t) * P));
    CK(hipMalloc(&vals, sizeof(uint32_t) * P));
    CK(hipMalloc(&vals2, sizeof(uint32_t) * P));

    k_bounds_init<<<1, 64, 0, st>>>(bounds);
    k_prep<<<blocks(P, 256), 256, 0, st>>>(P, dT, dtype, dtwos, rec_u, nrm, frame, box, bounds);
    k_morton<<<blocks(P, 256), 256, 0, st>>>(P, dtype, box, bounds, keys, vals);
    size_t tmp_bytes = 0;
    CK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys2, vals, vals2, (int)P, 0, 64, st));
    void *tmp;
    CK(hipMalloc(&tmp, tmp_bytes + 16));
    CK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys2, vals, vals2, (int)P, 0, 64, st));
    uint32_t hb[16];
    CK(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    uint32_t nleaf = hb[12];
    float vmin[3], vmax[3];
    for (int a = 0; a < 3; a++) { vmin[a] = o2f(hb[6 + a]); vmax[a] = o2f(hb[9 + a]); }
    make_periodic(S.h, vmin, vmax);

    float4 *rec = nullptr, *node = nullptr;
    uint32_t depth = 0;
    if (nleaf > 0) {
        float *lbox;
        CK(hipMalloc(&rec, sizeof(float4) * 4 * (size_t)nleaf));
        S.own.push_back(rec);
        CK(hipMalloc(&lbox, sizeof(float) * 6 * (size_t)nleaf));
        k_leaves<<<blocks(nleaf, 256), 256, 0, st>>>(nleaf, vals2, rec_u, box, bounds, rec, lbox);
        if (nleaf > 1) {
            uint32_t *child, *parent, *lparent, *visit, *dmax;
            float *ibox;
            CK(hipMalloc(&child, sizeof(uint32_t) * 2 * (size_t)(nleaf - 1)));
            CK(hipMalloc(&parent, sizeof(uint32_t) * (size_t)(nleaf - 1)));
            CK(hipMalloc(&lparent, sizeof(uint32_t) * (size_t)nleaf));
            CK(hipMalloc(&visit, sizeof(uint32_t) * (size_t)(nleaf - 1)));
            CK(hipMalloc(&dmax, sizeof(uint32_t)));
            CK(hipMalloc(&ibox, sizeof(float) * 6 * (size_t)(nleaf - 1)));
            CK(hipMalloc(&node, sizeof(float4) * 4 * (size_t)(nleaf - 1)));
            S.own.push_back(node);
            CK(hipMemsetAsync(visit, 0, sizeof(uint32_t) * (size_t)(nleaf - 1), st));
            CK(hipMemsetAsync(parent, 0xFF, sizeof(uint32_t) * (size_t)(nleaf - 1), st));
            CK(hipMemsetAsync(lparent, 0xFF, sizeof(uint32_t) * (size_t)nleaf, st));
            CK(hipMemsetAsync(child, 0, sizeof(uint32_t) * 2 * (size_t)(nleaf - 1), st));
            CK(hipMemsetAsync(dmax, 0, sizeof(uint32_t), st));
            k_karras<<<blocks(nleaf - 1, 256), 256, 0, st>>>((int)nleaf, keys2, child, parent, lparent);
            k_refit<<<blocks(nleaf, 256), 256, 0, st>>>((int)nleaf, child, parent, lparent, lbox, ibox, visit, dmax);
            k_emit<<<blocks(nleaf - 1, 256), 256, 0, st>>>((int)nleaf, child, lbox, ibox, node);
            CK(hipMemcpyAsync(&depth, dmax, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            CK(hipStreamSynchronize(st));
            CK(hipFree(child)); CK(hipFree(parent)); CK(hipFree(lparent)); CK(hipFree(visit));
            CK(hipFree(dmax)); CK(hipFree(ibox));
        } else {
            CK(hipStreamSynchronize(st));
        }
        CK(hipFree(lbox));
    }
    CK(hipFree(rec_u)); CK(hipFree(box)); CK(hipFree(bounds)); CK(hipFree(keys)); CK(hipFree(keys2));
    CK(hipFree(vals)); CK(hipFree(vals2)); CK(hipFree(tmp));
    if (depth + 1 > RT_STACK) {
        set_error("LBVH depth " + std::to_string(depth) + " exceeds the traversal stack (" + std::to_string(RT_STACK) + ")");
        scene_free(S);
        return -1;
    }
    S.h.rec = rec;
    S.h.node = node;
    S.h.nrm = nrm;
    S.h.frame = frame;
    S.h.T = dT;
    S.h.type = dtype;
    S.h.uuid = duuid;
    S.h.nleaf = nleaf;
    S.h.depth = depth;
    S.twos = dtwos;
    CK(hipMalloc(&S.d, sizeof(DevScene)));
    CK(hipMemcpyAsync(S.d, &S.h, sizeof(DevScene), hipMemcpyHostToDevice, st));
    CK(hipStreamSynchronize(st));
    S.bytes = (size_t)P * (64 + 4 * 3 + 16 * 2) + (size_t)nleaf * 64 + (nleaf > 1 ? (size_t)(nleaf - 1) * 64 : 0);
    return 0;
}

void scene_free(Scene &S) {
    for (void *p : S.own) (void)hipFree(p);
    S.own.clear();
    if (S.d) (void)hipFree(S.d);
    S.d = nullptr;
    S.twos = nullptr;
    S.bytes = 0;
    memset(&S.h, 0, sizeof(S.h));
}

}  // namespace rt
