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
// primitives that are leaves of the tree: the traced surfaces and voxels (castRays only)
__device__ __forceinline__ bool traced_t(uint32_t t) { return t == 0 || t == 1 || t == 3 || t == 4; }

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
    // raw AABB (inflated later) and bounds
    int nv = patch ? 4 : 3;
    float lo[3] = {v[0].x, v[0].y, v[0].z}, hi[3] = {v[0].x, v[0].y, v[0].z};
    for (int k = 1; k < nv; k++) {
        float c[3] = {v[k].x, v[k].y, v[k].z};
        for (int a = 0; a < 3; a++) { lo[a] = fminf(lo[a], c[a]); hi[a] = fmaxf(hi[a], c[a]); }
    }
    if (ty == 4) {
        // a voxel: the box of its eight transformed unit-cube corners, as castRays takes it from the
        // Context's vertices (core/src/Context_primitive.cpp:917-929, in that corner order) with a running
        // min/max that keeps the earlier of equal values (CD/src/CollisionDetection_RayTracing.cpp:1627-1637)
        for (int k = 0; k < 8; k++) {
            const float cx = ((k & 3) == 1 || (k & 3) == 2) ? 0.5f : -0.5f, cy = (k & 2) ? 0.5f : -0.5f;
            const v3f c = xform3(t, cx, cy, (k & 4) ? 0.5f : -0.5f);
            const float cc[3] = {c.x, c.y, c.z};
            for (int a = 0; a < 3; a++) {
                lo[a] = (k == 0 || cc[a] < lo[a]) ? cc[a] : lo[a];
                hi[a] = (k == 0 || cc[a] > hi[a]) ? cc[a] : hi[a];
            }
        }
        tag = i | RT_VOXEL_BIT;
        rr[0] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(tag));
        rr[1] = make_float4(hi[0], hi[1], hi[2], 0.f);
        rr[2] = make_float4(0.f, 0.f, 0.f, 0.f);
        rr[3] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        rr[0] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(tag));
        rr[1] = make_float4(e1.x, e1.y, e1.z, aa);
        rr[2] = make_float4(e2.x, e2.y, e2.z, bb);
        rr[3] = make_float4(pn.x, pn.y, pn.z, 0.f);
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

// ------------------------------------------------------------------------------------------------
// PLOC: parallel locally-ordered clustering (Meister & Bittner, IEEE TVCG 2018).  Clusters start as the
// leaves in Morton order.  Each round, every cluster finds the neighbour within +-PLOC_R positions whose
// union box has the smallest surface area (ties to the lower position); mutual nearest neighbours merge
// into a new internal node kept at the lower position; the array is compacted with a prefix sum.  The
// minimal pair under (area, positions) is always mutual, so every round merges at least once.  Internal
// nodes are numbered from the top (the last merge, the root, is node 0), and ranks come from the prefix
// sum, so the tree is the same on every run.
// ------------------------------------------------------------------------------------------------
#ifndef PLOC_R
#define PLOC_R 16
#endif
#define PLOC_BLOCK 256

__device__ __forceinline__ float half_area(const float a[6]) {
    float dx = a[3] - a[0], dy = a[4] - a[1], dz = a[5] - a[2];
    return dx * dy + dy * dz + dz * dx;
}

__global__ void k_ploc_init(uint32_t n, const float *__restrict__ lbox, float *__restrict__ cbox,
                            uint32_t *__restrict__ cid, uint32_t *__restrict__ ch) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int q = 0; q < 6; q++) cbox[(size_t)k * 6 + q] = lbox[(size_t)k * 6 + q];
    cid[k] = RT_LEAF_BIT | k;
    ch[k] = 0;
}

__global__ __launch_bounds__(PLOC_BLOCK) void k_ploc_nn(uint32_t n, const float *__restrict__ cbox,
                                                        uint32_t *__restrict__ nn) {
    __shared__ float sb[(PLOC_BLOCK + 2 * PLOC_R) * 6];
    const int base = (int)(blockIdx.x * PLOC_BLOCK) - PLOC_R;
    for (int t = threadIdx.x; t < (PLOC_BLOCK + 2 * PLOC_R) * 6; t += PLOC_BLOCK) {
        int j = base + t / 6;
        sb[t] = (j >= 0 && j < (int)n) ? cbox[(size_t)j * 6 + t % 6] : 0.f;
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * PLOC_BLOCK + threadIdx.x);
    if (i >= (int)n) return;
    const float *bi = sb + (threadIdx.x + PLOC_R) * 6;
    uint64_t best = ~0ull;
    const int lo = max(0, i - PLOC_R), hi = min((int)n - 1, i + PLOC_R);
    for (int j = lo; j <= hi; j++) {
        if (j == i) continue;
        const float *bj = sb + (j - base) * 6;
        float u[6];
        for (int a = 0; a < 3; a++) { u[a] = fminf(bi[a], bj[a]); u[3 + a] = fmaxf(bi[3 + a], bj[3 + a]); }
        uint64_t key = ((uint64_t)__float_as_uint(half_area(u)) << 32) | (uint32_t)j;
        best = key < best ? key : best;
    }
    nn[i] = (uint32_t)best;
}

// flags: high word = merges here (mutual pair, lower position), low word = cluster survives
__global__ void k_ploc_flags(uint32_t n, const uint32_t *__restrict__ nn, uint64_t *__restrict__ flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t j = nn[i];
    bool mutual = nn[j] == i;
    uint64_t merge = (mutual && i < j) ? 1ull : 0ull;
    uint64_t keep = (mutual && i > j) ? 0ull : 1ull;
    flags[i] = (merge << 32) | keep;
}

__global__ void k_ploc_emit(uint32_t n, uint32_t node_base, const uint32_t *__restrict__ nn,
                            const uint64_t *__restrict__ flags, const uint64_t *__restrict__ pref,
                            const float *__restrict__ cbox, const uint32_t *__restrict__ cid,
                            const uint32_t *__restrict__ ch, float *__restrict__ obox, uint32_t *__restrict__ oid,
                            uint32_t *__restrict__ oh, uint32_t *__restrict__ child, float *__restrict__ ibox) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t f = flags[i];
    if (!(f & 1ull)) return;
    uint32_t pos = (uint32_t)(pref[i] & 0xFFFFFFFFull);
    const float *bi = cbox + (size_t)i * 6;
    if (f >> 32) {
        uint32_t j = nn[i];
        uint32_t node = node_base - 1u - (uint32_t)(pref[i] >> 32);
        const float *bj = cbox + (size_t)j * 6;
        float u[6];
        for (int a = 0; a < 3; a++) { u[a] = fminf(bi[a], bj[a]); u[3 + a] = fmaxf(bi[3 + a], bj[3 + a]); }
        child[2 * (size_t)node] = cid[i];
        child[2 * (size_t)node + 1] = cid[j];
        for (int q = 0; q < 6; q++) { ibox[(size_t)node * 6 + q] = u[q]; obox[(size_t)pos * 6 + q] = u[q]; }
        oid[pos] = node;
        oh[pos] = 1u + max(ch[i], ch[j]);
    } else {
        for (int q = 0; q < 6; q++) obox[(size_t)pos * 6 + q] = bi[q];
        oid[pos] = cid[i];
        oh[pos] = ch[i];
    }
}

// ------------------------------------------------------------------------------------------------
// BVH4 collapse of the binary tree, level by level from the root.  A 4-wide node takes its binary
// node's two children and repeatedly opens the internal child with the largest surface area until it
// has four children or only leaves.  Nodes are numbered breadth-first (the root is 0).
// Node layout (8 float4, 128 B): lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4], links[4]
// (leaf bit 31 | leaf slot, else node index; unused slots 0x7FFFFFFF with a box at +inf), (child count, 0, 0, 0).
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ const float *box_of(uint32_t c, const float *lbox, const float *ibox) {
    return (c & RT_LEAF_BIT) ? lbox + (size_t)(c & ~RT_LEAF_BIT) * 6 : ibox + (size_t)c * 6;
}

__device__ int collapse_pick(uint32_t node2, const uint32_t *child, const float *ibox, uint32_t out[4]) {
    out[0] = child[2 * (size_t)node2];
    out[1] = child[2 * (size_t)node2 + 1];
    int cnt = 2;
    while (cnt < 4) {
        int best = -1;
        float ba = -1.f;
        for (int k = 0; k < cnt; k++)
            if (!(out[k] & RT_LEAF_BIT)) {
                float a = half_area(ibox + (size_t)out[k] * 6);
                if (a > ba) { ba = a; best = k; }
            }
        if (best < 0) break;
        uint32_t c = out[best];
        out[best] = child[2 * (size_t)c];
        out[cnt++] = child[2 * (size_t)c + 1];
    }
    return cnt;
}

__global__ void k_c4_count(uint32_t count, const uint32_t *__restrict__ list, const uint32_t *__restrict__ child,
                           const float *__restrict__ ibox, uint32_t *__restrict__ nint) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    uint32_t c[4];
    int cnt = collapse_pick(list[t], child, ibox, c);
    uint32_t m = 0;
    for (int k = 0; k < cnt; k++) m += (c[k] & RT_LEAF_BIT) ? 0u : 1u;
    nint[t] = m;
}

__global__ void k_c4_write(uint32_t count, uint32_t level_base, uint32_t next_base, const uint32_t *__restrict__ list,
                           const uint32_t *__restrict__ child, const float *__restrict__ lbox,
                           const float *__restrict__ ibox, const uint32_t *__restrict__ pref,
                           uint32_t *__restrict__ next_list, float4 *__restrict__ node4) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    uint32_t c[4];
    int cnt = collapse_pick(list[t], child, ibox, c);
    float lo[3][4], hi[3][4];
    uint32_t link[4];
    uint32_t r = pref[t];
    for (int k = 0; k < 4; k++) {
        if (k < cnt) {
            const float *b = box_of(c[k], lbox, ibox);
            for (int a = 0; a < 3; a++) { lo[a][k] = b[a]; hi[a][k] = b[3 + a]; }
            if (c[k] & RT_LEAF_BIT) {
                link[k] = c[k];
            } else {
                next_list[r] = c[k];
                link[k] = next_base + r;
                r++;
            }
        } else {
            // unused slot: a point box at +inf, which every slab test rejects (each axis gives +-inf entry and
            // exit distances: mixed signs make entry > exit, all +inf is beyond any tmax, all -inf is behind
            // the origin), and a link without the leaf bit that no traversal follows
            for (int a = 0; a < 3; a++) { lo[a][k] = INFINITY; hi[a][k] = INFINITY; }
            link[k] = 0x7FFFFFFFu;
        }
    }
    float4 *o = node4 + (size_t)(level_base + t) * 8;
    for (int a = 0; a < 3; a++) {
        o[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
        o[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
    o[6] = make_float4(__uint_as_float(link[0]), __uint_as_float(link[1]), __uint_as_float(link[2]),
                       __uint_as_float(link[3]));
    o[7] = make_float4(__uint_as_float((uint32_t)cnt), 0.f, 0.f, 0.f);
}

// 8-bit BVH4 nodes (64 B) from the float ones (128 B).  Per node: the union box of its children as the
// origin o and a power-of-two step s_a = 2^e_a per axis with 255 s_a >= extent; a child bound is stored as
// floor((lo - o) / s) and ceil((hi - o) / s), so the decoded box o + q s contains the float box exactly
// (the differences and quotients are exact in double: floats differ by a double-exact amount, s is a power
// of two).  Layout, as uint words:
//   q0 = (o.x, o.y, o.z bits, e_x+127 | e_y+127 << 8 | e_z+127 << 16 | children << 24)
//   q1 = (lo.x, hi.x, lo.y, hi.y), q2 = (lo.z, hi.z, link0, link1), q3 = (link2, link3, 0, 0)
// where each bound word holds the four children's bytes, child k in bits 8k..8k+7.  Unused slots (k >=
// children) are never tested.
__device__ __forceinline__ int q_exponent(double ext) {
    // smallest e in [-126, 127] with 255 * 2^e >= ext
    if (!(ext > 0.0)) return -126;
    int x;
    (void)frexp(ext / 255.0, &x);  // ext / 255 = m 2^x with m in [0.5, 1): 2^x > ext / 255
    return x < -126 ? -126 : (x > 127 ? 127 : x);
}

__global__ void k_q4(uint32_t nn4, const float4 *__restrict__ node4, uint4 *__restrict__ q4, int *__restrict__ bad) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn4) return;
    const float4 *o = node4 + (size_t)t * 8;
    const uint32_t cnt = __float_as_uint(o[7].x);
    float lo[3][4], hi[3][4];
    for (int a = 0; a < 3; a++) {
        const float4 l = o[2 * a], h = o[2 * a + 1];
        lo[a][0] = l.x; lo[a][1] = l.y; lo[a][2] = l.z; lo[a][3] = l.w;
        hi[a][0] = h.x; hi[a][1] = h.y; hi[a][2] = h.z; hi[a][3] = h.w;
    }
    uint32_t org[3], meta = cnt << 24, wl[3] = {0u, 0u, 0u}, wh[3] = {0u, 0u, 0u};
    for (int a = 0; a < 3; a++) {
        float u0 = INFINITY, u1 = -INFINITY;
        for (uint32_t k = 0; k < cnt; k++) { u0 = fminf(u0, lo[a][k]); u1 = fmaxf(u1, hi[a][k]); }
        const int e = q_exponent((double)u1 - (double)u0);
        const double s = ldexp(1.0, e);
        for (uint32_t k = 0; k < cnt; k++) {
            const double ql = floor(((double)lo[a][k] - (double)u0) / s);
            const double qh = ceil(((double)hi[a][k] - (double)u0) / s);
            if (!(ql >= 0.0 && qh <= 255.0 && ql <= qh)) atomicOr(bad, 1);
            wl[a] |= (uint32_t)fmin(fmax(ql, 0.0), 255.0) << (8 * k);
            wh[a] |= (uint32_t)fmin(fmax(qh, 0.0), 255.0) << (8 * k);
        }
        org[a] = __float_as_uint(u0);
        meta |= (uint32_t)(e + 127) << (8 * a);
    }
    const float4 lk = o[6];
    uint4 *q = q4 + (size_t)t * 4;
    q[0] = make_uint4(org[0], org[1], org[2], meta);
    q[1] = make_uint4(wl[0], wh[0], wl[1], wh[1]);
    q[2] = make_uint4(wl[2], wh[2], __float_as_uint(lk.x), __float_as_uint(lk.y));
    q[3] = make_uint4(__float_as_uint(lk.z), __float_as_uint(lk.w), 0u, 0u);
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

// PLOC rounds on the device.  Writes child[2*(n-1)] and ibox[(n-1)*6] and the tree height.  Returns 1 on
// success, 0 if a round made no progress (never expected; the caller then builds the Karras tree), -1 on a
// HIP error.
static int build_ploc(uint32_t n, const float *lbox, uint32_t *child, float *ibox, uint32_t &height, hipStream_t st) {
    float *box[2];
    uint32_t *id[2], *h[2], *nn;
    uint64_t *flags, *pref;
    CK(hipMalloc(&box[0], sizeof(float) * 6 * (size_t)n));
    CK(hipMalloc(&box[1], sizeof(float) * 6 * (size_t)n));
    CK(hipMalloc(&id[0], sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&id[1], sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&h[0], sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&h[1], sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&nn, sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&flags, sizeof(uint64_t) * (size_t)n));
    CK(hipMalloc(&pref, sizeof(uint64_t) * (size_t)n));
    size_t tmp_bytes = 0;
    CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, flags, pref, (int)n, st));
    void *tmp;
    CK(hipMalloc(&tmp, tmp_bytes + 16));
    k_ploc_init<<<blocks(n, 256), 256, 0, st>>>(n, lbox, box[0], id[0], h[0]);
    uint32_t cur = 0, count = n, node_base = n - 1;
    int ok = 1;
    while (count > 1) {
        k_ploc_nn<<<blocks(count, PLOC_BLOCK), PLOC_BLOCK, 0, st>>>(count, box[cur], nn);
        k_ploc_flags<<<blocks(count, 256), 256, 0, st>>>(count, nn, flags);
        CK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, flags, pref, (int)count, st));
        k_ploc_emit<<<blocks(count, 256), 256, 0, st>>>(count, node_base, nn, flags, pref, box[cur], id[cur], h[cur],
                                                        box[cur ^ 1], id[cur ^ 1], h[cur ^ 1], child, ibox);
        uint64_t last[2];
        CK(hipMemcpyAsync(&last[0], pref + count - 1, 8, hipMemcpyDeviceToHost, st));
        CK(hipMemcpyAsync(&last[1], flags + count - 1, 8, hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        uint32_t kept = (uint32_t)(last[0] & 0xFFFFFFFFull) + (uint32_t)(last[1] & 0xFFFFFFFFull);
        uint32_t merged = (uint32_t)(last[0] >> 32) + (uint32_t)(last[1] >> 32);
        if (merged == 0 || merged > node_base) { ok = 0; break; }
        node_base -= merged;
        count = kept;
        cur ^= 1;
    }
    uint32_t root_h = 0;
    if (ok) {
        CK(hipMemcpyAsync(&root_h, h[cur], 4, hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        if (node_base != 0) ok = 0;
    }
    height = root_h;
    CK(hipFree(box[0])); CK(hipFree(box[1])); CK(hipFree(id[0])); CK(hipFree(id[1]));
    CK(hipFree(h[0])); CK(hipFree(h[1])); CK(hipFree(nn)); CK(hipFree(flags)); CK(hipFree(pref)); CK(hipFree(tmp));
    return ok;
}

// BVH4 from the binary tree (child/ibox over n leaves with boxes lbox).  node4 holds n-1 slots.
static int build_bvh4(uint32_t n, const uint32_t *child, const float *lbox, const float *ibox, float4 *node4,
                      uint32_t &nnode4, uint32_t &levels, hipStream_t st) {
    uint32_t *list[2], *nint, *pref;
    CK(hipMalloc(&list[0], sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&list[1], sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&nint, sizeof(uint32_t) * (size_t)n));
    CK(hipMalloc(&pref, sizeof(uint32_t) * (size_t)n));
    size_t tmp_bytes = 0;
    CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, nint, pref, (int)n, st));
    void *tmp;
    CK(hipMalloc(&tmp, tmp_bytes + 16));
    CK(hipMemsetAsync(list[0], 0, sizeof(uint32_t), st));  // the binary root
    uint32_t count = 1, base = 0, cur = 0;
    levels = 0;
    while (count > 0) {
        if (base + count > n - 1) { set_error("BVH4 collapse produced too many nodes"); return -1; }
        k_c4_count<<<blocks(count, 256), 256, 0, st>>>(count, list[cur], child, ibox, nint);
        CK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, nint, pref, (int)count, st));
        k_c4_write<<<blocks(count, 256), 256, 0, st>>>(count, base, base + count, list[cur], child, lbox, ibox, pref,
                                                       list[cur ^ 1], node4);
        uint32_t last[2];
        CK(hipMemcpyAsync(&last[0], pref + count - 1, 4, hipMemcpyDeviceToHost, st));
        CK(hipMemcpyAsync(&last[1], nint + count - 1, 4, hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        base += count;
        count = last[0] + last[1];
        cur ^= 1;
        levels++;
    }
    nnode4 = base;
    CK(hipFree(list[0])); CK(hipFree(list[1])); CK(hipFree(nint)); CK(hipFree(pref)); CK(hipFree(tmp));
    return 0;
}

int validate_geometry(uint64_t P, const uint32_t *types, const char *who) {
    if (P > RT_MAX_PRIMS) {
        set_error(std::string(who) + ": " + std::to_string(P) + " primitives (at most 2^29)");
        return -1;
    }
    for (uint64_t i = 0; i < P; i++) {
        const uint32_t t = types[i];
        if (t == 0u || t == 1u || t == 3u || t == 4u) continue;
        const char *what = t == 2u ? "a disk (type 2): disks are not traced by this backend (Helios disks are "
                                     "triangle fans; pass their triangles)"
                                   : "an unknown type (0 patch, 1 triangle, 3 tile, 4 voxel)";
        set_error(std::string(who) + ": primitive " + std::to_string(i) + " is " + what);
        return -1;
    }
    return 0;
}

int check_work_items(uint64_t items, const char *what) {
    if (items <= RT_MAX_WORK_ITEMS) return 0;
    set_error(std::string(what) + ": " + std::to_string(items) + " work items in one launch (at most " +
              std::to_string(RT_MAX_WORK_ITEMS) + "); split the batch");
    return -1;
}

int scene_build(Scene &S, uint32_t P, const float *T, const uint32_t *type, const uint32_t *uuid,
                const uint32_t *twos, int px, int py, hipStream_t st) {
    if (validate_geometry(P, type, "scene")) return -1;
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
    S.h.smax = 0.f;
    for (int a = 0; a < 3; a++) S.h.smax = fmaxf(S.h.smax, fmaxf(fabsf(vmin[a]), fabsf(vmax[a])));

    float4 *rec = nullptr, *node = nullptr, *node4 = nullptr, *qnode4 = nullptr;
    uint32_t depth = 0, nnode4 = 0, depth4 = 0;
    if (nleaf > 0) {
        float *lbox;
        CK(hipMalloc(&rec, sizeof(float4) * 4 * (size_t)nleaf));
        S.own.push_back(rec);
        CK(hipMalloc(&lbox, sizeof(float) * 6 * (size_t)nleaf));
        k_leaves<<<blocks(nleaf, 256), 256, 0, st>>>(nleaf, vals2, rec_u, box, bounds, rec, lbox);
        if (nleaf > 1) {
            uint32_t *child;
            float *ibox;
            CK(hipMalloc(&child, sizeof(uint32_t) * 2 * (size_t)(nleaf - 1)));
            CK(hipMalloc(&ibox, sizeof(float) * 6 * (size_t)(nleaf - 1)));
            CK(hipMalloc(&node, sizeof(float4) * 4 * (size_t)(nleaf - 1)));
            S.own.push_back(node);
            int built = build_ploc(nleaf, lbox, child, ibox, depth, st);
            if (built < 0) return -1;
            if (built == 0 || depth + 1 > RT_STACK) {
                // PLOC tree too deep for the traversal stack: Karras LBVH over the same Morton order
                uint32_t *parent, *lparent, *visit, *dmax;
                CK(hipMalloc(&parent, sizeof(uint32_t) * (size_t)(nleaf - 1)));
                CK(hipMalloc(&lparent, sizeof(uint32_t) * (size_t)nleaf));
                CK(hipMalloc(&visit, sizeof(uint32_t) * (size_t)(nleaf - 1)));
                CK(hipMalloc(&dmax, sizeof(uint32_t)));
                CK(hipMemsetAsync(visit, 0, sizeof(uint32_t) * (size_t)(nleaf - 1), st));
                CK(hipMemsetAsync(parent, 0xFF, sizeof(uint32_t) * (size_t)(nleaf - 1), st));
                CK(hipMemsetAsync(lparent, 0xFF, sizeof(uint32_t) * (size_t)nleaf, st));
                CK(hipMemsetAsync(child, 0, sizeof(uint32_t) * 2 * (size_t)(nleaf - 1), st));
                CK(hipMemsetAsync(dmax, 0, sizeof(uint32_t), st));
                k_karras<<<blocks(nleaf - 1, 256), 256, 0, st>>>((int)nleaf, keys2, child, parent, lparent);
                k_refit<<<blocks(nleaf, 256), 256, 0, st>>>((int)nleaf, child, parent, lparent, lbox, ibox, visit,
                                                             dmax);
                CK(hipMemcpyAsync(&depth, dmax, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
                CK(hipStreamSynchronize(st));
                CK(hipFree(parent)); CK(hipFree(lparent)); CK(hipFree(visit)); CK(hipFree(dmax));
                S.builder = 0;
            } else {
                S.builder = 1;
            }
            k_emit<<<blocks(nleaf - 1, 256), 256, 0, st>>>((int)nleaf, child, lbox, ibox, node);
            CK(hipMalloc(&node4, sizeof(float4) * 8 * (size_t)(nleaf - 1)));
            S.own.push_back(node4);
            if (build_bvh4(nleaf, child, lbox, ibox, node4, nnode4, depth4, st)) return -1;
            CK(hipMalloc(&qnode4, sizeof(uint4) * 4 * (size_t)nnode4 + sizeof(int)));
            S.own.push_back(qnode4);
            int *qbad = (int *)((uint4 *)qnode4 + 4 * (size_t)nnode4);
            CK(hipMemsetAsync(qbad, 0, sizeof(int), st));
            k_q4<<<blocks(nnode4, 256), 256, 0, st>>>(nnode4, node4, (uint4 *)qnode4, qbad);
            int hbad = 0;
            CK(hipMemcpyAsync(&hbad, qbad, sizeof(int), hipMemcpyDeviceToHost, st));
            CK(hipStreamSynchronize(st));
            if (hbad) {
                set_error("8-bit BVH4 node quantisation failed (a child box outside its node's grid)");
                return -1;
            }
            CK(hipFree(child)); CK(hipFree(ibox));
        } else {
            CK(hipStreamSynchronize(st));
        }
        CK(hipFree(lbox));
    }
    CK(hipFree(rec_u)); CK(hipFree(box)); CK(hipFree(bounds)); CK(hipFree(keys)); CK(hipFree(keys2));
    CK(hipFree(vals)); CK(hipFree(tmp));
    S.own.push_back(vals2);  // traced primitives in leaf order, then the rest: the solve's work order
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
    S.h.order = vals2;
    // a BVH4 traversal pushes at most 3 entries per level: use it only when that fits the wave stack (128)
    S.h.node4 = (node4 && 3 * depth4 + 1 <= 128) ? node4 : nullptr;
    S.h.qnode4 = S.h.node4 ? qnode4 : nullptr;
    S.h.nnode4 = nnode4;
    S.h.depth4 = depth4;
    S.h.nleaf = nleaf;
    S.h.depth = depth;
    S.twos = dtwos;
    CK(hipMalloc(&S.d, sizeof(DevScene)));
    CK(hipMemcpyAsync(S.d, &S.h, sizeof(DevScene), hipMemcpyHostToDevice, st));
    CK(hipStreamSynchronize(st));
    S.bytes = (size_t)P * (64 + 4 * 4 + 16 * 2) + (size_t)nleaf * 64 + (nleaf > 1 ? (size_t)(nleaf - 1) * 64 : 0) +
              (size_t)(nleaf > 1 ? nleaf - 1 : 0) * 128 + (size_t)nnode4 * 64;
    return 0;
}

static inline float bits_f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__global__ void k_mask_tags(uint32_t nleaf, float4 *__restrict__ rec, const uint32_t *__restrict__ mdesc) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nleaf) return;
    float4 *r = rec + (size_t)k * 4;
    uint32_t tag = __float_as_uint(r[0].w) & ~RT_MASK_BIT;
    const uint32_t m = mdesc ? mdesc[tag & RT_ID_BITS] : RT_MISS;
    float w = 0.f;
    if (m != RT_MISS) {
        tag |= RT_MASK_BIT;
        w = __uint_as_float(m);
    }
    r[0].w = __uint_as_float(tag);
    r[3].w = w;
}

int scene_set_masks(Scene &S, uint32_t nmasks, const uint8_t *texels, const uint32_t *sizes, const int32_t *mask_id,
                    const float *uv, const int32_t *uv_id, hipStream_t st) {
    const uint32_t P = S.h.P, nleaf = S.h.nleaf;
    if (!S.d) { set_error("texture masks: no scene"); return -1; }
    if (P > RT_ID_BITS) { set_error("texture masks: more than 2^29 primitives"); return -1; }
    if (nmasks && (!texels || !sizes)) { set_error("texture masks: texels and sizes are required"); return -1; }
    if (nmasks && !mask_id) { set_error("texture masks: mask_id is required"); return -1; }
    std::vector<uint64_t> toff(nmasks + 1, 0);
    for (uint32_t m = 0; m < nmasks; m++) {
        const uint64_t w = sizes[2 * m], h = sizes[2 * m + 1];
        if (w == 0 || h == 0 || w > 65536 || h > 65536) {
            set_error("texture masks: mask " + std::to_string(m) + " has size " + std::to_string(w) + "x" +
                      std::to_string(h) + " (1..65536 texels per side)");
            return -1;
        }
        toff[m + 1] = toff[m] + w * h;
    }
    std::vector<uint32_t> type(P);
    if (P) CK(hipMemcpy(type.data(), S.h.type, sizeof(uint32_t) * P, hipMemcpyDeviceToHost));
    std::vector<uint32_t> mdesc(P, RT_MISS);
    std::vector<float4> desc;
    std::vector<uint32_t> dmask;  // mask of each descriptor
    const uint64_t rec_f4 = (uint64_t)nleaf * 4;
    for (uint32_t i = 0; i < P && nmasks; i++) {
        const int32_t mid = mask_id[i];
        if (mid < 0) continue;
        if ((uint32_t)mid >= nmasks) {
            set_error("texture masks: primitive " + std::to_string(i) + " names mask " + std::to_string(mid) + " of " +
                      std::to_string(nmasks));
            return -1;
        }
        const uint32_t ty = type[i];
        if (!(ty == 0u || ty == 1u || ty == 3u)) continue;  // not traced
        // triangles always interpolate their vertex UVs; patches use them when uv_id >= 0 (UVs read as
        // zeros when uv is NULL, as the reference's zero-filled per-primitive UV buffer)
        const uint32_t mode = ty == 1u ? 2u : ((uv_id && uv_id[i] >= 0) ? 1u : 0u);
        mdesc[i] = (uint32_t)(rec_f4 + desc.size());  // float4 index of this descriptor
        const float *u = uv ? uv + (size_t)i * 8 : nullptr;
        desc.push_back(make_float4(0.f, bits_f(sizes[2 * mid]), bits_f(sizes[2 * mid + 1]),
                                   bits_f(mode)));
        desc.push_back(u ? make_float4(u[0], u[1], u[2], u[3]) : make_float4(0.f, 0.f, 0.f, 0.f));
        desc.push_back(u ? make_float4(u[4], u[5], u[6], u[7]) : make_float4(0.f, 0.f, 0.f, 0.f));
        dmask.push_back((uint32_t)mid);
    }
    const uint64_t tex_base = (rec_f4 + desc.size()) * sizeof(float4);
    const uint64_t total = tex_base + toff[nmasks];
    if (total >= (1ull << 32)) {
        set_error("texture masks: records, descriptors and texels exceed 4 GiB");
        return -1;
    }
    for (size_t j = 0; j < dmask.size(); j++) desc[3 * j].x = bits_f((uint32_t)(tex_base + toff[dmask[j]]));
    if (nleaf == 0) return 0;
    float4 *R = nullptr;
    uint32_t *dm = nullptr;
    CK(hipMalloc(&R, (total + 15) & ~(uint64_t)15));
    CK(hipMemcpyAsync(R, S.h.rec, rec_f4 * sizeof(float4), hipMemcpyDeviceToDevice, st));
    if (!desc.empty()) {
        CK(hipMemcpyAsync(R + rec_f4, desc.data(), desc.size() * sizeof(float4), hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync((char *)R + tex_base, texels, toff[nmasks], hipMemcpyHostToDevice, st));
        CK(hipMalloc(&dm, sizeof(uint32_t) * P));
        CK(hipMemcpyAsync(dm, mdesc.data(), sizeof(uint32_t) * P, hipMemcpyHostToDevice, st));
    }
    k_mask_tags<<<blocks(nleaf, 256), 256, 0, st>>>(nleaf, R, dm);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(st));
    // replace the record buffer and the descriptor map
    auto drop = [&S](const void *p) {
        for (size_t k = 0; k < S.own.size(); k++)
            if (S.own[k] == p) { (void)hipFree(S.own[k]); S.own.erase(S.own.begin() + k); return; }
    };
    drop(S.h.rec);
    if (S.h.mdesc) drop(S.h.mdesc);
    S.own.push_back(R);
    if (dm) S.own.push_back(dm);
    S.h.rec = R;
    S.h.mdesc = dm;
    CK(hipMemcpyAsync(S.d, &S.h, sizeof(DevScene), hipMemcpyHostToDevice, st));
    CK(hipStreamSynchronize(st));
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
