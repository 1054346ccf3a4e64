// rt_trace.hip -- BVH traversal and the wavefront-persistent ray kernels (pieces b, c, d).
//
//   trace_closest   stack traversal of the 64-byte BVH2 (both child boxes per node), LDS stack,
//                   closest hit with the reference acceptance rule (t > tmin, skip the origin
//                   primitive; R/shaders/common/bvh_traversal.glsl:68-170) and a canonical tie rule
//                   (equal t -> lower primitive index) instead of traversal order.
//   trace_any       occlusion query: any accepted hit with t <= tmax ends the traversal.  The direct
//                   pass only needs "occluded or not"; with the nearest periodic face as tmax this
//                   reproduces the reference decision (closest hit vs periodic face) exactly.
//   k_direct        direct rays (R/shaders/direct_raygen.comp:323-731): one wave per (job, primitive),
//                   d*d stratified samples, occlusion trace with periodic wrap, compensated sums of
//                   ray strength split by lit face.
//   k_diffuse       diffuse/scatter rays (R/shaders/diffuse_raygen.comp:330-560): one wave per
//                   (pass, primitive), both faces, writes one hit word per ray.
//   k_gather_*      per-primitive compensated reduction of gathered radiation_out / sky flux over a
//                   hit table (diffuse_raygen.comp:562-652), deterministic (no atomics).
//   k_cast          CollisionDetection::castRaysSoA (CD/src/CollisionDetection_RayTracing.cpp:551-680,
//                   980-1080): one thread per ray, CD predicates, face-forwarded normal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include <cfloat>
#include <string>

#include "rt_internal.h"

// 4-wide tree for diffuse rays (1) or the binary tree (0); children of a 4-wide node fully sorted by entry
// distance (1) or only the nearest taken first (0).  On the 1 M-primitive orchard the binary tree is faster
// (916 ms per series of 5 passes against 956 sorted and 1021 nearest-first): the 4-wide step needs 112 B of
// node per lane and more registers (6 instead of 7 waves per SIMD) for ~40 % fewer steps.
#ifndef RT_DIFFUSE_BVH4
#define RT_DIFFUSE_BVH4 0
#endif
#ifndef RT_DIFFUSE_QBVH4
#define RT_DIFFUSE_QBVH4 1
#endif
#ifndef RT_CAST_QBVH4
#define RT_CAST_QBVH4 1
#endif
#ifndef RT_Q4_EAGER
#define RT_Q4_EAGER 1
#endif
// BVH4 pushes without branches into a spare LDS row per lane (the per-lane stacks then hold LSTK + 1 rows)
#ifndef RT_PUSH_SPARE
#define RT_PUSH_SPARE 1
#endif
#ifndef RT_BVH4_SORT
#define RT_BVH4_SORT 1
#endif
// 8-bit node steps check the step budget and the link without a branch (the kernel reports a failed lane
// when it finishes) and test the stack bound only on the pushes that can exceed it: 626 -> 611 ms for the
// orchard's five diffuse passes
#ifndef RT_Q4_LATE_ERR
#define RT_Q4_LATE_ERR 1
#endif
#define RT_BUDGET_BAD 0xFFFFFFFFu
// Diffuse rays take the 8-bit node step with the pop read issued beside the node load and no branch on the
// hit-child count (611 -> 600 ms for the orchard's five passes); castRays keep the branching step, which is
// faster for their long rays (3,112 against 2,700 Mrays/s)
#ifndef RT_Q4_FLAT_DIFFUSE
#define RT_Q4_FLAT_DIFFUSE 1
#endif

namespace rt {

__device__ __forceinline__ float min3f(float a, float b, float c) { return fminf(fminf(a, b), c); }
__device__ __forceinline__ float max3f(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }

// traversal work counters (node fetches, primitive tests) of one lane.  The counter type also carries the
// traversal's flags MK: bit 0 = honour texture masks (off compiles the mask tests away for scenes without
// masks, keeping the leaf code and its register use as small as before masks existed), bit 1 = watertight
// intersection predicates instead of the reference's (rt_set_predicates).
template <int MK>
struct TCntT {
    uint32_t nodes, prims;
    static constexpr bool mask = (MK & RT_FL_MASK) != 0;
    static constexpr bool wt = (MK & RT_FL_WT) != 0;
    static constexpr int flags = MK;
};
// the same interface with every update compiled away (timed kernels)
struct NopCounter {
    __device__ __forceinline__ NopCounter operator++(int) { return *this; }
    __device__ __forceinline__ NopCounter &operator+=(uint32_t) { return *this; }
};
template <int MK>
struct NoCntT {
    NopCounter nodes, prims;
    static constexpr bool mask = (MK & RT_FL_MASK) != 0;
    static constexpr bool wt = (MK & RT_FL_WT) != 0;
    static constexpr int flags = MK;
};
using TCnt = TCntT<RT_FL_MASK>;
using NoCnt = NoCntT<RT_FL_MASK>;
template <bool COUNT, int MK>
using Cnt = typename std::conditional<COUNT, TCntT<MK>, NoCntT<MK>>::type;

struct RayPre {
    float ix, iy, iz, ox, oy, oz;
};

__device__ __forceinline__ RayPre ray_pre(v3f o, v3f d) {
    RayPre r;
    r.ix = fabsf(d.x) > 1e-30f ? 1.0f / d.x : 1e30f;
    r.iy = fabsf(d.y) > 1e-30f ? 1.0f / d.y : 1e30f;
    r.iz = fabsf(d.z) > 1e-30f ? 1.0f / d.z : 1e30f;
    r.ox = o.x * r.ix;
    r.oy = o.y * r.iy;
    r.oz = o.z * r.iz;
    return r;
}

// BVH nodes and primitive records do not change while a trace kernel runs: loading them through the
// constant address space lets the compiler use scalar loads when the index is wave-uniform (the
// packet traversal) and ordinary vector loads otherwise.
__device__ __forceinline__ float4 ldc(const float4 *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(4))) const float cfloat_t;
    const cfloat_t *q = (const cfloat_t *)(const float *)p;
    return make_float4(q[0], q[1], q[2], q[3]);
#else
    return *p;
#endif
}

// wave-uniform copy of a pointer (lets the compiler keep it, and addresses built from uniform
// indices, in scalar registers)
template <typename T>
__device__ __forceinline__ T *uptr(T *p) {
    uint64_t v = (uint64_t)p;
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T *)(((uint64_t)hi << 32) | lo);
}

// Load of a value at a wave-uniform address through the constant address space (scalar load).
template <typename T>
__device__ __forceinline__ T ldu(const T *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(4))) const T cT;
    return *(const cT *)p;
#else
    return *p;
#endif
}

// Intersect the primitive in leaf slot k; returns its index (RT_MISS if skipped) and the hit.
template <bool CD, int MK = RT_FL_MASK>
__device__ __forceinline__ bool leaf_hit_p(const float4 *recs, uint32_t k, v3f o, v3f d, uint32_t skip, uint32_t &id,
                                           float &t) {
    const float4 *r = recs + (size_t)k * 4;
    float4 r0 = ldc(r);
    uint32_t tag = __float_as_uint(r0.w);
    id = tag & RT_ID_BITS;
    bool patch = (tag & RT_PATCH_BIT) != 0;
    const bool masked = (MK & RT_FL_MASK) && !CD && (tag & RT_MASK_BIT) != 0;
    float4 r1 = ldc(r + 1), r2 = ldc(r + 2);
    // a voxel is a box that castRays hits (CollisionDetection_RayTracing.cpp:1625-1711) and the radiation
    // traversal never does (bvh_traversal.glsl:118-140 tests patches, tiles and triangles only)
    if (CD && (tag & RT_VOXEL_BIT)) return voxel_hit_cd(r0, r1, o, d, t) && id != skip;
    float4 r3 = (patch || masked) ? ldc(r + 3) : make_float4(0.f, 0.f, 0.f, 0.f);
    float ua, ub;
    bool hit = isect_rec_uv<CD, (MK & RT_FL_WT) != 0>(r0, r1, r2, r3, patch, o, d, t, ua, ub) && id != skip &&
               (CD || !(tag & RT_VOXEL_BIT));
    // texture mask (bvh_traversal.glsl:143-148): a transparent texel lets the ray pass
    if (masked && hit) hit = mask_hit_opaque(recs, __float_as_uint(r3.w), patch, ua, ub, r1.w, r2.w);
    return hit;
}

template <bool CD, int MK = RT_FL_MASK>
__device__ __forceinline__ bool leaf_hit(const DevScene &sc, uint32_t k, v3f o, v3f d, uint32_t skip, uint32_t &id,
                                         float &t) {
    return leaf_hit_p<CD, MK>(sc.rec, k, o, d, skip, id, t);
}

template <bool CD, int MK = RT_FL_MASK>
__device__ __forceinline__ void test_leaf(const float4 *recs, uint32_t k, v3f o, v3f d, float tmin, uint32_t skip,
                                          float &best_t, uint32_t &best, uint32_t &best_slot) {
    uint32_t id;
    float t;
    // castRays takes a patch or triangle beyond 1e-6 and a voxel at 1e-6 or beyond
    // (CD/src/CollisionDetection_RayTracing.cpp:1562, 1595, 1681)
    if (leaf_hit_p<CD, MK>(recs, k, o, d, skip, id, t) &&
        (t > tmin || (CD && t == tmin && (__float_as_uint(ldc(recs + (size_t)k * 4).w) & RT_VOXEL_BIT)))) {
        if (t < best_t || (t == best_t && best != RT_MISS && id < best)) {
            best_t = t;
            best = id;
            best_slot = k;
        }
    }
}

// slab tests of both children of node ni against [.., tlim]
struct NodeTest {
    float tn0, tn1;
    uint32_t c0, c1;
    bool h0, h1;
};

__device__ __forceinline__ NodeTest node_test_p(const float4 *nodes, uint32_t ni, const RayPre &rp, float tlim);

__device__ __forceinline__ NodeTest node_test(const DevScene &sc, uint32_t ni, const RayPre &rp, float tlim) {
    return node_test_p(sc.node, ni, rp, tlim);
}

__device__ __forceinline__ NodeTest node_test_p(const float4 *nodes, uint32_t ni, const RayPre &rp, float tlim) {
    const float4 *nd = nodes + (size_t)ni * 4;
    float4 n0 = ldc(nd), n1 = ldc(nd + 1), n2 = ldc(nd + 2), n3 = ldc(nd + 3);
    float ax = fmaf(n0.x, rp.ix, -rp.ox), bx = fmaf(n0.y, rp.ix, -rp.ox);
    float ay = fmaf(n0.z, rp.iy, -rp.oy), by = fmaf(n0.w, rp.iy, -rp.oy);
    float az = fmaf(n1.x, rp.iz, -rp.oz), bz = fmaf(n1.y, rp.iz, -rp.oz);
    float tn0 = max3f(fminf(ax, bx), fminf(ay, by), fminf(az, bz));
    float tf0 = min3f(fmaxf(ax, bx), fmaxf(ay, by), fmaxf(az, bz));
    float cx = fmaf(n1.z, rp.ix, -rp.ox), dx = fmaf(n1.w, rp.ix, -rp.ox);
    float cy = fmaf(n2.x, rp.iy, -rp.oy), dy = fmaf(n2.y, rp.iy, -rp.oy);
    float cz = fmaf(n2.z, rp.iz, -rp.oz), dz = fmaf(n2.w, rp.iz, -rp.oz);
    float tn1 = max3f(fminf(cx, dx), fminf(cy, dy), fminf(cz, dz));
    float tf1 = min3f(fmaxf(cx, dx), fmaxf(cy, dy), fmaxf(cz, dz));
    NodeTest r;
    r.tn0 = tn0;
    r.tn1 = tn1;
    // tn <= tf && tf >= 0 && tn <= tlim, as one comparison (tlim >= 0; NaN only ever widens the test)
    r.h0 = fmaxf(tn0, 0.f) <= fminf(tf0, tlim);
    r.h1 = fmaxf(tn1, 0.f) <= fminf(tf1, tlim);
    r.c0 = __float_as_uint(n3.x);
    r.c1 = __float_as_uint(n3.y);
    return r;
}

// Closest accepted hit (t > tmin, t < tstart, ties -> lower primitive index).
template <bool CD, int STK, typename C>
__device__ uint32_t trace_closest(const DevScene &sc, v3f o, v3f d, float tmin, uint32_t skip, float tstart,
                                  float &t_out, uint32_t &slot_out, uint32_t *stk, uint32_t stride, int *err,
                                  C &cnt) {
    float best_t = tstart;
    uint32_t best = RT_MISS, best_slot = RT_MISS;
    // scene fields in registers: the stack stores below go through a generic pointer, and the compiler
    // would otherwise reload them from the descriptor on every step
    const float4 *nodes = sc.node, *recs = sc.rec;
    const uint32_t nleaf = sc.nleaf;
    if (nleaf <= 1) {
        if (nleaf == 1) { cnt.prims++; test_leaf<CD, C::flags>(recs, 0, o, d, tmin, skip, best_t, best, best_slot); }
        t_out = best_t;
        slot_out = best_slot;
        return best;
    }
    RayPre rp = ray_pre(o, d);
    uint32_t ni = 0;
    int sp = 0;
    // every internal node is entered at most once per trace; the cap only guards malformed trees
    uint32_t budget = nleaf + 8;
    while (true) {
        if (budget-- == 0 || ni >= nleaf - 1) {
            atomicOr(err, 2);
            break;
        }
        cnt.nodes++;
        NodeTest nt = node_test_p(nodes, ni, rp, best_t);
        bool l0 = (nt.c0 & RT_LEAF_BIT) != 0, l1 = (nt.c1 & RT_LEAF_BIT) != 0;
        if (nt.h0 && l0) { cnt.prims++; test_leaf<CD, C::flags>(recs, nt.c0 & ~RT_LEAF_BIT, o, d, tmin, skip, best_t, best, best_slot); }
        if (nt.h1 && l1) { cnt.prims++; test_leaf<CD, C::flags>(recs, nt.c1 & ~RT_LEAF_BIT, o, d, tmin, skip, best_t, best, best_slot); }
        bool g0 = nt.h0 && !l0, g1 = nt.h1 && !l1;
        if (g0 && g1) {
            bool first0 = nt.tn0 <= nt.tn1;
            if (sp >= STK) {
                atomicOr(err, 1);
                break;
            }
            stk[(uint32_t)sp * stride] = first0 ? nt.c1 : nt.c0;
            sp++;
            ni = first0 ? nt.c0 : nt.c1;
        } else if (g0) {
            ni = nt.c0;
        } else if (g1) {
            ni = nt.c1;
        } else {
            if (sp == 0) break;
            sp--;
            ni = stk[(uint32_t)sp * stride];
        }
    }
    t_out = best_t;
    slot_out = best_slot;
    return best;
}

// Any accepted radiation-predicate hit with tmin < t <= tmax?
template <int STK, typename C>
__device__ bool trace_any(const DevScene &sc, v3f o, v3f d, float tmin, uint32_t skip, float tmax, uint32_t *stk,
                          uint32_t stride, int *err, C &cnt) {
    uint32_t id;
    float t;
    const float4 *nodes = sc.node, *recs = sc.rec;
    const uint32_t nleaf = sc.nleaf;
    if (nleaf <= 1) {
        if (nleaf == 1) {
            cnt.prims++;
            return leaf_hit_p<false, C::flags>(recs, 0, o, d, skip, id, t) && t > tmin && t <= tmax;
        }
        return false;
    }
    RayPre rp = ray_pre(o, d);
    uint32_t ni = 0;
    int sp = 0;
    uint32_t budget = nleaf + 8;
    while (true) {
        if (budget-- == 0 || ni >= nleaf - 1) {
            atomicOr(err, 2);
            return false;
        }
        cnt.nodes++;
        NodeTest nt = node_test_p(nodes, ni, rp, tmax);
        bool l0 = (nt.c0 & RT_LEAF_BIT) != 0, l1 = (nt.c1 & RT_LEAF_BIT) != 0;
        if (nt.h0 && l0) {
            cnt.prims++;
            if (leaf_hit_p<false, C::flags>(recs, nt.c0 & ~RT_LEAF_BIT, o, d, skip, id, t) && t > tmin && t <= tmax) return true;
        }
        if (nt.h1 && l1) {
            cnt.prims++;
            if (leaf_hit_p<false, C::flags>(recs, nt.c1 & ~RT_LEAF_BIT, o, d, skip, id, t) && t > tmin && t <= tmax) return true;
        }
        bool g0 = nt.h0 && !l0, g1 = nt.h1 && !l1;
        if (g0 && g1) {
            bool first0 = nt.tn0 <= nt.tn1;
            if (sp >= STK) {
                atomicOr(err, 1);
                return false;
            }
            stk[(uint32_t)sp * stride] = first0 ? nt.c1 : nt.c0;
            sp++;
            ni = first0 ? nt.c0 : nt.c1;
        } else if (g0) {
            ni = nt.c0;
        } else if (g1) {
            ni = nt.c1;
        } else {
            if (sp == 0) return false;
            sp--;
            ni = stk[(uint32_t)sp * stride];
        }
    }
}

// Wave-coherent any-hit query: the 64 lanes (rays of one primitive towards one source, hence nearly
// parallel and close together) walk the BVH together.  The node index and the stack are wave-uniform
// (node fetched once per wave through scalar loads, stack in a per-wave LDS column); every lane tests
// the child boxes and primitives itself and stops searching at its first accepted hit.  A lane reaches
// every node whose box chain it intersects, so its answer equals the per-lane traversal's.
template <int STK, typename C>
__device__ bool packet_any(const DevScene &sc, v3f o, v3f d, float tmin, uint32_t skip, float tmax, bool active,
                           uint32_t *wstk, int *err, C &cnt) {
    bool act = active, found = false;
    uint32_t id;
    float t;
    if (sc.nleaf <= 1) {
        if (sc.nleaf == 1 && act) {
            cnt.prims++;
            found = leaf_hit<false, C::flags>(sc, 0, o, d, skip, id, t) && t > tmin && t <= tmax;
        }
        return found;
    }
    RayPre rp = ray_pre(o, d);
    const float4 *nodes = uptr(sc.node), *recs = uptr(sc.rec);
    const uint32_t nleaf = __builtin_amdgcn_readfirstlane(sc.nleaf);
    uint32_t ni = 0;
    uint32_t sp = 0;
    uint32_t budget = nleaf + 8;
    while (__ballot(act) != 0) {
        ni = __builtin_amdgcn_readfirstlane(ni);
        if (budget-- == 0 || ni >= nleaf - 1) {
            atomicOr(err, 2);
            break;
        }
        if (act) cnt.nodes++;  // per-ray visits (one fetch serves the whole wave)
        NodeTest nt = node_test_p(nodes, ni, rp, tmax);
        uint32_t c0 = __builtin_amdgcn_readfirstlane(nt.c0), c1 = __builtin_amdgcn_readfirstlane(nt.c1);
        nt.c0 = c0;
        nt.c1 = c1;
        bool l0 = (nt.c0 & RT_LEAF_BIT) != 0, l1 = (nt.c1 & RT_LEAF_BIT) != 0;
        bool h0 = act && nt.h0, h1 = act && nt.h1;
        if (l0 && h0) {
            cnt.prims++;
            if (leaf_hit_p<false, C::flags>(recs, nt.c0 & ~RT_LEAF_BIT, o, d, skip, id, t) && t > tmin && t <= tmax) {
                found = true;
                act = false;
            }
        }
        if (l1 && h1 && act) {
            cnt.prims++;
            if (leaf_hit_p<false, C::flags>(recs, nt.c1 & ~RT_LEAF_BIT, o, d, skip, id, t) && t > tmin && t <= tmax) {
                found = true;
                act = false;
            }
        }
        bool g0 = !l0 && __ballot(h0 && act) != 0;
        bool g1 = !l1 && __ballot(h1 && act) != 0;
        if (g0 && g1) {
            // nearer child first, as seen by the wave's first still-searching lane
            uint64_t m = __ballot(act && nt.tn0 <= nt.tn1);
            uint64_t a = __ballot(act);
            bool first0 = (m & (a & (~a + 1))) != 0;
            if (sp >= STK) {
                atomicOr(err, 1);
                break;
            }
            wstk[sp] = first0 ? nt.c1 : nt.c0;  // same value from every lane
            sp++;
            ni = first0 ? nt.c0 : nt.c1;
        } else if (g0) {
            ni = nt.c0;
        } else if (g1) {
            ni = nt.c1;
        } else {
            if (sp == 0) break;
            sp--;
            ni = wstk[sp];
        }
    }
    return found;
}

// Slab test of one child box of a BVH4 node.
__device__ __forceinline__ void slab4(float lx, float hx, float ly, float hy, float lz, float hz, const RayPre &rp,
                                      float &tn, float &tf) {
    float ax = fmaf(lx, rp.ix, -rp.ox), bx = fmaf(hx, rp.ix, -rp.ox);
    float ay = fmaf(ly, rp.iy, -rp.oy), by = fmaf(hy, rp.iy, -rp.oy);
    float az = fmaf(lz, rp.iz, -rp.oz), bz = fmaf(hz, rp.iz, -rp.oz);
    tn = max3f(fminf(ax, bx), fminf(ay, by), fminf(az, bz));
    tf = min3f(fmaxf(ax, bx), fmaxf(ay, by), fmaxf(az, bz));
}

// Ascending compare-exchange of two (key, link) pairs held in scalar registers.
__device__ __forceinline__ void cswap(float &ka, uint32_t &la, float &kb, uint32_t &lb) {
    bool sw = kb < ka;
    float k = sw ? kb : ka;
    uint32_t l = sw ? lb : la;
    kb = sw ? ka : kb;
    lb = sw ? la : lb;
    ka = k;
    la = l;
}

// packet_any() on the BVH4: one 128-byte node per step (two scalar loads), four child boxes tested by every
// lane.  Leaf children are tested by the lanes whose box test passed; internal children hit by any still
// searching lane are visited nearest-first as seen by the wave's first searching lane (a sorting network on
// at most four scalar keys), the others pushed on the per-wave LDS stack.  A lane still reaches every node
// whose box chain it intersects, so each lane's answer equals the single-ray traversal's.
#define RT_WSTK4 128
template <typename C>
__device__ bool packet_any4(const DevScene &sc, v3f o, v3f d, float tmin, uint32_t skip, float tmax, bool active,
                            uint32_t *wstk, int *err, C &cnt) {
    bool act = active, found = false;
    uint32_t id;
    float t;
    if (sc.nleaf <= 1) {
        if (sc.nleaf == 1 && act) {
            cnt.prims++;
            found = leaf_hit<false, C::flags>(sc, 0, o, d, skip, id, t) && t > tmin && t <= tmax;
        }
        return found;
    }
    RayPre rp = ray_pre(o, d);
    const float4 *nodes = uptr(sc.node4), *recs = uptr(sc.rec);
    const uint32_t nn4 = __builtin_amdgcn_readfirstlane(sc.nnode4);
    uint32_t ni = 0, sp = 0;
    uint32_t budget = nn4 + 8;
    while (__ballot(act) != 0) {
        ni = __builtin_amdgcn_readfirstlane(ni);
        if (budget-- == 0 || ni >= nn4) {
            atomicOr(err, 2);
            break;
        }
        if (act) cnt.nodes++;
        const float4 *nd = nodes + (size_t)ni * 8;
        const float4 lx = ldc(nd), hx = ldc(nd + 1), ly = ldc(nd + 2), hy = ldc(nd + 3);
        const float4 lz = ldc(nd + 4), hz = ldc(nd + 5), lk = ldc(nd + 6);
        uint32_t lnk[4];
        lnk[0] = __builtin_amdgcn_readfirstlane(__float_as_uint(lk.x));
        lnk[1] = __builtin_amdgcn_readfirstlane(__float_as_uint(lk.y));
        lnk[2] = __builtin_amdgcn_readfirstlane(__float_as_uint(lk.z));
        lnk[3] = __builtin_amdgcn_readfirstlane(__float_as_uint(lk.w));
        float tn[4], tf[4];
        slab4(lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, rp, tn[0], tf[0]);
        slab4(lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, rp, tn[1], tf[1]);
        slab4(lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, rp, tn[2], tf[2]);
        slab4(lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, rp, tn[3], tf[3]);
        bool h[4];
#pragma unroll
        for (int k = 0; k < 4; k++) h[k] = fmaxf(tn[k], 0.f) <= fminf(tf[k], tmax);  // unused slots never pass
        // leaf children: a wave-uniform branch per leaf, the test itself predicated per lane
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (lnk[k] & RT_LEAF_BIT) {
                const bool want = h[k] && act;
                if (__ballot(want) != 0) {
                    cnt.prims += want ? 1u : 0u;
                    const bool hit = leaf_hit_p<false, C::flags>(recs, lnk[k] & ~RT_LEAF_BIT, o, d, skip, id, t) && t > tmin &&
                                     t <= tmax;
                    const bool take = want && hit;
                    found = found || take;
                    act = act && !take;
                }
            }
        }
        const uint64_t am = __ballot(act);
        if (am == 0) break;
        const uint32_t leader = (uint32_t)__builtin_ctzll(am);
        // internal children some searching lane hits, keyed by the leader's entry distance
        float key[4];
        uint32_t lk4[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            bool inner = !(lnk[k] & RT_LEAF_BIT);
            bool any = inner && __ballot(h[k] && act) != 0;
            const int kk = __builtin_amdgcn_readlane(__float_as_int(h[k] ? tn[k] : 3.0e38f), (int)leader);
            key[k] = any ? __int_as_float(kk) : INFINITY;
            lk4[k] = any ? lnk[k] : 0xFFFFFFFFu;
        }
        cswap(key[0], lk4[0], key[1], lk4[1]);
        cswap(key[2], lk4[2], key[3], lk4[3]);
        cswap(key[0], lk4[0], key[2], lk4[2]);
        cswap(key[1], lk4[1], key[3], lk4[3]);
        cswap(key[1], lk4[1], key[2], lk4[2]);
        const uint32_t m = (lk4[0] != 0xFFFFFFFFu) + (lk4[1] != 0xFFFFFFFFu) + (lk4[2] != 0xFFFFFFFFu) +
                           (lk4[3] != 0xFFFFFFFFu);
        if (m == 0) {
            if (sp == 0) break;
            sp--;
            ni = wstk[sp];
        } else {
            if (sp + m - 1 > RT_WSTK4) {
                atomicOr(err, 1);
                break;
            }
            if (m >= 4) wstk[sp++] = lk4[3];
            if (m >= 3) wstk[sp++] = lk4[2];
            if (m >= 2) wstk[sp++] = lk4[1];
            ni = lk4[0];
        }
    }
    return found;
}

// R/shaders/common/periodic_boundary.glsl:28-59
__device__ __forceinline__ float isect_bbox_face(const BBoxFace &f, v3f o, v3f d) {
    v3f n = mk3(f.n[0], f.n[1], f.n[2]);
    v3f v0 = mk3(f.v0[0], f.v0[1], f.v0[2]);
    float denom = dot3(d, n);
    if (fabsf(denom) < 1e-8f) return -1.0f;
    float t = rdiv(dot3(sub3(v0, o), n), denom);
    if (t <= 1e-5f || t > 1e8f) return -1.0f;
    v3f p = add3(o, mul3(d, t));
    v3f dd = sub3(p, v0);
    float da = dot3(dd, mk3(f.a[0], f.a[1], f.a[2]));
    float db = dot3(dd, mk3(f.b[0], f.b[1], f.b[2]));
    if (da > 0.0f && da < f.aa && db > 0.0f && db < f.bb) return t;
    return -1.0f;
}

// nearest periodic face strictly below tlim (periodic_boundary.glsl:84-92); returns its t or -1
__device__ __forceinline__ float nearest_face(const DevScene &sc, v3f o, v3f d, float tlim) {
    float bt = tlim;
    int cf = -1;
    for (uint32_t f = 0; f < sc.nbbox; f++) {
        float t = isect_bbox_face(sc.bbox[f], o, d);
        if (t > 0.0f && t < bt) { bt = t; cf = (int)f; }
    }
    return cf < 0 ? -1.0f : bt;
}

// wrapped origin for a face hit at distance bt (periodic_boundary.glsl:94-111)
__device__ __forceinline__ v3f wrap_at(const DevScene &sc, v3f o, v3f d, float bt) {
    v3f hp = add3(o, mul3(d, bt));
    const float eps = 1e-5f;
    float wx = sc.dom[1] - sc.dom[0], wy = sc.dom[3] - sc.dom[2];
    v3f w = hp;
    if (sc.px && fabsf(hp.x - sc.dom[0]) <= eps) w.x += wx - eps;
    else if (sc.px && fabsf(hp.x - sc.dom[1]) <= eps) w.x -= wx - eps;
    else if (sc.py && fabsf(hp.y - sc.dom[2]) <= eps) w.y += wy - eps;
    else if (sc.py && fabsf(hp.y - sc.dom[3]) <= eps) w.y -= wy - eps;
    return w;
}

// Closest-hit trace with up to 32 periodic wraps (diffuse_raygen.comp:498-560 without translucent
// covers).  Returns the last traversal result.
template <int STK, typename C>
__device__ __forceinline__ uint32_t trace_loop(const DevScene &sc, v3f o, v3f d, float tmin, uint32_t skip,
                                               bool &exhausted, uint32_t *stk, uint32_t stride, int *err, C &cnt) {
    uint32_t hit = RT_MISS;
    v3f cur = o;
    exhausted = true;
    for (int step = 0; step < 32; ++step) {
        float t;
        uint32_t slot;
        hit = trace_closest<false, STK>(sc, cur, d, tmin, skip, 1e30f, t, slot, stk, stride, err, cnt);
        if (sc.nbbox) {
            float bt = nearest_face(sc, cur, d, t);
            if (bt > 0.f) { cur = wrap_at(sc, cur, d, bt); continue; }
        }
        exhausted = false;
        break;
    }
    return hit;
}

// Occlusion towards a collimated source with periodic wraps (direct_raygen.comp:603-672 without
// translucent covers).  The reference wraps when the nearest periodic face is strictly closer than
// the closest hit, and is occluded otherwise if anything was hit; "some hit with t <= t_face" is the
// same decision, so an any-hit query suffices.  32 wraps without resolution count as unoccluded.
template <int STK, typename C>
__device__ __forceinline__ bool occluded(const DevScene &sc, v3f o, v3f d, float tmin, uint32_t skip, uint32_t *stk,
                                         uint32_t stride, int *err, C &cnt) {
    v3f cur = o;
    for (int step = 0; step < 32; ++step) {
        float tf = sc.nbbox ? nearest_face(sc, cur, d, 1e30f) : -1.0f;
        float tmax = tf > 0.f ? tf : 1e30f;
        if (trace_any<STK>(sc, cur, d, tmin, skip, tmax, stk, stride, err, cnt)) return true;
        if (tf <= 0.f) return false;
        cur = wrap_at(sc, cur, d, tf);
    }
    return false;
}

// occluded() for a whole wave with packet_any4() (packet_any() when the scene has no usable BVH4); each
// lane keeps its own wrap state.
template <int STK, typename C>
__device__ __forceinline__ bool packet_occluded(const DevScene &sc, v3f o, v3f d, float tmin, uint32_t skip,
                                                bool active, uint32_t *wstk, uint32_t *wstk4, int *err, C &cnt) {
    v3f cur = o;
    bool need = active, occl = false;
    for (int step = 0; step < 32; ++step) {
        if (__ballot(need) == 0) break;
        float tf = (need && sc.nbbox) ? nearest_face(sc, cur, d, 1e30f) : -1.0f;
        float tmax = tf > 0.f ? tf : 1e30f;
        bool f = sc.node4 ? packet_any4(sc, cur, d, tmin, skip, tmax, need, wstk4, err, cnt)
                          : packet_any<STK>(sc, cur, d, tmin, skip, tmax, need, wstk, err, cnt);
        if (need) {
            if (f) { occl = true; need = false; }
            else if (tf <= 0.f) need = false;
            else cur = wrap_at(sc, cur, d, tf);
        }
    }
    return occl;
}

// Next work item of this wave.  No lane-divergent branch around the fetch: every lane issues the
// atomic (lane 0 adds 1, the others 0) and lane 0's old value is broadcast as a scalar.  A
// `if (lane == 0) atomicAdd` + shuffle lets the compiler split the persistent loop into a lane-0
// loop and a lanes-1..63 loop that re-processes item 0 forever.
__device__ __forceinline__ uint32_t wave_item(uint32_t *counter, uint32_t chunk) {
    uint32_t lane = threadIdx.x & 63;
    uint32_t item = atomicAdd(counter, lane == 0 ? chunk : 0u);
    return __builtin_amdgcn_readfirstlane(item);
}

template <typename J>
__device__ __forceinline__ uint32_t find_job(const J *jobs, uint32_t njobs, uint32_t item) {
    uint32_t lo = 0, hi = njobs;  // last job with item_begin <= item
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].item_begin <= item) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int off = 32; off >= 1; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
    return v;
}

// Sum of k rays of strength s, each added once: k s is exact in double (k < 2^24), then rounded to float.
__device__ __forceinline__ float lit_sum(uint32_t k, float s) { return (float)((double)k * (double)s); }

__device__ __forceinline__ void wave_reduce(ksum &k) {
    for (int off = 32; off >= 1; off >>= 1) {
        float os = __shfl_xor(k.s, off, 64);
        float oc = __shfl_xor(k.c, off, 64);
        k.merge(os, oc);
    }
}

__device__ __forceinline__ bool traced_type(uint32_t t) { return t == 0 || t == 1 || t == 3; }

template <int MK>
__device__ __forceinline__ void add_stats(unsigned long long *, const NoCntT<MK> &, uint32_t) {}
template <int MK>
__device__ __forceinline__ void add_stats(unsigned long long *stats, const TCntT<MK> &c, uint32_t rays) {
    if (stats) {
        atomicAdd(&stats[0], (unsigned long long)c.nodes);
        atomicAdd(&stats[1], (unsigned long long)c.prims);
        atomicAdd(&stats[2], (unsigned long long)rays);
    }
}

// Occlusion of one shadow ray through the source's grid: test the primitives binned in the cell of the
// origin's projection with the BVH traversal's predicate (t in (1e-4, 1e30], not the origin primitive) and
// stop at the first hit.  Non-periodic scenes only (the grid is not built otherwise).
// Candidate range [b, e) of the grid cell holding origin o's projection.
__device__ __forceinline__ void grid_cell(const SunGrid &G, v3f o, uint32_t &b, uint32_t &e) {
    const float qu = o.x * G.u[0] + o.y * G.u[1] + o.z * G.u[2];
    const float qv = o.x * G.v[0] + o.y * G.v[1] + o.z * G.v[2];
    float fx = floorf((qu - G.g0u) * G.inv_cs), fy = floorf((qv - G.g0v) * G.inv_cs);
    const int cx = fx < 0.f ? 0 : (fx > (float)(G.nx - 1) ? (int)(G.nx - 1) : (int)fx);
    const int cy = fy < 0.f ? 0 : (fy > (float)(G.ny - 1) ? (int)(G.ny - 1) : (int)fy);
    const uint32_t c = (uint32_t)cy * G.nx + (uint32_t)cx;
    b = G.start[c] - G.sbase;
    e = G.start[c + 1] - G.sbase;
}

// Candidates p and p + 1 (when p + 1 < e) of a cell list: is either accepted as an occluder?  Both
// candidates' loads are issued before either test (the list entry -> record chain is latency-bound); a slot
// that is out of the list or unfilled loads record 0 and is ignored.  An entry is one 8-byte load (slot and
// key), and a record's fourth float4 (the patch plane normal, the mask descriptor) is loaded only for the
// entries flagged as needing it, so a triangle candidate costs four vector loads instead of six.  A
// candidate whose depth key (an upper bound of the depth of its points along the source direction, padded
// by 1e-5 of the scene's largest coordinate) is below the ray origin's depth lies wholly behind the origin:
// no FP32 hit of it can have t > 1e-4 (the FP32 errors of t, of the key and of the depth are a few ulps of
// that coordinate), so it is skipped without loading its record.
// Returns 1 when a candidate occludes the ray, 2 when the list is known to hold no further occluder (a
// sorted list whose next key is below the origin's depth), else 0.
template <typename C>
__device__ __forceinline__ int grid_pair(const uint2 *ents, bool sorted, float depth, const float4 *recs,
                                         uint32_t nleaf, uint32_t p, uint32_t e, v3f o, v3f d, uint32_t skip,
                                         C &cnt) {
    const bool two = p + 1 < e;
    const uint2 e0 = ents[p];
    const uint2 e1 = two ? ents[p + 1] : make_uint2(RT_MISS, __float_as_uint(-INFINITY));
    const float key0 = __uint_as_float(e0.y), key1 = __uint_as_float(e1.y);
    const uint32_t k0 = e0.x & RT_GRID_SLOT, k1 = e1.x & RT_GRID_SLOT;
    const bool v0 = e0.x != RT_MISS && k0 < nleaf && !(key0 < depth);
    const bool v1 = e1.x != RT_MISS && k1 < nleaf && !(key1 < depth);
    const float4 *q0 = recs + (size_t)(v0 ? k0 : 0u) * 4, *q1 = recs + (size_t)(v1 ? k1 : 0u) * 4;
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 a0 = ldc(q0), a1 = ldc(q0 + 1), a2 = ldc(q0 + 2);
    const float4 a3 = (v0 && (e0.x & RT_GRID_REC3)) ? ldc(q0 + 3) : zero;
    const float4 b0 = ldc(q1), b1 = ldc(q1 + 1), b2 = ldc(q1 + 2);
    const float4 b3 = (v1 && (e1.x & RT_GRID_REC3)) ? ldc(q1 + 3) : zero;
    const uint32_t ta = __float_as_uint(a0.w), tb = __float_as_uint(b0.w);
    float t0, t1, ua0, ub0, ua1, ub1;
    bool h0 = v0 && isect_rec_uv<false, C::wt>(a0, a1, a2, a3, (ta & RT_PATCH_BIT) != 0, o, d, t0, ua0, ub0) &&
              (ta & RT_ID_BITS) != skip && t0 > 1e-4f && t0 <= 1e30f;
    bool h1 = v1 && isect_rec_uv<false, C::wt>(b0, b1, b2, b3, (tb & RT_PATCH_BIT) != 0, o, d, t1, ua1, ub1) &&
              (tb & RT_ID_BITS) != skip && t1 > 1e-4f && t1 <= 1e30f;
    if (C::mask && h0 && (ta & RT_MASK_BIT))
        h0 = mask_hit_opaque(recs, __float_as_uint(a3.w), (ta & RT_PATCH_BIT) != 0, ua0, ub0, a1.w, a2.w);
    if (C::mask && h1 && (tb & RT_MASK_BIT))
        h1 = mask_hit_opaque(recs, __float_as_uint(b3.w), (tb & RT_PATCH_BIT) != 0, ua1, ub1, b1.w, b2.w);
    cnt.prims += (v0 ? 1u : 0u) + ((v1 && !h0) ? 1u : 0u);  // the tests a one-at-a-time scan makes
    if (h0 || h1) return 1;
    return (sorted && (!(key0 >= depth) || (two && !(key1 >= depth)))) ? 2 : 0;
}

// ------------------------------------------------------------------------------------------------
// direct rays
// ------------------------------------------------------------------------------------------------
// COUNT: the counting variant (stats != nullptr) traces every ray on its own (per-lane LDS stack) so the
// counters hold single-ray node visits and primitive tests -- the algorithmic work per ray -- rather than
// the packet's visits.  Both variants make the same occlusion decision for every ray.
#ifndef RT_DIRECT_REFILL
#define RT_DIRECT_REFILL 24
#endif
// Origin of direct ray r of primitive i (stratified sample, then the texture-mask rejection when the
// primitive has a mask: mdi = its descriptor, RT_MISS = none); false when the ray is dropped.
__device__ __forceinline__ bool direct_origin(const float4 *recs, uint32_t mdi, uint32_t i, uint32_t r, uint32_t d,
                                              uint32_t seed, bool patch, float &spx, float &spy) {
    uint32_t st;
    direct_sample_st(i, r % d, r / d, d, seed, patch, spx, spy, st);
    return mdi == RT_MISS || origin_mask(recs, mdi, patch, false, st, spx, spy);
}

// Waves per SIMD the compiler must fit the unmasked direct kernel into: 7 (72 VGPRs, 6 without the bound)
// takes the orchard's 64-sun direct pass from 407 to 393 ms; 8 spills in the candidate loop (419 ms)
#ifndef RT_DIRECT_WAVES_PER_EU
#define RT_DIRECT_WAVES_PER_EU 7
#endif
template <int WPB, int STK, bool COUNT, int MK>
__global__ __launch_bounds__(WPB * 64, (COUNT || MK) ? 1 : RT_DIRECT_WAVES_PER_EU) void k_direct(const DirectJob *__restrict__ jobs, uint32_t njobs, uint32_t total,
                                                     uint32_t *counter, uint32_t chunk, int *err,
                                                     unsigned long long *stats) {
    __shared__ uint32_t wstk_all[COUNT ? STK * WPB * 64 : STK * WPB];
    __shared__ uint32_t wstk4_all[COUNT ? 1 : RT_WSTK4 * WPB];
    uint32_t *wstk = COUNT ? wstk_all + threadIdx.x : wstk_all + (threadIdx.x >> 6) * STK;
    uint32_t *wstk4 = COUNT ? wstk4_all : wstk4_all + (threadIdx.x >> 6) * RT_WSTK4;
    const uint32_t lane = threadIdx.x & 63;
    Cnt<COUNT, MK> cnt{};
    uint32_t nray = 0;
    for (uint32_t first = wave_item(counter, chunk); first < total; first = wave_item(counter, chunk)) {
        const uint32_t end = min(first + chunk, total);
        for (uint32_t item = first; item < end; ++item) {
            // job, scene and primitive data are wave-uniform: scalar loads keep them out of VGPRs
            uint32_t j = find_job(jobs, njobs, item);
            const DirectJob &J = jobs[j];
            const DevScene &sc = *ldu(&J.sc);
            uint32_t i = item - ldu(&J.item_begin);
            const uint32_t *jorder = ldu(&J.order);
            if (jorder) i = ldu(jorder + i);
            i = __builtin_amdgcn_readfirstlane(i);
            uint32_t ty = ldu(ldu(&sc.type) + i);
            float sum_out = 0.f;
            bool top = true;
            const SunGrid *grid = ldu(&J.grid);
            const float4 *recs = ldu(&sc.rec);
            const uint32_t *mdesc = ldu(&sc.mdesc);
            const uint32_t mdi = (MK && mdesc) ? ldu(mdesc + i) : RT_MISS;
            if (traced_type(ty)) {
                float4 n4 = ldu(ldu(&sc.nrm) + i);
                v3f n = mk3(n4.x, n4.y, n4.z);
                uint32_t tw = __float_as_uint(n4.w);
                v3f dir = nrm3(mk3(ldu(&J.sdir[0]), ldu(&J.sdir[1]), ldu(&J.sdir[2])));
                uint32_t d = ldu(&J.d), Nrays = d * d;
                float strength = rdiv(fabsf(dot3(n, dir)), (float)Nrays);
                top = dot3(dir, n) > 0.0f;
                if (!(tw == 3 || (!top && tw == 0))) {
                    bool patch = ty == 0 || ty == 3;
                    float tm[12];
                    const float *Tp = ldu(&sc.T) + (size_t)i * 16;
#pragma unroll
                    for (int q = 0; q < 12; q++) tm[q] = ldu(Tp + q);
                    // every ray of the primitive carries the same strength: the lit rays are counted and the
                    // sum is count x strength, rounded once (exact up to that rounding)
                    uint32_t lit = 0;
                    const uint32_t seed = ldu(&J.seed);
                    if (grid) {
                        // Occlusion grid: each lane scans its ray's cell list two candidates per step.  A lane
                        // whose ray ends (occluder found, or list exhausted: lit) takes the primitive's next
                        // ray; idle lanes are refilled together once fewer than RT_DIRECT_REFILL lanes still
                        // scan (every ray is tested exactly as on its own; only the ray -> lane assignment
                        // and the summation grouping differ).
                        const SunGrid &G = *grid;
                        const uint2 *ents = G.ents;
                        const bool sorted = G.sorted != 0;
                        const uint32_t nleaf = ldu(&sc.nleaf);
                        const float gw0 = G.w[0], gw1 = G.w[1], gw2 = G.w[2];
                        uint32_t r = lane, pc = 0, pe = 0;
                        v3f o = mk3(0.f, 0.f, 0.f);
                        float depth = 0.f;  // origin depth along the source direction
                        bool act = r < Nrays;
                        if (act) {
                            float spx, spy;
                            act = direct_origin(recs, mdi, i, r, d, seed, patch, spx, spy);
                            o = xform3(tm, spx, spy, 0.0f);
                            grid_cell(G, o, pc, pe);
                            depth = o.x * gw0 + o.y * gw1 + o.z * gw2;
                            cnt.nodes++;
                            nray++;
                        }
                        uint32_t qnext = min(64u, Nrays);
                        while (true) {
                            const uint32_t floor_lanes = qnext < Nrays ? RT_DIRECT_REFILL : 1u;
                            while (true) {
                                if (act) {
                                    const int res = pc >= pe ? 2
                                                             : grid_pair(ents, sorted, depth, recs, nleaf, pc, pe,
                                                                         o, dir, i, cnt);
                                    if (res == 2) lit++;
                                    if (res != 0) act = false;
                                    else pc += 2;
                                }
                                if ((uint32_t)__popcll(__ballot(act)) < floor_lanes) break;
                            }
                            const uint64_t idle = __ballot(!act);
                            if (qnext >= Nrays) {
                                if (idle == ~0ull) break;
                                continue;
                            }
                            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                            const uint32_t take = min((uint32_t)__popcll(idle), Nrays - qnext);
                            if (!act && rank < take) {
                                r = qnext + rank;
                                float spx, spy;
                                act = direct_origin(recs, mdi, i, r, d, seed, patch, spx, spy);
                                o = xform3(tm, spx, spy, 0.0f);
                                grid_cell(G, o, pc, pe);
                                depth = o.x * gw0 + o.y * gw1 + o.z * gw2;
                                cnt.nodes++;
                                nray++;
                            }
                            qnext += take;
                        }
                    } else {
                        for (uint32_t base = 0; base < Nrays; base += 64) {
                            uint32_t r = base + lane;
                            bool need = r < Nrays;
                            v3f o = mk3(0.f, 0.f, 0.f);
                            if (need) {
                                float spx, spy;
                                need = direct_origin(recs, mdi, i, r, d, seed, patch, spx, spy);
                                o = xform3(tm, spx, spy, 0.0f);
                                nray++;
                            }
                            bool occl;
                            if (COUNT) occl = need && occluded<STK>(sc, o, dir, 1e-4f, i, wstk, WPB * 64, err, cnt);
                            else occl = packet_occluded<STK>(sc, o, dir, 1e-4f, i, need, wstk, wstk4, err, cnt);
                            if (need && !occl) lit++;
                        }
                    }
                    sum_out = lit_sum(wave_sum_u32(lit), strength);
                }
            }
            if (lane == 0) {
                float *out = ldu(&J.out);
                out[(size_t)i * 2 + 0] = top ? sum_out : 0.f;
                out[(size_t)i * 2 + 1] = top ? 0.f : sum_out;
            }
        }
    }
    add_stats(stats, cnt, nray);
}

// Collimated direct rays when every job of the launch has an occlusion grid.  The lanes of a wave are refilled
// across the primitives of a job, as the diffuse kernel refills them: a lane whose ray ends takes the next
// ray of the open primitive, or of the next primitive of the wave's work queue, so no lane idles until the
// longest ray of one primitive ends.  Each primitive with rays in flight holds a slot of the wave's ring in
// LDS that counts its lit rays; the primitive's sums are written once its last ray has ended.  A job change
// (another direction, grid and output) waits until the rays in flight have ended.  Every ray is tested
// exactly as in k_direct, and the sums are the same count x strength.
#define RT_DQ_SLOTS 8
template <int WPB, int MK>
__global__ __launch_bounds__(WPB * 64, MK ? 1 : RT_DIRECT_WAVES_PER_EU) void k_direct_q(const DirectJob *__restrict__ jobs,
                                                                                        uint32_t njobs, uint32_t total,
                                                                                        uint32_t *counter, uint32_t chunk) {
    __shared__ uint32_t ring_all[3 * RT_DQ_SLOTS * WPB];  // per wave: lit count, primitive, strength of each slot
    uint32_t *lit = ring_all + (threadIdx.x >> 6) * 3 * RT_DQ_SLOTS;
    uint32_t *sprim = lit + RT_DQ_SLOTS;
    float *sstr = (float *)(lit + 2 * RT_DQ_SLOTS);
    const uint32_t lane = threadIdx.x & 63;
    Cnt<false, MK> cnt{};
    // work queue (wave-uniform)
    uint32_t it = 0, it_end = 0;
    bool drained = false;
    // job (wave-uniform): loaded when the wave first takes one of its items
    uint32_t jcur = RT_MISS, jbegin = 0, jend = 0, d = 0, Nrays = 0, seed = 0, nleaf = 0;
    const DevScene *scp = nullptr;
    const SunGrid *grid = nullptr;
    const uint32_t *order = nullptr;
    float *out = nullptr;
    v3f dir = mk3(0.f, 0.f, 0.f);
    // open primitive (wave-uniform) and the slots in use (bit mask; a slot's bit 8 + k: its sums are lit from
    // the top)
    uint32_t op_i = 0, op_next = 0, op_cnt = 0, op_slot = 0, op_mdi = RT_MISS, slots = 0;
    bool op_patch = false;
    // lane: its ray (origin, depth along the source direction, cell list position), primitive and slot
    bool act = false;
    v3f o = mk3(0.f, 0.f, 0.f);
    float depth = 0.f;
    uint32_t pc = 0, pe = 0, my_i = 0, my_slot = 0;
    while (true) {
        bool blocked = false;
        while (true) {
            const uint64_t idle = __ballot(!act);
            if (idle == 0) break;
            if (op_next == op_cnt) {
                // the open primitive has no rays left: open the next item of the queue
                if (it == it_end) {
                    if (drained) break;
                    it = wave_item(counter, chunk);
                    if (it >= total) { drained = true; it = it_end = 0; break; }
                    it_end = min(it + chunk, total);
                }
                if (jcur == RT_MISS || it < jbegin || it >= jend) {
                    if (__ballot(act) != 0 || (slots & 0xFFu) != 0) { blocked = true; break; }
                    jcur = find_job(jobs, njobs, it);
                    const DirectJob &J = jobs[jcur];
                    jbegin = ldu(&J.item_begin);
                    jend = jcur + 1 < njobs ? ldu(&jobs[jcur + 1].item_begin) : total;
                    scp = ldu(&J.sc);
                    grid = ldu(&J.grid);
                    order = ldu(&J.order);
                    out = ldu(&J.out);
                    d = ldu(&J.d);
                    Nrays = d * d;
                    seed = ldu(&J.seed);
                    dir = nrm3(mk3(ldu(&J.sdir[0]), ldu(&J.sdir[1]), ldu(&J.sdir[2])));
                    nleaf = ldu(&scp->nleaf);
                }
                if ((slots & 0xFFu) == 0xFFu) { blocked = true; break; }
                uint32_t i = it - jbegin;
                if (order) i = ldu(order + i);
                i = __builtin_amdgcn_readfirstlane(i);
                it++;
                const uint32_t ty = ldu(ldu(&scp->type) + i);
                const float4 n4 = ldu(ldu(&scp->nrm) + i);
                const v3f n = mk3(n4.x, n4.y, n4.z);
                const uint32_t tw = __float_as_uint(n4.w);
                const bool top = dot3(dir, n) > 0.0f;
                if (!traced_type(ty) || tw == 3 || (!top && tw == 0)) {
                    if (lane == 0) { out[(size_t)i * 2] = 0.f; out[(size_t)i * 2 + 1] = 0.f; }
                    continue;
                }
                const uint32_t sl = (uint32_t)__builtin_ctz(~slots & 0xFFu);
                slots |= (1u << sl) | (top ? (0x100u << sl) : 0u);
                lit[sl] = 0u;
                sprim[sl] = i;
                sstr[sl] = rdiv(fabsf(dot3(n, dir)), (float)Nrays);
                const uint32_t *mdesc = ldu(&scp->mdesc);
                op_mdi = (MK && mdesc) ? ldu(mdesc + i) : RT_MISS;
                op_i = i;
                op_next = 0;
                op_cnt = Nrays;
                op_slot = sl;
                op_patch = ty == 0 || ty == 3;
                continue;
            }
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            const uint32_t take = min((uint32_t)__popcll(idle), op_cnt - op_next);
            if (!act && rank < take) {
                const uint32_t r = op_next + rank;
                const float4 *recs = ldu(&scp->rec);
                float spx, spy;
                act = direct_origin(recs, op_mdi, op_i, r, d, seed, op_patch, spx, spy);
                float tm[12];
                const float *Tp = ldu(&scp->T) + (size_t)op_i * 16;
#pragma unroll
                for (int q = 0; q < 12; q++) tm[q] = ldu(Tp + q);
                o = xform3(tm, spx, spy, 0.0f);
                grid_cell(*grid, o, pc, pe);
                depth = o.x * grid->w[0] + o.y * grid->w[1] + o.z * grid->w[2];
                my_i = op_i;
                my_slot = op_slot;
            }
            op_next += take;
        }
        // write the sums of the primitives none of whose rays remain (the open one once all its rays are out)
        for (uint32_t pend = slots & 0xFFu; pend != 0u; pend &= pend - 1u) {
            const uint32_t sl = (uint32_t)__builtin_ctz(pend);
            if (sl == op_slot && op_next < op_cnt) continue;
            if (__ballot(act && my_slot == sl) != 0) continue;
            const uint32_t i = sprim[sl];
            const float sum = lit_sum(lit[sl], sstr[sl]);
            const bool top = (slots >> (8 + sl)) & 1u;
            if (lane == 0) {
                out[(size_t)i * 2] = top ? sum : 0.f;
                out[(size_t)i * 2 + 1] = top ? 0.f : sum;
            }
            slots &= ~((1u << sl) | (0x100u << sl));
        }
        if (__ballot(act) == 0) {
            if (drained && op_next == op_cnt && (slots & 0xFFu) == 0u) break;
            continue;
        }
        // trace until fewer than RT_DIRECT_REFILL lanes scan (all of them when no ray can be handed out)
        const uint32_t floor_lanes = (blocked || (drained && op_next == op_cnt)) ? 1u : RT_DIRECT_REFILL;
        const uint2 *ents = grid->ents;
        const bool sorted = grid->sorted != 0;
        const float4 *recs = ldu(&scp->rec);
        while (true) {
            if (act) {
                const int res = pc >= pe ? 2 : grid_pair(ents, sorted, depth, recs, nleaf, pc, pe, o, dir, my_i, cnt);
                if (res == 2) atomicAdd(&lit[my_slot], 1u);
                if (res != 0) act = false;
                else pc += 2;
            }
            if ((uint32_t)__popcll(__ballot(act)) < floor_lanes) break;
        }
    }
}

// Occlusion of a ray towards a non-collimated source, literally as the reference loop
// (direct_raygen.comp:603-628 without glass covers): closest hit, periodic wrap when a face is strictly
// nearer, and a hit at or beyond the distance from the original origin to the source centre does not block.
template <int STK, typename C>
__device__ __forceinline__ bool source_occluded(const DevScene &sc, v3f o, v3f d, uint32_t skip, float dist0,
                                                uint32_t *stk, uint32_t stride, int *err, C &cnt) {
    v3f cur = o;
    for (int step = 0; step < 32; ++step) {
        float t;
        uint32_t slot;
        uint32_t h = trace_closest<false, STK>(sc, cur, d, 1e-4f, skip, 1e30f, t, slot, stk, stride, err, cnt);
        if (sc.nbbox) {
            float bt = nearest_face(sc, cur, d, t);
            if (bt > 0.f) { cur = wrap_at(sc, cur, d, bt); continue; }
        }
        if (h == RT_MISS) return false;
        return t < dist0;
    }
    return false;
}

// Direct rays towards sphere, sun-sphere, rectangle and disk sources (direct_raygen.comp:482-561): one
// wave per (job, primitive), d*d stratified origins as k_direct, per-ray direction, strength and lit face,
// per-lane closest-hit traversal.  Items [item_lo, item_lo + nitems) of jobs whose item_begin ascend.
template <int STK, bool COUNT, int MK>
__global__ __launch_bounds__(256) void k_direct_src(const DirectJob *__restrict__ jobs, uint32_t njobs, uint32_t item_lo,
                                                    uint32_t nitems, uint32_t *counter, int *err,
                                                    unsigned long long *stats) {
    __shared__ uint32_t stk_all[STK * 256];
    uint32_t *stk = stk_all + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    Cnt<COUNT, MK> cnt{};
    uint32_t nray = 0;
    for (uint32_t first = wave_item(counter, 1u); first < nitems; first = wave_item(counter, 1u)) {
        const uint32_t item = item_lo + first;
        const DirectJob &J = jobs[find_job(jobs, njobs, item)];
        const DevScene &sc = *J.sc;
        uint32_t i = item - J.item_begin;
        if (J.order) i = J.order[i];
        const uint32_t ty = sc.type[i];
        ksum top_acc, bot_acc;
        top_acc.init();
        bot_acc.init();
        if (traced_type(ty)) {
            const float4 n4 = sc.nrm[i];
            const v3f n = mk3(n4.x, n4.y, n4.z);
            const uint32_t tw = __float_as_uint(n4.w);
            const bool patch = ty == 0 || ty == 3;
            const uint32_t d = J.d, Nrays = d * d;
            float tm[12];
            const float *Tp = sc.T + (size_t)i * 16;
#pragma unroll
            for (int q = 0; q < 12; q++) tm[q] = Tp[q];
            const v3f sp3 = mk3(J.sdir[0], J.sdir[1], J.sdir[2]);
            for (uint32_t base = 0; base < Nrays; base += 64) {
                const uint32_t r = base + lane;
                if (r >= Nrays) continue;
                float spx, spy;
                uint32_t st;
                direct_sample_st(i, r % d, r / d, d, J.seed, patch, spx, spy, st);
                const uint32_t mdi = (MK && sc.mdesc) ? sc.mdesc[i] : RT_MISS;
                if (MK && mdi != RT_MISS && !origin_mask(sc.rec, mdi, patch, false, st, spx, spy)) continue;
                const v3f o = xform3(tm, spx, spy, 0.0f);
                for (uint32_t k = 0; k < J.skip_draws; k++) (void)rnd24(st);
                nray++;
                v3f dir;
                float strength;
                if (!source_ray(J.stype, J.sdir, J.width, J.rot, o, n, Nrays, st, dir, strength)) continue;
                const bool top = dot3(dir, n) > 0.0f;
                if (tw == 3 || (!top && tw == 0)) continue;
                const v3f dv = sub3(sp3, o);
                const float dist0 = rsqrt_cr(dot3(dv, dv));
                if (source_occluded<STK>(sc, o, dir, i, dist0, stk, 256u, err, cnt)) continue;
                if (top) top_acc.add(strength);
                else bot_acc.add(strength);
            }
        }
        wave_reduce(top_acc);
        wave_reduce(bot_acc);
        if (lane == 0) {
            J.out[(size_t)i * 2 + 0] = top_acc.value();
            J.out[(size_t)i * 2 + 1] = bot_acc.value();
        }
    }
    add_stats(stats, cnt, nray);
}

// Resumable closest-hit traversal state of one lane (BVH2), for kernels that refill finished lanes.
struct LaneRay {
    v3f o, d;
    RayPre rp;
    float best_t;
    uint32_t best, ni, budget;
    int sp;
};

__device__ __forceinline__ void lane_start(LaneRay &L, uint32_t nleaf) {
    L.rp = ray_pre(L.o, L.d);
    L.best_t = 1e30f;
    L.best = RT_MISS;
    L.ni = 0;
    L.sp = 0;
    L.budget = nleaf + 8;
}

// Starts the closest-hit trace of L.o, L.d; returns false when it is already complete (a scene with at
// most one traced primitive needs no node steps), as trace_closest() with tstart = 1e30.
template <typename C>
__device__ __forceinline__ bool lane_begin(LaneRay &L, const float4 *recs, uint32_t nleaf, float tmin, uint32_t skip,
                                           C &cnt) {
    lane_start(L, nleaf);
    if (nleaf > 1) return true;
    uint32_t slot_unused;
    if (nleaf == 1) { cnt.prims++; test_leaf<false, C::flags>(recs, 0, L.o, L.d, tmin, skip, L.best_t, L.best, slot_unused); }
    return false;
}

// LDS pointer: accesses compile to ds_read/ds_write (a generic pointer chosen between LDS and private
// memory would compile to slower flat accesses that also wait on vector memory)
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) uint32_t lds_u32;
#else
typedef uint32_t lds_u32;
#endif

// Traversal entries are inner nodes or leaves (RT_LEAF_BIT).  The stack holds its first LSTK entries in
// LDS (stk, lane stride `stride`) and the rest, up to STK, in the lane's private array ovf (deep stacks
// are rare; LDS per wave bounds the occupancy).  lane_pop() returns true when the stack is empty.
template <int STK, int LSTK>
__device__ __forceinline__ bool lane_pop(LaneRay &L, const lds_u32 *stk, const uint32_t *ovf, uint32_t stride) {
    if (L.sp == 0) return true;
    L.sp--;
    if (LSTK >= STK || L.sp < LSTK) L.ni = stk[(uint32_t)L.sp * stride];
    else L.ni = ovf[L.sp - LSTK];
    return false;
}

// Visit the inner node L.ni: both child boxes against [0, best_t]; the nearer hit child (inner or leaf)
// becomes the next entry and the farther one is pushed.  Returns true when the traversal has ended.
template <int STK, int LSTK, typename C>
__device__ __forceinline__ bool lane_node(const float4 *nodes, uint32_t nleaf, LaneRay &L, lds_u32 *stk, uint32_t *ovf,
                                          uint32_t stride, int *err, C &cnt) {
    if (L.budget-- == 0 || L.ni >= nleaf - 1) {
        atomicOr(err, 2);
        return true;
    }
    cnt.nodes++;
    NodeTest nt = node_test_p(nodes, L.ni, L.rp, L.best_t);
    if (nt.h0 && nt.h1) {
        bool first0 = nt.tn0 <= nt.tn1;
        if (L.sp >= STK) {
            atomicOr(err, 1);
            return true;
        }
        const uint32_t far = first0 ? nt.c1 : nt.c0;
        if (LSTK >= STK || L.sp < LSTK) stk[(uint32_t)L.sp * stride] = far;
        else ovf[L.sp - LSTK] = far;
        L.sp++;
        L.ni = first0 ? nt.c0 : nt.c1;
    } else if (nt.h0) {
        L.ni = nt.c0;
    } else if (nt.h1) {
        L.ni = nt.c1;
    } else {
        return lane_pop<STK, LSTK>(L, stk, ovf, stride);
    }
    return false;
}

// The hit children of a BVH4 visit (key = entry distance, +inf = missed): the nearest becomes the next
// entry and the others are pushed (farthest first with RT_BVH4_SORT, else in slot order); the stack empty
// after no hit ends the traversal (returns true), as does a stack overflow (error bit 1).
template <int STK, int LSTK>
__device__ __forceinline__ bool node4_order(LaneRay &L, float *key, uint32_t *lnk, uint32_t nhit, lds_u32 *stk,
                                            uint32_t *ovf, uint32_t stride, int *err) {
    if (nhit == 0) return lane_pop<STK, LSTK>(L, stk, ovf, stride);
#if RT_BVH4_SORT
    // ascending sort of the four (key, link) pairs; missed children (key +inf) end up last
    cswap(key[0], lnk[0], key[1], lnk[1]);
    cswap(key[2], lnk[2], key[3], lnk[3]);
    cswap(key[0], lnk[0], key[2], lnk[2]);
    cswap(key[1], lnk[1], key[3], lnk[3]);
    cswap(key[1], lnk[1], key[2], lnk[2]);
#if !RT_Q4_LATE_ERR
    if ((uint32_t)L.sp + nhit - 1 > (uint32_t)STK) {
        atomicOr(err, 1);
        return true;
    }
#endif
#if RT_PUSH_SPARE
    // (with the stack bound tested after this path, the path needs sp + 3 <= LSTK even when LSTK >= STK)
    if ((!RT_Q4_LATE_ERR && LSTK >= STK) || (uint32_t)L.sp + 3u <= (uint32_t)LSTK) {
        // every slot a push may use is in LDS: three stores without branches, those beyond the pushed
        // children aimed at the lane's spare row (index LSTK, which nothing reads); child k < nhit goes to
        // slot sp + nhit - 1 - k, so the farthest lies deepest and child 1 on top
#pragma unroll
        for (int k = 3; k >= 1; k--) {
            const uint32_t at = (uint32_t)k < nhit ? (uint32_t)L.sp + nhit - 1u - (uint32_t)k : (uint32_t)LSTK;
            stk[at * stride] = lnk[k];
        }
        L.sp += (int)nhit - 1;
        L.ni = lnk[0];
        return false;
    }
#endif
#if RT_Q4_LATE_ERR
    // (the branch-free pushes above cannot overflow: sp + 3 <= LSTK <= STK)
    if ((uint32_t)L.sp + nhit - 1 > (uint32_t)STK) {
        atomicOr(err, 1);
        return true;
    }
#endif
#pragma unroll
    for (int k = 3; k >= 1; k--) {
        if ((uint32_t)k < nhit) {
            if (LSTK >= STK || L.sp < LSTK) stk[(uint32_t)L.sp * stride] = lnk[k];
            else ovf[L.sp - LSTK] = lnk[k];
            L.sp++;
        }
    }
    L.ni = lnk[0];
    return false;
#else
    // the nearest hit child next; the other hit children pushed in slot order
    int m = 0;
    float km = key[0];
#pragma unroll
    for (int k = 1; k < 4; k++)
        if (key[k] < km) { km = key[k]; m = k; }
    if ((uint32_t)L.sp + nhit - 1 > (uint32_t)STK) {
        atomicOr(err, 1);
        return true;
    }
    uint32_t next = lnk[0];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k == m) next = lnk[k];
        else if (key[k] < INFINITY) {
            if (LSTK >= STK || L.sp < LSTK) stk[(uint32_t)L.sp * stride] = lnk[k];
            else ovf[L.sp - LSTK] = lnk[k];
            L.sp++;
        }
    }
    L.ni = next;
    return false;
#endif
}

// Visit the BVH4 node L.ni: four child boxes against [0, best_t]; the hit children (inner nodes or leaves)
// in ascending entry distance -- the nearest becomes the next entry, the others are pushed farthest first.
// Unused slots are +inf point boxes that never pass the test.  Returns true when the traversal has ended.
template <int STK, int LSTK, typename C>
__device__ __forceinline__ bool lane_node4(const float4 *nodes4, uint32_t nn4, LaneRay &L, lds_u32 *stk, uint32_t *ovf,
                                           uint32_t stride, int *err, C &cnt) {
    if (L.budget-- == 0 || L.ni >= nn4) {
        atomicOr(err, 2);
        return true;
    }
    cnt.nodes++;
    const float4 *nd = nodes4 + (size_t)L.ni * 8;
    const float4 lx = ldc(nd), hx = ldc(nd + 1), ly = ldc(nd + 2), hy = ldc(nd + 3);
    const float4 lz = ldc(nd + 4), hz = ldc(nd + 5), lk = ldc(nd + 6);
    float tn[4], tf[4];
    slab4(lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, L.rp, tn[0], tf[0]);
    slab4(lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, L.rp, tn[1], tf[1]);
    slab4(lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, L.rp, tn[2], tf[2]);
    slab4(lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, L.rp, tn[3], tf[3]);
    float key[4];
    uint32_t lnk[4] = {__float_as_uint(lk.x), __float_as_uint(lk.y), __float_as_uint(lk.z), __float_as_uint(lk.w)};
    uint32_t nhit = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool h = fmaxf(tn[k], 0.f) <= fminf(tf[k], L.best_t);
        key[k] = h ? tn[k] : INFINITY;
        nhit += h ? 1u : 0u;
    }
    return node4_order<STK, LSTK>(L, key, lnk, nhit, stk, ovf, stride, err);
}

// Visit the 8-bit BVH4 node L.ni (64 B: see k_q4 in rt_scene.hip).  A child bound q decodes in t-space as
// (o + q 2^e - ray origin) / dir = q (2^e / dir) + (o - ray origin) / dir: one convert and one fma per bound
// after a multiply and an fma per axis.  The decoded boxes contain the float boxes, and the t rounding is a
// few ulps of the scene scale, far inside the leaf boxes' padding, so no ray a float box passes is culled
// and the hits are those of the float trees.
__device__ __forceinline__ float qbyte(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xFFu); }

template <int STK, int LSTK, bool FLAT, typename C>
__device__ __forceinline__ bool lane_node4q(const float4 *q4, uint32_t nn4, LaneRay &L, lds_u32 *stk, uint32_t *ovf,
                                            uint32_t stride, int *err, C &cnt) {
#if RT_Q4_LATE_ERR
    // a spent step budget or a link outside the tree ends the trace without a branch: the read goes to
    // node 0 and the budget becomes RT_BUDGET_BAD, which the kernel reports when the lane finishes
    const bool bad = L.budget == 0u || L.ni >= nn4;
    L.budget = bad ? RT_BUDGET_BAD : L.budget - 1u;
    const uint32_t ni = bad ? 0u : L.ni;
#else
    if (L.budget-- == 0 || L.ni >= nn4) {
        atomicOr(err, 2);
        return true;
    }
    const uint32_t ni = L.ni;
    constexpr bool bad = false;
#endif
    cnt.nodes++;
    const float4 *nd = q4 + (size_t)ni * 4;
    const float4 f0 = ldc(nd), f1 = ldc(nd + 1), f2 = ldc(nd + 2), f3 = ldc(nd + 3);
    // FLAT: the entry a pop would take is read now, so its latency hides under the node's (a lane whose
    // stack top is not in LDS reads its spare row and takes the general path below)
    const uint32_t sp0 = (uint32_t)L.sp;
    uint32_t popv = 0u;
    if (FLAT) popv = stk[((sp0 != 0u && sp0 <= (uint32_t)LSTK) ? sp0 - 1u : (uint32_t)LSTK) * stride];
    const uint32_t meta = __float_as_uint(f0.w);
    const uint32_t nch = meta >> 24;
    // 2^e as a float from the biased exponent byte (e + 127 in [1, 254])
    const float sx = __uint_as_float((meta & 0xFFu) << 23), sy = __uint_as_float(((meta >> 8) & 0xFFu) << 23),
                sz = __uint_as_float(((meta >> 16) & 0xFFu) << 23);
    const float Ax = sx * L.rp.ix, Ay = sy * L.rp.iy, Az = sz * L.rp.iz;
    const float Bx = fmaf(f0.x, L.rp.ix, -L.rp.ox), By = fmaf(f0.y, L.rp.iy, -L.rp.oy),
                Bz = fmaf(f0.z, L.rp.iz, -L.rp.oz);
#if RT_Q4_EAGER
    // keep the link words' load beside the box words' (the compiler would otherwise sink it into the
    // branch that pushes, a second dependent round trip per step)
    asm volatile("" ::"v"(f2.z), "v"(f2.w), "v"(f3.x), "v"(f3.y));
#endif
    const uint32_t wlx = __float_as_uint(f1.x), whx = __float_as_uint(f1.y), wly = __float_as_uint(f1.z),
                   why = __float_as_uint(f1.w), wlz = __float_as_uint(f2.x), whz = __float_as_uint(f2.y);
    float key[4];
    uint32_t lnk[4] = {__float_as_uint(f2.z), __float_as_uint(f2.w), __float_as_uint(f3.x), __float_as_uint(f3.y)};
    uint32_t nhit = 0;
    // per axis, the bound word nearer the origin: the lower bounds when the direction is positive (A > 0:
    // t rises with q), else the upper ones.  t(near) = min(t(lo), t(hi)) exactly, so the entry and exit
    // distances equal the min/max form's with one select per axis instead of a min and a max per child.
    const bool sx_pos = L.rp.ix >= 0.f, sy_pos = L.rp.iy >= 0.f, sz_pos = L.rp.iz >= 0.f;
    const uint32_t nwx = sx_pos ? wlx : whx, fwx = sx_pos ? whx : wlx;
    const uint32_t nwy = sy_pos ? wly : why, fwy = sy_pos ? why : wly;
    const uint32_t nwz = sz_pos ? wlz : whz, fwz = sz_pos ? whz : wlz;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float tn = max3f(fmaf(qbyte(nwx, k), Ax, Bx), fmaf(qbyte(nwy, k), Ay, By), fmaf(qbyte(nwz, k), Az, Bz));
        const float tf = min3f(fmaf(qbyte(fwx, k), Ax, Bx), fmaf(qbyte(fwy, k), Ay, By), fmaf(qbyte(fwz, k), Az, Bz));
        const bool h = (uint32_t)k < nch && fmaxf(tn, 0.f) <= fminf(tf, L.best_t);
        key[k] = h ? tn : INFINITY;
        nhit += h ? 1u : 0u;
    }
    if (FLAT && sp0 + 3u <= (uint32_t)LSTK) {
        // no branch on the number of hit children: the sorted children are pushed into the rows above the
        // stack top (those beyond the hit children into the spare row), and the next entry is the nearest
        // hit child or, with none, the popped entry; the pushes write rows >= sp0, never the one popped
        cswap(key[0], lnk[0], key[1], lnk[1]);
        cswap(key[2], lnk[2], key[3], lnk[3]);
        cswap(key[0], lnk[0], key[2], lnk[2]);
        cswap(key[1], lnk[1], key[3], lnk[3]);
        cswap(key[1], lnk[1], key[2], lnk[2]);
#pragma unroll
        for (int k = 3; k >= 1; k--) {
            const uint32_t at = (uint32_t)k < nhit ? sp0 + nhit - 1u - (uint32_t)k : (uint32_t)LSTK;
            stk[at * stride] = lnk[k];
        }
        L.ni = nhit != 0u ? lnk[0] : popv;
        L.sp = (int)(nhit != 0u ? sp0 + nhit - 1u : (sp0 != 0u ? sp0 - 1u : 0u));
        return (nhit == 0u && sp0 == 0u) || bad;
    }
#if RT_Q4_LATE_ERR
    return node4_order<STK, LSTK>(L, key, lnk, nhit, stk, ovf, stride, err) || bad;
#else
    return node4_order<STK, LSTK>(L, key, lnk, nhit, stk, ovf, stride, err);
#endif
}

// Test the leaf entry L.ni (closest accepted hit, ties to the lower primitive index), then pop.
template <int STK, int LSTK, typename C>
__device__ __forceinline__ bool lane_leaf(const float4 *recs, LaneRay &L, float tmin, uint32_t skip,
                                          const lds_u32 *stk, const uint32_t *ovf, uint32_t stride, C &cnt) {
    cnt.prims++;
    uint32_t slot_unused;
    test_leaf<false, C::flags>(recs, L.ni & ~RT_LEAF_BIT, L.o, L.d, tmin, skip, L.best_t, L.best, slot_unused);
    return lane_pop<STK, LSTK>(L, stk, ovf, stride);
}

// The result does not depend on the visiting order: a primitive hit at t <= the final closest t lies in
// boxes whose entry distance is <= t <= every intermediate best_t, so no step culls it, and the tie rule
// is order-free.  The entries visited equal trace_closest()'s up to that order.

// ------------------------------------------------------------------------------------------------
// diffuse / scattering rays -> hit words
// ------------------------------------------------------------------------------------------------
// A wave keeps 64 rays in flight.  Rays of one primitive are handed to idle lanes in ray order; a lane
// whose traversal ends writes its hit word and takes the primitive's next ray, and the wave opens the
// next primitive of the work queue when the current one has no rays left.  The traversal loop runs until
// fewer than RT_REFILL_BELOW lanes are still traversing (all of them once the queue is empty), so lanes do
// not idle for the longest ray of a fixed group of 64.  Every ray is traced exactly as a plain
// per-ray trace_loop() would trace it; only the assignment of rays to lanes and steps differs.
// The traversal stack keeps its first RT_DIFFUSE_LDS_STACK entries in LDS and deeper ones in private
// memory: 4 KB of LDS per wave leaves the register file (about 70 VGPRs) as the occupancy limit.
#ifndef RT_REFILL_BELOW
#define RT_REFILL_BELOW 48
#endif
#ifndef RT_LEAF_BATCH
#define RT_LEAF_BATCH 16
#endif
// Diffuse lanes park the first leaf they reach and keep stepping (585 against 598 ms for the orchard's five
// passes)
#ifndef RT_LEAF_DEFER
#define RT_LEAF_DEFER 1
#endif
#define RT_DEFER_END 0xFFFFFFFFu
// With parked leaves, the leaf phase tests one leaf per waiting lane and returns to the node phase (lanes
// that reach further leaves park them there), and starts once 8 lanes wait: 589 -> 554 ms for the
// orchard's five passes (batches of 2-24 lanes swept; the full leaf phase with 16 lanes was the 589)
#ifndef RT_LEAF_ONE_ROUND
#define RT_LEAF_ONE_ROUND 1
#endif
#ifndef RT_LEAF_BATCH_PARKED
#define RT_LEAF_BATCH_PARKED 8
#endif
#ifndef RT_DIFFUSE_LDS_STACK
#define RT_DIFFUSE_LDS_STACK 16
#endif
// Waves per SIMD the compiler must fit the timed diffuse kernel into (the default launch shape, 16 two-wave
// workgroups per CU, runs 8 per SIMD): on the orchard 8 waves with a few spills beat 5 waves without (622
// against 726 ms per diurnal series).  With texture masks, 8 as well (the homogeneous-canopy benchmark's
// diffuse runBand 0.67 s at 5 waves, 0.56 s at 8); with watertight predicates, 6 (one plain orchard solve
// 804 ms at 8, 732 ms at 6: the larger leaf test spills more); the counting variants keep their registers.
#ifndef RT_DIFFUSE_WAVES_PER_EU
#define RT_DIFFUSE_WAVES_PER_EU 8
#endif
#ifndef RT_DIFFUSE_WAVES_PER_EU_WT
#define RT_DIFFUSE_WAVES_PER_EU_WT 6
#endif
#ifndef RT_DIFFUSE_WAVES_PER_EU_MASKED
#define RT_DIFFUSE_WAVES_PER_EU_MASKED 8
#endif
template <int WPB, int STK, bool COUNT, int MK>
__global__ __launch_bounds__(WPB * 64, COUNT ? 1 : ((MK & RT_FL_MASK) ? RT_DIFFUSE_WAVES_PER_EU_MASKED
                                                                    : (MK & RT_FL_WT) ? RT_DIFFUSE_WAVES_PER_EU_WT
                                                                                      : RT_DIFFUSE_WAVES_PER_EU))
void k_diffuse(const DiffuseJob *__restrict__ job, uint32_t total,
                                                      uint32_t *counter, uint32_t chunk, int *err,
                                                      unsigned long long *stats) {
    constexpr int LSTK = STK < RT_DIFFUSE_LDS_STACK ? STK : RT_DIFFUSE_LDS_STACK;
    __shared__ uint32_t stk_all[(LSTK + 1) * WPB * 64];  // + the spare row of the branch-free BVH4 pushes
    // overflow entries: sized for the deepest stack (RT_STACK) so the array stays in private memory
    // instead of being promoted to registers, which would cost the occupancy the small LDS part buys
    uint32_t ovf[STK > LSTK ? RT_STACK - LSTK : 1];
    lds_u32 *stk = (lds_u32 *)stk_all + threadIdx.x;
    const uint32_t stride = WPB * 64;
    const uint32_t lane = threadIdx.x & 63;
    Cnt<COUNT, MK> cnt{};
    uint32_t nray = 0;
    // job and scene (wave-uniform, scalar loads)
    const DevScene &sc = *ldu(&job->sc);
    const float4 *nodes = ldu(&sc.node), *recs = ldu(&sc.rec), *nrm = ldu(&sc.nrm), *frame = ldu(&sc.frame);
    // the 4-wide tree when it exists and its worst-case stack fits (3 pushes per level); the counting
    // variant walks the binary tree, whose node visits are the canonical work measure
    const float4 *nodes4 = ldu(&sc.node4);
    const uint32_t nn4 = ldu(&sc.nnode4);
    const bool use4 = RT_DIFFUSE_BVH4 && !COUNT && nodes4 != nullptr && 3u * ldu(&sc.depth4) + 1u <= (uint32_t)STK;
    // the 8-bit 4-wide tree (64 B per node): half the node fetches of the binary tree per ray
    const float4 *q4 = ldu(&sc.qnode4);
    const bool useq = RT_DIFFUSE_QBVH4 && !COUNT && q4 != nullptr && 3u * ldu(&sc.depth4) + 1u <= (uint32_t)STK;
    const float *T = ldu(&sc.T);
    const uint32_t *type = ldu(&sc.type);
    const uint32_t *mdesc = ldu(&sc.mdesc);
    const uint32_t nleaf = ldu(&sc.nleaf), nbbox = ldu(&sc.nbbox);
    const uint32_t n = ldu(&job->n), nn = n * n, seed = ldu(&job->seed);
    const uint32_t faces = ldu(&job->faces) ? ldu(&job->faces) : 3u;
    const uint32_t item_begin = ldu(&job->item_begin);
    const uint32_t *order = ldu(&job->order);
    uint32_t *hits = ldu(&job->hits);
    // work queue and open primitive (wave-uniform); the primitive's frame and transform are reloaded
    // (scalar) whenever rays are generated, so they hold no registers across the traversal
    uint32_t it = 0, it_end = 0;
    bool drained = false;
    uint32_t op_i = 0, op_next = 0, op_cnt = 0, op_base = 0, op_m = RT_MISS;
    bool patch = false, flip = false;
    // lane state: busy = holds a ray, trav = its traversal has not ended
    bool busy = false, trav = false;
    LaneRay L;
    L.o = L.d = mk3(0.f, 0.f, 0.f);
    lane_start(L, nleaf);
    uint32_t my_i = 0, my_r = 0, wraps = 0;
    uint32_t pend = 0u;  // a parked leaf entry (RT_LEAF_DEFER), 0 = none
    while (true) {
        // finish lanes whose traversal ended: periodic wrap (as trace_loop: at most 32 traces), else write
        if (busy && !trav) {
            bool done = true;
            if (RT_Q4_LATE_ERR && L.budget == RT_BUDGET_BAD) atomicOr(err, 2);
            if (nbbox) {
                float bt = nearest_face(sc, L.o, L.d, L.best_t);
                if (bt > 0.f && wraps < 31) {
                    L.o = wrap_at(sc, L.o, L.d, bt);
                    wraps++;
                    trav = lane_begin(L, recs, nleaf, 1e-5f, my_i, cnt);
                    done = false;
                }
            }
            if (done) {
                uint32_t word = RT_MISS;
                if (L.best != RT_MISS) {
                    float4 hn = nrm[L.best];
                    bool hit_top = dot3(mk3(hn.x, hn.y, hn.z), L.d) < 0.0f;
                    word = L.best | (hit_top ? 0u : RT_HIT_BOTTOM);
                }
                hits[(size_t)my_i * 2 * nn + my_r] = word;
                busy = false;
            }
        }
        // refill idle lanes
        while (true) {
            const uint64_t idle = __ballot(!busy);
            if (idle == 0) break;
            if (op_next == op_cnt) {
                // open the next primitive of the queue
                if (it == it_end) {
                    if (drained) break;
                    it = wave_item(counter, chunk);
                    if (it >= total) { drained = true; it = it_end = 0; break; }
                    it_end = min(it + chunk, total);
                }
                uint32_t i = it - item_begin;
                if (order) i = ldu(order + i);
                i = __builtin_amdgcn_readfirstlane(i);
                it++;
                uint32_t *w = hits + (size_t)i * 2 * nn;
                const uint32_t ty = ldu(type + i);
                if (!traced_type(ty)) {
                    for (uint32_t r = lane; r < 2 * nn; r += 64) w[r] = RT_NOT_LAUNCHED;
                    continue;
                }
                const uint32_t tw = __float_as_uint(ldu(&nrm[i].w));
                const bool s0 = (faces & 1u) != 0, s1 = ((faces >> 1) & 1u) != 0 && tw != 0;
                if (!s0) for (uint32_t r = lane; r < nn; r += 64) w[r] = RT_NOT_LAUNCHED;
                if (!s1) for (uint32_t r = nn + lane; r < 2 * nn; r += 64) w[r] = RT_NOT_LAUNCHED;
                op_i = i;
                op_m = (MK && mdesc) ? ldu(mdesc + i) : RT_MISS;
                op_next = 0;
                op_cnt = ((s0 ? 1u : 0u) + (s1 ? 1u : 0u)) * nn;
                op_base = s0 ? 0u : nn;
                patch = ty == 0 || ty == 3;
                flip = tw == 1;
                continue;
            }
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            const uint32_t take = min((uint32_t)__popcll(idle), op_cnt - op_next);
            if (!busy && rank < take) {
                const uint32_t r = op_base + op_next + rank;
                const uint32_t slot = r >= nn ? 1u : 0u;  // slot 0 = top launch (face 1), slot 1 = bottom (face 0)
                const uint32_t rr = r - slot * nn;
                float spx, spy;
                v3f dl;
                if (diffuse_sample_m(op_i, rr % n, rr / n, n, seed, patch, recs, op_m, spx, spy, dl)) {
                    float tm[12];
                    const float *Tp = T + (size_t)op_i * 16;
#pragma unroll
                    for (int q = 0; q < 12; q++) tm[q] = ldu(Tp + q);
                    const float4 fr = ldu(frame + op_i);
                    L.o = xform3(tm, spx, spy, 0.0f);
                    v3f dir = frame_rotate(dl, fr.x, fr.y, fr.z, fr.w);
                    if (slot == 1 && flip) dir = mk3(-dir.x, -dir.y, -dir.z);
                    L.d = dir;
                    my_i = op_i;
                    my_r = r;
                    wraps = 0;
                    nray++;
                    busy = true;
                    trav = lane_begin(L, recs, nleaf, 1e-5f, my_i, cnt);
                } else {
                    // origin on transparent texels: the ray is dropped (adds nothing to the gather); the lane
                    // stays idle and takes another ray
                    hits[(size_t)op_i * 2 * nn + r] = RT_NOT_LAUNCHED;
                }
            }
            op_next += take;
        }
        if (__ballot(busy) == 0) break;  // queue empty and nothing in flight
        // traverse until fewer than RT_REFILL_BELOW lanes (once the queue is empty: no lane) still traverse
        // node phase: inner-node steps until no lane needs one or RT_LEAF_BATCH lanes wait at a leaf; leaf
        // phase: the waiting lanes test their leaves together (the intersection code runs with many lanes
        // active instead of once per node step for the few lanes that reached a leaf)
        const uint32_t floor_lanes = drained ? 1u : RT_REFILL_BELOW;
        if (RT_LEAF_DEFER && !COUNT) {
        // A lane that reaches a leaf parks it (pend) and goes on with its next entry, so it keeps taking node
        // steps; it waits only when it reaches a second leaf or runs out of entries (L.ni = RT_DEFER_END).  The
        // leaf phase tests the parked leaves and the leaves lanes wait at, many lanes at a time, one leaf per
        // lane per phase.  (The counting variant keeps the plain order, whose node visits are the canonical
        // work measure.)
        while (true) {
            while (true) {
                if (trav && pend == 0u && (L.ni & RT_LEAF_BIT) && L.ni != RT_DEFER_END) {
                    pend = L.ni;
                    if (lane_pop<STK, LSTK>(L, stk, ovf, stride)) L.ni = RT_DEFER_END;
                }
                const bool inner = trav && !(L.ni & RT_LEAF_BIT);
                const uint64_t mi = __ballot(inner);
                if (mi == 0 || (uint32_t)__popcll(__ballot(trav && !inner)) >= RT_LEAF_BATCH_PARKED) break;
                if (inner && (useq   ? lane_node4q<STK, LSTK, RT_Q4_FLAT_DIFFUSE != 0>(q4, nn4, L, stk, ovf, stride, err, cnt)
                              : use4 ? lane_node4<STK, LSTK>(nodes4, nn4, L, stk, ovf, stride, err, cnt)
                                     : lane_node<STK, LSTK>(nodes, nleaf, L, stk, ovf, stride, err, cnt))) {
                    if (pend != 0u) L.ni = RT_DEFER_END;
                    else trav = false;
                }
            }
            for (int round = 0; !RT_LEAF_ONE_ROUND || round < 1; round++) {
                const bool atleaf = trav && (L.ni & RT_LEAF_BIT) && L.ni != RT_DEFER_END;
                const bool work = trav && (pend != 0u || atleaf);
                if (__ballot(work) == 0) break;
                if (work) {
                    uint32_t slot_unused;
                    cnt.prims++;
                    if (pend != 0u) {
                        test_leaf<false, MK>(recs, pend & ~RT_LEAF_BIT, L.o, L.d, 1e-5f, my_i,
                                                              L.best_t, L.best, slot_unused);
                        pend = 0u;
                    } else {
                        test_leaf<false, MK>(recs, L.ni & ~RT_LEAF_BIT, L.o, L.d, 1e-5f, my_i,
                                                              L.best_t, L.best, slot_unused);
                        if (lane_pop<STK, LSTK>(L, stk, ovf, stride)) L.ni = RT_DEFER_END;
                    }
                }
                if (trav && pend == 0u && L.ni == RT_DEFER_END) trav = false;
            }
            if ((uint32_t)__popcll(__ballot(trav)) < floor_lanes) break;
        }
        } else {
        while (true) {
            while (true) {
                const bool inner = trav && !(L.ni & RT_LEAF_BIT);
                const uint64_t mi = __ballot(inner);
                if (mi == 0 || (uint32_t)__popcll(__ballot(trav && !inner)) >= RT_LEAF_BATCH) break;
                if (inner && (useq   ? lane_node4q<STK, LSTK, RT_Q4_FLAT_DIFFUSE != 0>(q4, nn4, L, stk, ovf, stride, err, cnt)
                              : use4 ? lane_node4<STK, LSTK>(nodes4, nn4, L, stk, ovf, stride, err, cnt)
                                     : lane_node<STK, LSTK>(nodes, nleaf, L, stk, ovf, stride, err, cnt)))
                    trav = false;
            }
            while (true) {
                const bool leaf = trav && (L.ni & RT_LEAF_BIT);
                if (__ballot(leaf) == 0) break;
                if (leaf && lane_leaf<STK, LSTK>(recs, L, 1e-5f, my_i, stk, ovf, stride, cnt)) trav = false;
            }
            if ((uint32_t)__popcll(__ballot(trav)) < floor_lanes) break;
        }
        }
    }
    add_stats(stats, cnt, nray);
}

// ------------------------------------------------------------------------------------------------
// gather: lanes over rays (few channels)
// ------------------------------------------------------------------------------------------------
// Direction of diffuse ray r (face slot r / nn) of primitive i, regenerated from the pass seed exactly as
// the trace kernel generated it: the gather weighs a sky hit by the sky distribution in that direction.
__device__ __forceinline__ v3f diffuse_dir(const DevScene &sc, uint32_t i, uint32_t r, uint32_t n, uint32_t seed) {
    const uint32_t nn = n * n, slot = r >= nn ? 1u : 0u, rr = r - slot * nn;
    const uint32_t ty = sc.type[i];
    const uint32_t mdi = sc.mdesc ? sc.mdesc[i] : RT_MISS;
    float spx, spy;
    v3f dl;
    (void)diffuse_sample_m(i, rr % n, rr / n, n, seed, ty == 0 || ty == 3, sc.rec, mdi, spx, spy, dl);
    const float4 fr = sc.frame[i];
    v3f dir = frame_rotate(dl, fr.x, fr.y, fr.z, fr.w);
    if (slot == 1 && __float_as_uint(sc.nrm[i].w) == 1u) dir = mk3(-dir.x, -dir.y, -dir.z);
    return dir;
}

// ACCT: also record the pass's incident and sky flux densities and the sky-bound ray counts (GatherJob.acct,
// .miss) for the energy bookkeeping; a separate instance, so the plain gather carries no extra registers.
template <int CM, bool ACCT>
__global__ __launch_bounds__(256) void k_gather_rays(const GatherJob *__restrict__ jobs, uint32_t njobs, uint32_t total,
                                                     uint32_t *counter, uint32_t chunk) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t first = wave_item(counter, chunk); first < total; first = wave_item(counter, chunk)) {
        const uint32_t end = min(first + chunk, total);
        for (uint32_t item = first; item < end; ++item) {
            uint32_t j = find_job(jobs, njobs, item);
            const GatherJob &J = jobs[j];
            uint32_t i = item - J.item_begin;
            if (J.order) i = J.order[i];
            uint32_t ty = J.sc->type[i];
            if (!traced_type(ty)) continue;
            uint32_t nn = J.n * J.n, C = J.C;
            const uint32_t *w = J.hits + (size_t)i * 2 * nn;
            ksum acc[2][CM];
            ksum asky[ACCT ? CM : 1];
            uint32_t mt = 0, mb = 0;
#pragma unroll
            for (int f = 0; f < 2; f++)
#pragma unroll
                for (int c = 0; c < CM; c++) acc[f][c].init();
#pragma unroll
            for (int c = 0; c < (ACCT ? CM : 1); c++) asky[c].init();
            for (uint32_t r = lane; r < 2 * nn; r += 64) {
                uint32_t word = w[r];
                if (word == RT_NOT_LAUNCHED) continue;
                uint32_t slot = r >= nn ? 1u : 0u;
                if (ACCT && word == RT_MISS) { if (slot) mb++; else mt++; }
                if (word == RT_MISS && J.sky) {
                    // anisotropic sky: fd(direction) x diffuse flux (diffuse_raygen.comp:613-625)
                    const v3f dir = diffuse_dir(*J.sc, i, r, J.n, J.seed);
#pragma unroll
                    for (int c = 0; c < CM; c++)
                        if ((uint32_t)c < C) {
                            float v = sky_weight(dir, J.sky + 9 * c) * J.dflux[c];
                            if (slot) acc[1][c].add(v); else acc[0][c].add(v);
                            if (ACCT) asky[ACCT ? c : 0].add(v);
                        }
                } else if (word == RT_MISS) {
#pragma unroll
                    for (int c = 0; c < CM; c++)
                        if ((uint32_t)c < C) {
                            float v = J.dflux[c];
                            if (slot) acc[1][c].add(v); else acc[0][c].add(v);
                            if (ACCT) asky[ACCT ? c : 0].add(v);
                        }
                } else {
                    uint32_t h = word & ~RT_HIT_BOTTOM;
                    const float *src = ((word & RT_HIT_BOTTOM) ? J.out_bot : J.out_top) + (size_t)h * C;
#pragma unroll
                    for (int c = 0; c < CM; c++)
                        if ((uint32_t)c < C) {
                            float v = src[c];
                            if (slot) acc[1][c].add(v); else acc[0][c].add(v);
                        }
                }
            }
            float strength = rdiv(1.0f, (float)nn);
#pragma unroll
            for (int c = 0; c < CM; c++) {
                wave_reduce(acc[0][c]);
                wave_reduce(acc[1][c]);
                if (ACCT) wave_reduce(asky[ACCT ? c : 0]);
            }
            if (ACCT) {
                for (int off = 32; off > 0; off >>= 1) { mt += __shfl_xor(mt, off); mb += __shfl_xor(mb, off); }
                if (lane == 0) { J.miss[(size_t)i * 2] = mt; J.miss[(size_t)i * 2 + 1] = mb; }
            }
            if (lane < C && lane < CM) {
                float I1 = 0.f, I0 = 0.f, Is = 0.f;
#pragma unroll
                for (int c = 0; c < CM; c++)
                    if ((uint32_t)c == lane) {
                        I1 = acc[0][c].value(); I0 = acc[1][c].value();
                        if (ACCT) Is = asky[ACCT ? c : 0].value();
                    }
                uint32_t c = lane;
                const bool act = J.active[c] != 0;
                I1 = I1 * strength;
                I0 = I0 * strength;
                if (act) {
                    uint32_t b = c % J.B;
                    float rho = J.rho[(size_t)i * J.B + b], tau = J.tau[(size_t)i * J.B + b];
                    size_t q = (size_t)i * C + c;
                    J.in[q] += (I1 + I0) * (1.0f - rho - tau);
                    J.st[q] += I1 * rho + I0 * tau;
                    J.sb[q] += I1 * tau + I0 * rho;
                }
                if (ACCT) {
                    const size_t q = ((size_t)i * C + c) * 2;
                    J.acct[q] = act ? I1 + I0 : 0.f;
                    J.acct[q + 1] = act ? Is * strength : 0.f;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// gather: lanes over channels (many channels, e.g. suns x bands of a diurnal batch)
// ------------------------------------------------------------------------------------------------
#ifndef RT_GATHER_UNROLL
#define RT_GATHER_UNROLL 4
#endif
// Waves per SIMD of the channel-parallel gather: 8 (64 VGPRs, a few spills) took the diurnal series'
// gathers from 180 to 166 ms against the 5 the compiler's register use allowed.
#ifndef RT_GATHER_WAVES_PER_EU
#define RT_GATHER_WAVES_PER_EU 8
#endif
template <int CB, bool ACCT>
__global__ __launch_bounds__(256, RT_GATHER_WAVES_PER_EU) void k_gather_chan(const GatherJob *__restrict__ jobs, uint32_t njobs, uint32_t total,
                                                     uint32_t *counter, uint32_t chunk) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t first = wave_item(counter, chunk); first < total; first = wave_item(counter, chunk)) {
        const uint32_t end = min(first + chunk, total);
        for (uint32_t item = first; item < end; ++item) {
            uint32_t j = find_job(jobs, njobs, item);
            const GatherJob &J = jobs[j];
            uint32_t i = item - J.item_begin;
            if (J.order) i = J.order[i];
            uint32_t ty = J.sc->type[i];
            if (!traced_type(ty)) continue;
            uint32_t nn = J.n * J.n, C = J.C;
            const uint32_t *w = J.hits + (size_t)i * 2 * nn;
            ksum acc[2][CB];
            ksum asky[ACCT ? CB : 1];
            uint32_t mt = 0, mb = 0;
#pragma unroll
            for (int k = 0; k < CB; k++) {
                acc[0][k].init();
                acc[1][k].init();
            }
#pragma unroll
            for (int k = 0; k < (ACCT ? CB : 1); k++) asky[k].init();
            const float *dfl = J.dflux, *otop = J.out_top, *obot = J.out_bot;
            const float *sky = J.sky;
            for (uint32_t base = 0; base < 2 * nn; base += 64) {
                uint32_t mine = base + lane < 2 * nn ? w[base + lane] : RT_NOT_LAUNCHED;
                uint32_t cnt = min(64u, 2 * nn - base);
                if (ACCT) {
                    mt += (uint32_t)__popcll(__ballot(mine == RT_MISS && base + lane < nn));
                    mb += (uint32_t)__popcll(__ballot(mine == RT_MISS && base + lane >= nn));
                }
                // anisotropic sky: each lane regenerates the direction of its own ray if it missed; the
                // directions are broadcast with the words
                v3f mdir = mk3(0.f, 0.f, 1.f);
                if (sky && mine == RT_MISS) mdir = diffuse_dir(*J.sc, i, base + lane, J.n, J.seed);
                // RT_GATHER_UNROLL hit words per step: every row (the hit primitive's radiation_out C-vector,
                // or the sky fluxes for a miss) is loaded before any is summed, so that many rows are in
                // flight per wave; the sums take them in ray order.  A face slot that was not launched adds
                // nothing (its value is 0, which leaves a compensated sum unchanged).
                for (uint32_t q0 = 0; q0 < cnt; q0 += RT_GATHER_UNROLL) {
                    float v[RT_GATHER_UNROLL][CB];
#pragma unroll
                    for (int u = 0; u < RT_GATHER_UNROLL; u++) {
                        const uint32_t q = q0 + u;
                        const uint32_t word = q < cnt ? __builtin_amdgcn_readlane(mine, q) : RT_NOT_LAUNCHED;
                        const bool skip = word == RT_NOT_LAUNCHED;
                        const float *src = word == RT_MISS ? dfl
                                           : ((word & RT_HIT_BOTTOM) ? obot : otop) + (size_t)(word & ~RT_HIT_BOTTOM) * C;
#pragma unroll
                        for (int k = 0; k < CB; k++) {
                            uint32_t c = lane + 64 * k;
                            v[u][k] = (!skip && c < C) ? src[c] : 0.f;
                        }
                        if (sky && word == RT_MISS) {
                            const v3f dq = mk3(__shfl(mdir.x, q), __shfl(mdir.y, q), __shfl(mdir.z, q));
#pragma unroll
                            for (int k = 0; k < CB; k++) {
                                uint32_t c = lane + 64 * k;
                                if (c < C) v[u][k] = sky_weight(dq, sky + 9 * c) * v[u][k];
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < RT_GATHER_UNROLL; u++) {
                        const uint32_t slot = base + q0 + u >= nn ? 1u : 0u;
#pragma unroll
                        for (int k = 0; k < CB; k++) {
                            if (slot) acc[1][k].add(v[u][k]); else acc[0][k].add(v[u][k]);
                        }
                        if (ACCT) {
                            const uint32_t q = q0 + u;
                            const bool is_miss = q < cnt && __builtin_amdgcn_readlane(mine, q) == RT_MISS;
#pragma unroll
                            for (int k = 0; k < (ACCT ? CB : 1); k++)
                                if (is_miss) asky[k].add(v[u][k]);
                        }
                    }
                }
            }
            float strength = rdiv(1.0f, (float)nn);
            if (ACCT && lane == 0) { J.miss[(size_t)i * 2] = mt; J.miss[(size_t)i * 2 + 1] = mb; }
#pragma unroll
            for (int k = 0; k < CB; k++) {
                uint32_t c = lane + 64 * k;
                if (ACCT && c < C) {
                    const size_t q = ((size_t)i * C + c) * 2;
                    const bool act = J.active[c] != 0;
                    J.acct[q] = act ? (acc[0][k].value() * strength + acc[1][k].value() * strength) : 0.f;
                    J.acct[q + 1] = act ? asky[ACCT ? k : 0].value() * strength : 0.f;
                }
                if (c >= C || !J.active[c]) continue;
                float I1 = acc[0][k].value() * strength, I0 = acc[1][k].value() * strength;
                uint32_t b = c % J.B;
                float rho = J.rho[(size_t)i * J.B + b], tau = J.tau[(size_t)i * J.B + b];
                size_t qq = (size_t)i * C + c;
                J.in[qq] += (I1 + I0) * (1.0f - rho - tau);
                J.st[qq] += I1 * rho + I0 * tau;
                J.sb[qq] += I1 * tau + I0 * rho;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// castRaysSoA
// ------------------------------------------------------------------------------------------------
// One closest-hit leaf test with the CollisionDetection predicates (tmin 1e-6, nothing skipped), keeping the
// hit's leaf slot for the normal, then pop.
template <int STK, int LSTK, typename C>
__device__ __forceinline__ bool lane_leaf_cd(const float4 *recs, LaneRay &L, uint32_t &slot, const lds_u32 *stk,
                                             const uint32_t *ovf, uint32_t stride, C &cnt) {
    cnt.prims++;
    test_leaf<true, C::flags>(recs, L.ni & ~RT_LEAF_BIT, L.o, L.d, 1e-6f, RT_MISS, L.best_t, L.best, slot);
    return lane_pop<STK, LSTK>(L, stk, ovf, stride);
}

// Next 64-ray chunk of a wave (wave-uniform 64-bit work counter).
__device__ __forceinline__ unsigned long long wave_chunk64(unsigned long long *counter) {
    const uint32_t lane = threadIdx.x & 63;
    unsigned long long f = atomicAdd(counter, lane == 0 ? 64ull : 0ull);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)f);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(f >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// castRaysSoA (CD/src/CollisionDetection_RayTracing.cpp:551-680, 980-1080): closest hit of each ray with
// the CD predicates.  Persistent waves take 64-ray chunks from a counter and keep 64 rays in flight: a lane
// whose trace ends writes its result and takes the chunk's next ray; traversal runs in node and leaf phases
// as k_diffuse's.  Each ray is traced exactly as on its own.
template <int STK, bool COUNT, int MK>
#ifndef RT_CAST_WAVES_PER_EU
#define RT_CAST_WAVES_PER_EU 1
#endif
__global__ __launch_bounds__(256, COUNT ? 1 : RT_CAST_WAVES_PER_EU) void k_cast(const DevScene *__restrict__ scp, uint64_t n, const float *__restrict__ O,
                                              const float *__restrict__ D, float maxd, float *__restrict__ T,
                                              float *__restrict__ N, uint32_t *__restrict__ U,
                                              unsigned long long *counter, int *err, unsigned long long *stats) {
    constexpr int LSTK = STK < RT_DIFFUSE_LDS_STACK ? STK : RT_DIFFUSE_LDS_STACK;
    __shared__ uint32_t stk_all[(LSTK + 1) * 256];  // + the spare row of the branch-free BVH4 pushes
    uint32_t ovf[STK > LSTK ? RT_STACK - LSTK : 1];
    lds_u32 *stk = (lds_u32 *)stk_all + threadIdx.x;
    const uint32_t stride = 256;
    const DevScene &sc = *scp;
    const float4 *nodes = ldu(&sc.node), *recs = ldu(&sc.rec);
    const uint32_t *uuid = ldu(&sc.uuid);
    const uint32_t nleaf = ldu(&sc.nleaf);
    const float tstart = maxd > 0.f ? maxd : FLT_MAX;
    // the 8-bit 4-wide tree when present and its worst-case stack fits; the counting variant walks the
    // binary tree, whose node visits are the canonical work measure
    const float4 *q4 = ldu(&sc.qnode4);
    const uint32_t nn4 = ldu(&sc.nnode4);
    const bool useq = RT_CAST_QBVH4 && !COUNT && q4 != nullptr && 3u * ldu(&sc.depth4) + 1u <= (uint32_t)STK;
    Cnt<COUNT, MK> cnt{};
    uint32_t nray = 0;
    unsigned long long q_next = 0, q_end = 0, my_r = 0;
    bool drained = false, busy = false, trav = false;
    LaneRay L;
    L.o = L.d = mk3(0.f, 0.f, 0.f);
    lane_start(L, nleaf);
    uint32_t slot = RT_MISS;
    while (true) {
        // finished lanes write their result
        if (busy && !trav) {
            if (RT_Q4_LATE_ERR && L.budget == RT_BUDGET_BAD) atomicOr(err, 2);
            const uint32_t h = L.best;
            if (h == RT_MISS) {
                U[my_r] = RT_MISS;
            } else {
                U[my_r] = uuid[h];
                T[my_r] = L.best_t;
                const float4 *rr = recs + (size_t)slot * 4;
                const float4 r0 = rr[0], r1 = rr[1], r2 = rr[2], r3 = rr[3];
                const uint32_t tag = __float_as_uint(r0.w);
                const bool patch = (tag & RT_PATCH_BIT) != 0;
                // CD/src/CollisionDetection_RayTracing.cpp:1572-1585: normal of cross(v1-v0, v2-v0), facing
                // the origin; a voxel's is the outward normal of the face hit (1685-1711), not turned
                v3f nn;
                bool ok;
                if (tag & RT_VOXEL_BIT) {
                    nn = voxel_normal_cd(r0, r1, add3(L.o, mul3(L.d, L.best_t)));
                    ok = false;
                } else if (patch) {
                    nn = mk3(r3.x, r3.y, r3.z);
                    v3f raw = cross3(mk3(r1.x, r1.y, r1.z), add3(mk3(r1.x, r1.y, r1.z), mk3(r2.x, r2.y, r2.z)));
                    ok = rsqrt_cr(raw.x * raw.x + raw.y * raw.y + raw.z * raw.z) > 1e-8f;
                } else {
                    nn = cross3(mk3(r1.x, r1.y, r1.z), mk3(r2.x, r2.y, r2.z));
                    float mag = rsqrt_cr(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
                    ok = mag > 1e-8f;
                    if (ok) nn = mul3(nn, rdiv(1.0f, mag));
                }
                if (ok) {
                    v3f ip = add3(L.o, mul3(L.d, L.best_t));
                    v3f to = sub3(L.o, ip);
                    if (nn.x * to.x + nn.y * to.y + nn.z * to.z < 0) nn = mul3(nn, -1.0f);
                } else if (!(tag & RT_VOXEL_BIT)) {
                    nn = nrm3(mk3(-L.d.x, -L.d.y, -L.d.z));
                }
                N[my_r * 3] = nn.x;
                N[my_r * 3 + 1] = nn.y;
                N[my_r * 3 + 2] = nn.z;
            }
            busy = false;
        }
        // idle lanes take the next rays of the wave's chunk
        while (true) {
            const uint64_t idle = __ballot(!busy);
            if (idle == 0) break;
            if (q_next == q_end) {
                if (drained) break;
                const unsigned long long f = wave_chunk64(counter);
                if (f >= n) { drained = true; break; }
                q_next = f;
                q_end = f + 64 < n ? f + 64 : n;
            }
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            const uint32_t left = (uint32_t)(q_end - q_next);
            const uint32_t take = min((uint32_t)__popcll(idle), left);
            if (!busy && rank < take) {
                const unsigned long long r = q_next + rank;
                L.o = mk3(O[r * 3], O[r * 3 + 1], O[r * 3 + 2]);
                L.d = mk3(D[r * 3], D[r * 3 + 1], D[r * 3 + 2]);
                lane_start(L, nleaf);
                L.best_t = tstart;
                slot = RT_MISS;
                trav = nleaf > 1;
                if (nleaf == 1) { cnt.prims++; test_leaf<true, MK>(recs, 0, L.o, L.d, 1e-6f, RT_MISS, L.best_t, L.best, slot); }
                my_r = r;
                busy = true;
                nray++;
            }
            q_next += take;
        }
        if (__ballot(busy) == 0) break;  // queue empty and nothing in flight
        const uint32_t floor_lanes = drained ? 1u : RT_REFILL_BELOW;
        while (true) {
            while (true) {
                const bool inner = trav && !(L.ni & RT_LEAF_BIT);
                const uint64_t mi = __ballot(inner);
                if (mi == 0 || (uint32_t)__popcll(__ballot(trav && !inner)) >= RT_LEAF_BATCH) break;
                if (inner && (useq ? lane_node4q<STK, LSTK, false>(q4, nn4, L, stk, ovf, stride, err, cnt)
                                   : lane_node<STK, LSTK>(nodes, nleaf, L, stk, ovf, stride, err, cnt)))
                    trav = false;
            }
            while (true) {
                const bool leaf = trav && (L.ni & RT_LEAF_BIT);
                if (__ballot(leaf) == 0) break;
                if (leaf && lane_leaf_cd<STK, LSTK>(recs, L, slot, stk, ovf, stride, cnt)) trav = false;
            }
            if ((uint32_t)__popcll(__ballot(trav)) < floor_lanes) break;
        }
    }
    add_stats(stats, cnt, nray);
}

// ------------------------------------------------------------------------------------------------
// cameras (camera_raygen.comp, pixel_label_raygen.comp)
// ------------------------------------------------------------------------------------------------
// Closest hit of a camera ray with up to 10 periodic wraps (camera_raygen.comp:266-285); a ray still
// crossing faces after the 10th trace counts as a miss.  depth accumulates the distance of the wraps.
// trace_closest() with radiation predicates and a hybrid stack: entries below LSTK in LDS (stride
// 256), deeper ones in the private array ovf.
template <int LSTK, int STK, typename C>
__device__ __forceinline__ uint32_t trace_closest_hyb(const DevScene &sc, v3f o, v3f d, float tmin, float &t_out,
                                                      lds_u32 *stk, uint32_t *ovf, int *err, C &cnt) {
    float best_t = 1e30f;
    uint32_t best = RT_MISS, best_slot = RT_MISS;
    const float4 *nodes = sc.node, *recs = sc.rec;
    const uint32_t nleaf = sc.nleaf;
    if (nleaf <= 1) {
        if (nleaf == 1) { cnt.prims++; test_leaf<false, C::flags>(recs, 0, o, d, tmin, RT_MISS, best_t, best, best_slot); }
        t_out = best_t;
        return best;
    }
    RayPre rp = ray_pre(o, d);
    uint32_t ni = 0;
    int sp = 0;
    uint32_t budget = nleaf + 8;
    while (true) {
        if (budget-- == 0 || ni >= nleaf - 1) {
            atomicOr(err, 2);
            break;
        }
        cnt.nodes++;
        NodeTest nt = node_test_p(nodes, ni, rp, best_t);
        bool l0 = (nt.c0 & RT_LEAF_BIT) != 0, l1 = (nt.c1 & RT_LEAF_BIT) != 0;
        if (nt.h0 && l0) { cnt.prims++; test_leaf<false, C::flags>(recs, nt.c0 & ~RT_LEAF_BIT, o, d, tmin, RT_MISS, best_t, best, best_slot); }
        if (nt.h1 && l1) { cnt.prims++; test_leaf<false, C::flags>(recs, nt.c1 & ~RT_LEAF_BIT, o, d, tmin, RT_MISS, best_t, best, best_slot); }
        bool g0 = nt.h0 && !l0, g1 = nt.h1 && !l1;
        if (g0 && g1) {
            bool first0 = nt.tn0 <= nt.tn1;
            if (sp >= STK) {
                atomicOr(err, 1);
                break;
            }
            const uint32_t far = first0 ? nt.c1 : nt.c0;
            if (sp < LSTK) stk[(uint32_t)sp * 256u] = far;
            else ovf[sp - LSTK] = far;
            sp++;
            ni = first0 ? nt.c0 : nt.c1;
        } else if (g0) {
            ni = nt.c0;
        } else if (g1) {
            ni = nt.c1;
        } else {
            if (sp == 0) break;
            sp--;
            ni = sp < LSTK ? stk[(uint32_t)sp * 256u] : ovf[sp - LSTK];
        }
    }
    t_out = best_t;
    return best;
}

#ifndef RT_CAMERA_LDS_STACK
#define RT_CAMERA_LDS_STACK 16
#endif

template <int STK, typename CT>
__device__ __forceinline__ uint32_t camera_trace(const DevScene &sc, v3f &o, v3f d, float &t, float &depth,
                                                 lds_u32 *stk, uint32_t *ovf, int *err, CT &cnt) {
    constexpr int LSTK = STK < RT_CAMERA_LDS_STACK ? STK : RT_CAMERA_LDS_STACK;
    uint32_t hit = RT_MISS;
    depth = 0.0f;
    for (int wrap = 0; wrap < 10; ++wrap) {
        hit = trace_closest_hyb<LSTK, STK>(sc, o, d, 1e-5f, t, stk, ovf, err, cnt);
        if (sc.nbbox == 0) break;
        const float bt = nearest_face(sc, o, d, t);
        if (!(bt > 0.f)) break;
        depth += bt;
        o = wrap_at(sc, o, d, bt);
        hit = RT_MISS;
    }
    return hit;
}

// Pixel of this thread: each wave takes an 8 x 8 block of the tile (blocks row by row), so the rays of
// a wave leave through neighbouring pixels and traverse much the same nodes.
__device__ __forceinline__ bool camera_pixel(const CameraLaunch &c, uint32_t &px, uint32_t &py) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    const uint32_t blk = q >> 6, lane = q & 63u, bx = (c.res_x + 7u) >> 3;
    px = (blk % bx) * 8u + (lane & 7u);
    py = (blk / bx) * 8u + (lane >> 3);
    return px < c.res_x && py < c.res_y;
}

// Prague sky pattern (sky_distribution.glsl:60-90), or 1 (isotropic) when its normalisation is not positive.
__device__ __forceinline__ float cam_sky_weight(v3f d, const float *sun, const float *prm) {
    if (!(prm[3] > 0.0f)) return 1.0f;
    return sky_prague(d, sun[0], sun[1], sun[2], prm);
}

// Radiance of the sources a camera ray meets before tmax, for band b, summed in source order
// (camera_raygen.comp:349-377 on a hit, 431-455 on a miss, where collimated and sun-sphere sources add the
// solar disk's radiance when the ray lies within it).
__device__ __forceinline__ float cam_source_radiance(const CameraLaunch &c, v3f o, v3f d, uint32_t b, float tmax,
                                                     bool miss) {
    float rad = 0.0f;
    for (uint32_t s = 0; s < c.S; s++) {
        const float flux = c.flux[(size_t)s * c.B + b];
        if (flux <= 0.0f) continue;
        const uint32_t ty = c.src_type[s];
        const v3f sp = mk3(c.src_pos[3 * s], c.src_pos[3 * s + 1], c.src_pos[3 * s + 2]);
        const float *w = c.src_width + 2 * s;
        float cosang;
        if (ty == 0u || ty == 2u) {
            if (miss && d.x * c.sun[0] + d.y * c.sun[1] + d.z * c.sun[2] >= c.solar_cos && c.solar_rad[b] > 0.0f)
                rad += c.solar_rad[b];
        } else if (ty == 1u) {
            const float r = w[0] * 0.5f;
            if (cam_sphere_visible(o, d, sp, r, tmax)) rad += rdiv(rdiv(flux, 4.0f * RT_PI_F * r * r), RT_PI_F);
        } else if (ty == 3u) {
            if (cam_planar_visible(false, o, d, sp, w, c.src_rot + 9 * s, tmax, cosang))
                rad += rdiv(rdiv(flux, w[0] * w[1]) * cosang, RT_PI_F);
        } else if (ty == 4u) {
            const float r = w[0];
            if (cam_planar_visible(true, o, d, sp, w, c.src_rot + 9 * s, tmax, cosang))
                rad += rdiv(rdiv(flux, RT_PI_F * r * r) * cosang, RT_PI_F);
        }
    }
    return rad;
}

// Blinn-Phong specular term of a camera hit on primitive h for band b (camera_raygen.comp:362-409): for
// each source, the radiation it delivered to h in the direct launch (radiation_specular) times the
// camera's weight for that source and band and a Fresnel reflectance of 0.25; if positive, times the
// primitive's scale (mode 2, else 1), (n . h)^e (e + 2) / (2 pi samples), with h the half vector between
// the light direction (a collimated or sun-sphere source's direction, else towards the source's centre
// from the hit point) and the reversed view direction, and n the primitive's stored normal.  0 when the
// primitive's exponent is not positive.  FP32, the reference's operation order.
__device__ __forceinline__ float cam_specular(const CameraLaunch &c, v3f o, v3f d, float t, uint32_t h, uint32_t b,
                                              v3f n) {
    float acc = 0.0f;
    const float e = c.spec_exp[h];
    if (!(e > 0.0f)) return 0.0f;
    for (uint32_t s = 0; s < c.S; s++) {
        float spec = c.spec[((size_t)s * c.spec_P + h) * c.B + b];
        spec = spec * (c.fcam[(size_t)s * c.B + b] * 0.25f);
        if (!(spec > 0.0f)) continue;
        const uint32_t ty = c.src_type[s];
        v3f L = mk3(c.src_pos[3 * s], c.src_pos[3 * s + 1], c.src_pos[3 * s + 2]);
        if (!(ty == 0u || ty == 2u)) L = sub3(L, add3(o, mul3(d, t)));
        L = nrm3(L);
        const v3f hv = nrm3(sub3(L, d));
        const float scale = c.spec_mode == 2u ? c.spec_scale[h] : 1.0f;
        const float ch = fmaxf(0.0f, hv.x * n.x + hv.y * n.y + hv.z * n.z);
        acc += rdiv(spec * scale * powf(ch, e) * (e + 2.0f), (float)c.spp * 2.0f * RT_PI_F);
    }
    return acc;
}

// Camera pixels: one thread per pixel of the tile traces its antialiasing samples in order and adds
// their contributions to camera_radiation[b][pixel] (a fixed order, so a render is reproducible; the
// reference adds the samples with float atomics).
template <int STK, bool COUNT, int MK, bool SPEC>
__global__ __launch_bounds__(256) void k_camera(CameraLaunch c, int *err, unsigned long long *stats) {
    constexpr int LSTK = STK < RT_CAMERA_LDS_STACK ? STK : RT_CAMERA_LDS_STACK;
    __shared__ uint32_t stk_all[LSTK * 256];
    uint32_t ovf[STK > LSTK ? RT_STACK - LSTK : 1];
    lds_u32 *stk = (lds_u32 *)stk_all + threadIdx.x;
    uint32_t px, py;
    if (!camera_pixel(c, px, py)) return;
    const DevScene &sc = *c.sc;
    const uint32_t ii = c.off_x + px, jj = c.off_y + py;
    const size_t pix = (size_t)jj * c.full_x + ii, npix = (size_t)c.full_x * c.full_y;
    const float strength = rdiv(1.0f, (float)c.spp);
    const uint32_t dimx = c.res_x * c.res_y, indx = c.res_x * py + px;
    Cnt<COUNT, MK> cnt{};
    for (uint32_t smp = 0; smp < c.spp; smp++) {
        uint32_t st = tea16(indx + dimx * smp, c.seed);
        const float Rx = rnd24(st), Ry = rnd24(st);
        float lx = 0.f, ly = 0.f;
        if (c.lens > 0.0f) unit_disk(rnd24(st), rnd24(st), lx, ly);
        v3f o, d;
        camera_ray(c.pos[0], c.pos[1], c.pos[2], c.st, c.ct, c.sp, c.cp, c.focal, c.lens, c.aspect, c.viewplane, ii,
                   jj, c.full_x, c.full_y, Rx, Ry, lx, ly, o, d);
        float t, depth;
        const uint32_t h = camera_trace<STK>(sc, o, d, t, depth, stk, ovf, err, cnt);
        if (h != RT_MISS) {
            const float4 n4 = sc.nrm[h];
            const bool face = n4.x * d.x + n4.y * d.y + n4.z * d.z < 0.0f;  // true: top face hit
            const bool top = face || sc.type[h] == 4u;
            for (uint32_t b = 0; b < c.B; b++) {
                const float srad = cam_source_radiance(c, o, d, b, t, false);
                float surf = (top ? c.out_top : c.out_bot)[(size_t)h * c.B + b] * strength;
                if (srad > 0.0f) surf += srad * strength;
                if (SPEC) surf += cam_specular(c, o, d, t, h, b, mk3(n4.x, n4.y, n4.z));
                c.pixels[(size_t)b * npix + pix] += rdiv(surf, RT_PI_F);
            }
        } else {
            for (uint32_t b = 0; b < c.B; b++) {
                float rad = cam_source_radiance(c, o, d, b, 1e30f, true);
                if (rad <= 0.0f && c.sky_rad[b] > 0.0f) rad = c.sky_rad[b] * cam_sky_weight(d, c.sun, c.sky_params + 4 * b);
                if (c.emis[b] != 0u && c.cam_dflux[b] > 0.0f) rad += rdiv(c.cam_dflux[b], RT_PI_F);
                if (rad > 0.0f) c.pixels[(size_t)b * npix + pix] += rad * strength;
            }
        }
    }
    add_stats(stats, cnt, c.spp);
}

// Camera pixels with lane refill: a lane keeps its pixel and runs its samples in order, but a lane whose
// ray has ended shades it and starts its next sample while the others still traverse (traversal in node
// and leaf phases as the diffuse kernel).  Same rays, same per-pixel summation order as k_camera.
// Selectable at build time; measured slower on the orchard image (1,094 vs 1,460 Mrays/s): lanes drift to
// different samples and the wave loses the coherence of neighbouring pixels' rays, so it is off.
#ifndef RT_CAMERA_REFILL
#define RT_CAMERA_REFILL 0
#endif
template <int STK, bool COUNT, int MK, bool SPEC>
__global__ __launch_bounds__(256) void k_camera_r(CameraLaunch c, int *err, unsigned long long *stats) {
    constexpr int LSTK = STK < RT_CAMERA_LDS_STACK ? STK : RT_CAMERA_LDS_STACK;
    __shared__ uint32_t stk_all[LSTK * 256];
    uint32_t ovf[STK > LSTK ? RT_STACK - LSTK : 1];
    lds_u32 *stk = (lds_u32 *)stk_all + threadIdx.x;
    const uint32_t stride = 256;
    Cnt<COUNT, MK> cnt{};
    uint32_t nray = 0;
    const DevScene &sc = *c.sc;
    const float4 *nodes = sc.node, *recs = sc.rec;
    const uint32_t nleaf = sc.nleaf, nbbox = sc.nbbox;
    uint32_t px, py;
    const bool valid = camera_pixel(c, px, py);
    const uint32_t ii = c.off_x + px, jj = c.off_y + py;
    const size_t pix = (size_t)jj * c.full_x + ii, npix = (size_t)c.full_x * c.full_y;
    const float strength = rdiv(1.0f, (float)c.spp);
    const uint32_t dimx = c.res_x * c.res_y, indx = c.res_x * py + px;
    LaneRay L;
    uint32_t smp = 0, wraps = 0;
    bool busy = false, trav = false;
    auto start = [&]() {
        uint32_t st = tea16(indx + dimx * smp, c.seed);
        const float Rx = rnd24(st), Ry = rnd24(st);
        float lx = 0.f, ly = 0.f;
        if (c.lens > 0.0f) unit_disk(rnd24(st), rnd24(st), lx, ly);
        camera_ray(c.pos[0], c.pos[1], c.pos[2], c.st, c.ct, c.sp, c.cp, c.focal, c.lens, c.aspect, c.viewplane, ii,
                   jj, c.full_x, c.full_y, Rx, Ry, lx, ly, L.o, L.d);
        wraps = 0;
        busy = true;
        nray++;
        trav = lane_begin(L, recs, nleaf, 1e-5f, RT_MISS, cnt);
    };
    if (valid && c.spp > 0) start();
    while (true) {
        if (busy && !trav) {
            // periodic wrap, up to 10 traces (camera_raygen.comp:266-285)
            bool done = true;
            if (nbbox) {
                const float bt = nearest_face(sc, L.o, L.d, L.best_t);
                if (bt > 0.f) {
                    L.o = wrap_at(sc, L.o, L.d, bt);
                    if (wraps < 9) {
                        wraps++;
                        trav = lane_begin(L, recs, nleaf, 1e-5f, RT_MISS, cnt);
                        done = false;
                    } else {
                        L.best = RT_MISS;
                    }
                }
            }
            if (done) {
                const uint32_t h = L.best;
                const v3f o = L.o, d = L.d;
                if (h != RT_MISS) {
                    const float4 n4 = sc.nrm[h];
                    const bool face = n4.x * d.x + n4.y * d.y + n4.z * d.z < 0.0f;
                    const bool top = face || sc.type[h] == 4u;
                    for (uint32_t b = 0; b < c.B; b++) {
                        const float srad = cam_source_radiance(c, o, d, b, L.best_t, false);
                        float surf = (top ? c.out_top : c.out_bot)[(size_t)h * c.B + b] * strength;
                        if (srad > 0.0f) surf += srad * strength;
                        if (SPEC) surf += cam_specular(c, o, d, L.best_t, h, b, mk3(n4.x, n4.y, n4.z));
                        c.pixels[(size_t)b * npix + pix] += rdiv(surf, RT_PI_F);
                    }
                } else {
                    for (uint32_t b = 0; b < c.B; b++) {
                        float rad = cam_source_radiance(c, o, d, b, 1e30f, true);
                        if (rad <= 0.0f && c.sky_rad[b] > 0.0f) rad = c.sky_rad[b] * cam_sky_weight(d, c.sun, c.sky_params + 4 * b);
                        if (c.emis[b] != 0u && c.cam_dflux[b] > 0.0f) rad += rdiv(c.cam_dflux[b], RT_PI_F);
                        if (rad > 0.0f) c.pixels[(size_t)b * npix + pix] += rad * strength;
                    }
                }
                smp++;
                if (smp < c.spp) start();
                else busy = false;
            }
        }
        if (__ballot(busy) == 0) break;
        // traverse until fewer than RT_REFILL_BELOW lanes still do while some lane has samples left to start
        const uint32_t floor_lanes = __ballot(busy && smp + 1 < c.spp) ? RT_REFILL_BELOW : 1u;
        while (true) {
            while (true) {
                const bool inner = trav && !(L.ni & RT_LEAF_BIT);
                const uint64_t mi = __ballot(inner);
                if (mi == 0 || (uint32_t)__popcll(__ballot(trav && !inner)) >= RT_LEAF_BATCH) break;
                if (inner && lane_node<STK, LSTK>(nodes, nleaf, L, stk, ovf, stride, err, cnt)) trav = false;
            }
            while (true) {
                const bool leaf = trav && (L.ni & RT_LEAF_BIT);
                if (__ballot(leaf) == 0) break;
                if (leaf && lane_leaf<STK, LSTK>(recs, L, 1e-5f, RT_MISS, stk, ovf, stride, cnt)) trav = false;
            }
            if ((uint32_t)__popcll(__ballot(trav)) < floor_lanes) break;
        }
    }
    add_stats(stats, cnt, nray);
}

// Pixel labels (pixel_label_raygen.comp:80-155): the pixel-centre pinhole ray's hit UUID + 1 (0 = sky) and
// its depth along the view direction (-1 on a miss).
template <int STK, int MK>
__global__ __launch_bounds__(256) void k_pixel_label(CameraLaunch c, int *err) {
    constexpr int LSTK = STK < RT_CAMERA_LDS_STACK ? STK : RT_CAMERA_LDS_STACK;
    __shared__ uint32_t stk_all[LSTK * 256];
    uint32_t ovf[STK > LSTK ? RT_STACK - LSTK : 1];
    lds_u32 *stk = (lds_u32 *)stk_all + threadIdx.x;
    uint32_t px, py;
    if (!camera_pixel(c, px, py)) return;
    const DevScene &sc = *c.sc;
    const uint32_t ii = c.off_x + px, jj = c.off_y + py;
    const size_t pix = (size_t)jj * c.full_x + ii;
    v3f o, d;
    camera_ray(c.pos[0], c.pos[1], c.pos[2], c.st, c.ct, c.sp, c.cp, c.focal, 0.0f, c.aspect, c.viewplane, ii, jj,
               c.full_x, c.full_y, 0.5f, 0.5f, 0.0f, 0.0f, o, d);
    float t, depth;
    NoCntT<MK> cnt{};
    const uint32_t h = camera_trace<STK>(sc, o, d, t, depth, stk, ovf, err, cnt);
    if (h != RT_MISS) {
        c.labels[pix] = sc.uuid[h] + 1u;
        const v3f fwd = camera_rotate(mk3(1.0f, 0.0f, 0.0f), c.st, c.ct, c.sp, c.cp);
        c.depths[pix] = fabsf(dot3(fwd, d)) * (depth + t);
    } else {
        c.depths[pix] = -1.0f;
    }
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

template <typename F>
static void with_flags(int fl, F &&f);

// ------------------------------------------------------------------------------------------------
// parity hooks: generated diffuse rays, and radiation-predicate traces of arbitrary rays
// ------------------------------------------------------------------------------------------------
__global__ void k_debug_rays(const DevScene *__restrict__ scp, uint32_t n, uint32_t seed, float *__restrict__ out) {
    const DevScene &sc = *scp;
    uint32_t nn = n * n;
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)sc.P * 2 * nn) return;
    uint32_t i = (uint32_t)(g / (2 * nn)), r = (uint32_t)(g % (2 * nn));
    uint32_t slot = r >= nn ? 1u : 0u, rr = r - slot * nn;
    uint32_t ty = sc.type[i];
    bool patch = ty == 0 || ty == 3;
    uint32_t tw = __float_as_uint(sc.nrm[i].w);
    float4 fr = sc.frame[i];
    float spx, spy;
    v3f dl;
    float *w = out + g * 6;
    const uint32_t mdi = sc.mdesc ? sc.mdesc[i] : RT_MISS;
    if (!diffuse_sample_m(i, rr % n, rr / n, n, seed, patch, sc.rec, mdi, spx, spy, dl)) {
        for (int k = 0; k < 6; k++) w[k] = __uint_as_float(0x7FC00000u);  // dropped ray: NaN
        return;
    }
    v3f o = xform3(sc.T + (size_t)i * 16, spx, spy, 0.0f);
    v3f dir = frame_rotate(dl, fr.x, fr.y, fr.z, fr.w);
    if (slot == 1 && tw == 1) dir = mk3(-dir.x, -dir.y, -dir.z);
    w[0] = o.x; w[1] = o.y; w[2] = o.z; w[3] = dir.x; w[4] = dir.y; w[5] = dir.z;
}

template <int MK>
__global__ __launch_bounds__(256) void k_debug_trace(const DevScene *__restrict__ scp, uint32_t n,
                                                     const float *__restrict__ rays, const uint32_t *__restrict__ skip,
                                                     float tmin, int brute, uint32_t *__restrict__ hit,
                                                     float *__restrict__ tout, int *err) {
    __shared__ uint32_t stk_all[RT_STACK * 256];
    TCntT<MK> cnt{0, 0};
    uint32_t *stk = stk_all + threadIdx.x;
    const DevScene &sc = *scp;
    uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    v3f o = mk3(rays[r * 6], rays[r * 6 + 1], rays[r * 6 + 2]);
    v3f d = mk3(rays[r * 6 + 3], rays[r * 6 + 4], rays[r * 6 + 5]);
    float t = 1e30f;
    uint32_t h = RT_MISS, slot;
    if (brute) {
        for (uint32_t k = 0; k < sc.nleaf; k++) test_leaf<false, MK>(sc.rec, k, o, d, tmin, skip[r], t, h, slot);
    } else {
        h = trace_closest<false, RT_STACK>(sc, o, d, tmin, skip[r], 1e30f, t, slot, stk, 256, err, cnt);
    }
    hit[r] = h;
    tout[r] = t;
}

int launch_debug_rays(const DevScene *sc, uint32_t P, uint32_t n, uint32_t seed, float *out, hipStream_t st) {
    uint64_t tot = (uint64_t)P * 2 * n * n;
    if (tot == 0) return 0;
    k_debug_rays<<<(uint32_t)((tot + 255) / 256), 256, 0, st>>>(sc, n, seed, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error(std::string("kernel launch failed: ") + hipGetErrorString(e)); return -1; }
    return 0;
}

int launch_debug_trace(const DevScene *sc, uint32_t n, const float *rays, const uint32_t *skip, float tmin, int brute,
                       uint32_t *hit, float *t, int *err, hipStream_t st, int flags) {
    if (n == 0) return 0;
    with_flags(flags | RT_FL_MASK, [&](auto fl) {
        k_debug_trace<decltype(fl)::value><<<(n + 255) / 256, 256, 0, st>>>(sc, n, rays, skip, tmin, brute, hit, t, err);
    });
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error(std::string("kernel launch failed: ") + hipGetErrorString(e)); return -1; }
    return 0;
}

// ------------------------------------------------------------------------------------------------
// translucent covers (R/shaders/common/glass_cover.glsl, direct_raygen.comp:587-664,
// diffuse_raygen.comp:484-560)
// ------------------------------------------------------------------------------------------------
// The reference's trace loop with cover pass-through: up to 32 closest-hit traces; a periodic face wraps the
// ray; a miss -- or, with dist0 >= 0, a hit at or beyond dist0 -- ends it (status 0); a hit on a primitive that
// is not a cover stops it (status 1, that primitive returned); a cover calls push(cover, top face, band,
// transmittance so far, rho, alpha) for every band, multiplies ctau[b] by its transmittance at the crossing
// angle and the ray goes on from just past it.  After 32 traces the last trace's hit is returned (status 2).
template <int STK, typename C, typename Push>
__device__ __forceinline__ uint32_t trace_covers(const DevScene &sc, const GlassCtx &G, v3f o, v3f d, float tmin,
                                                 uint32_t skip, float dist0, float *ctau, uint32_t *stk,
                                                 uint32_t stride, int *err, C &cnt, int &status, Push &&push) {
    v3f cur = o;
    uint32_t h = RT_MISS;
    for (int step = 0; step < 32; ++step) {
        float t;
        uint32_t slot;
        h = trace_closest<false, STK>(sc, cur, d, tmin, skip, 1e30f, t, slot, stk, stride, err, cnt);
        if (sc.nbbox) {
            const float bt = nearest_face(sc, cur, d, t);
            if (bt > 0.f) { cur = wrap_at(sc, cur, d, bt); continue; }
        }
        if (h == RT_MISS) { status = 0; return RT_MISS; }
        if (dist0 >= 0.f && t >= dist0) { status = 0; return h; }
        if (!G.cover[h]) { status = 1; return h; }
        const float4 hn = sc.nrm[h];
        const v3f nh = mk3(hn.x, hn.y, hn.z);
        const float c = fabsf(dot3(d, nh));
        const bool top = dot3(nh, d) < 0.0f;
        for (uint32_t b = 0; b < G.B; b++) {
            float tau, rho, alpha;
            glass_tra(c, G.gn[(size_t)h * G.B + b], G.gkl[(size_t)h * G.B + b], tau, rho, alpha);
            push(h, top, b, ctau[b], rho, alpha);
            ctau[b] *= tau;
        }
        cur = add3(cur, mul3(d, t + 1e-4f));
    }
    status = 2;
    return h;
}

// pushed share of a cover: FP64 atomic into GlassCtx.push[cover][channel][slot]
__device__ __forceinline__ void glass_push(const GlassCtx &G, uint32_t h, uint32_t c, bool top, float inc, float rho,
                                           float alpha) {
    if (!(inc > 0.0f)) return;
    double *q = G.push + ((size_t)h * G.C + c) * 3;
    atomicAdd(q + 0, (double)(inc * alpha));
    atomicAdd(q + (top ? 1 : 2), (double)(inc * rho));
}

// Direct rays of a solve with covers (direct_raygen.comp:413-731 with the glass loop): one wave per (job,
// primitive); d*d stratified origins, every source type; an unblocked ray adds strength x (transmittance of the
// covers crossed, per band) to its lit face's sums, and every cover crossed takes its shares of
// flux x strength x transmittance so far.  gsum[i][b][face] of each job is deposited like the strength sums.
template <int STK, int MK>
__global__ __launch_bounds__(256) void k_direct_glass(const GlassDirectJob *__restrict__ jobs, uint32_t njobs,
                                                      uint32_t total, uint32_t *counter, const GlassCtx *__restrict__ gp,
                                                      int *err) {
    __shared__ uint32_t stk_all[STK * 256];
    uint32_t *stk = stk_all + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    const GlassCtx G = *gp;
    NoCntT<MK> cnt{};
    for (uint32_t item = wave_item(counter, 1u); item < total; item = wave_item(counter, 1u)) {
        uint32_t j = 0;
        {
            uint32_t lo = 0, hi = njobs;  // last job with dj.item_begin <= item
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (jobs[mid].dj.item_begin <= item) lo = mid; else hi = mid;
            }
            j = lo;
        }
        const GlassDirectJob &GJ = jobs[j];
        const DirectJob &J = GJ.dj;
        const DevScene &sc = *J.sc;
        uint32_t i = item - J.item_begin;
        if (J.order) i = J.order[i];
        const uint32_t ty = sc.type[i];
        ksum acc[2][RT_GLASS_MAX_CH];
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int b = 0; b < RT_GLASS_MAX_CH; b++) acc[f][b].init();
        if (traced_type(ty)) {
            const float4 n4 = sc.nrm[i];
            const v3f n = mk3(n4.x, n4.y, n4.z);
            const uint32_t tw = __float_as_uint(n4.w);
            const bool patch = ty == 0 || ty == 3;
            const uint32_t d = J.d, Nrays = d * d;
            float tm[12];
            const float *Tp = sc.T + (size_t)i * 16;
#pragma unroll
            for (int q = 0; q < 12; q++) tm[q] = Tp[q];
            const v3f sp3 = mk3(J.sdir[0], J.sdir[1], J.sdir[2]);
            for (uint32_t base = 0; base < Nrays; base += 64) {
                const uint32_t r = base + lane;
                if (r >= Nrays) continue;
                float spx, spy;
                uint32_t st;
                direct_sample_st(i, r % d, r / d, d, J.seed, patch, spx, spy, st);
                const uint32_t mdi = (MK && sc.mdesc) ? sc.mdesc[i] : RT_MISS;
                if (MK && mdi != RT_MISS && !origin_mask(sc.rec, mdi, patch, false, st, spx, spy)) continue;
                const v3f o = xform3(tm, spx, spy, 0.0f);
                for (uint32_t k = 0; k < J.skip_draws; k++) (void)rnd24(st);
                v3f dir;
                float strength;
                float dist0 = -1.0f;
                if (J.stype == 0u) {
                    dir = nrm3(sp3);
                    strength = rdiv(fabsf(dot3(n, dir)), (float)Nrays);
                } else {
                    if (!source_ray(J.stype, J.sdir, J.width, J.rot, o, n, Nrays, st, dir, strength)) continue;
                    const v3f dv = sub3(sp3, o);
                    dist0 = rsqrt_cr(dot3(dv, dv));
                }
                const bool top = dot3(dir, n) > 0.0f;
                if (tw == 3 || (!top && tw == 0)) continue;
                float ctau[RT_GLASS_MAX_CH];
                for (uint32_t b = 0; b < G.B; b++) ctau[b] = 1.0f;
                int status;
                (void)trace_covers<STK>(sc, G, o, dir, 1e-4f, i, dist0, ctau, stk, 256u, err, cnt, status,
                                        [&](uint32_t h, bool ctop, uint32_t b, float tr, float rho, float alpha) {
                                            glass_push(G, h, GJ.chan0 + b, ctop, GJ.flux[b] * strength * tr, rho,
                                                       alpha);
                                        });
                if (status == 1) continue;  // an opaque primitive blocks the source
#pragma unroll
                for (int b = 0; b < RT_GLASS_MAX_CH; b++)
                    if ((uint32_t)b < G.B) {
                        if (top) acc[0][b].add(strength * ctau[b]);
                        else acc[1][b].add(strength * ctau[b]);
                    }
            }
        }
#pragma unroll
        for (int b = 0; b < RT_GLASS_MAX_CH; b++) {
            wave_reduce(acc[0][b]);
            wave_reduce(acc[1][b]);
        }
        if (lane == 0) {
            float *o = GJ.gsum + (size_t)i * G.B * 2;
#pragma unroll
            for (int b = 0; b < RT_GLASS_MAX_CH; b++)
                if ((uint32_t)b < G.B) { o[2 * b] = acc[0][b].value(); o[2 * b + 1] = acc[1][b].value(); }
        }
    }
}

// Diffuse / scattering pass of a solve with covers: trace and gather in one kernel (diffuse_raygen.comp:330-656
// with the glass loop).  One wave per primitive, both launched faces, rays in rounds of 64; a ray passes
// through covers, which take their shares of sky flux x strength x transmittance so far (only a pass with sky
// flux pushes anything); a sky ray brings the sky flux times the covers' transmittance, a surface hit the hit
// face's radiation_out unattenuated.  Each channel's sum per face is deposited as the gather deposits it.
template <int STK, int MK>
__global__ __launch_bounds__(256) void k_diffuse_glass(const DiffuseJob *__restrict__ fj, const GatherJob *__restrict__ gj,
                                                       uint32_t total, uint32_t *counter,
                                                       const GlassCtx *__restrict__ gp, int *err) {
    __shared__ uint32_t stk_all[STK * 256];
    uint32_t *stk = stk_all + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    const GlassCtx G = *gp;
    const DiffuseJob &F = *fj;
    const GatherJob &J = *gj;
    const DevScene &sc = *F.sc;
    const uint32_t n = F.n, nn = n * n, seed = F.seed, C = J.C;
    const float strength = rdiv(1.0f, (float)nn);
    NoCntT<MK> cnt{};
    for (uint32_t item = wave_item(counter, 1u); item < total; item = wave_item(counter, 1u)) {
        uint32_t i = item - F.item_begin;
        if (F.order) i = F.order[i];
        const uint32_t ty = sc.type[i];
        if (!traced_type(ty)) continue;
        const float4 n4 = sc.nrm[i];
        const uint32_t tw = __float_as_uint(n4.w);
        const bool patch = ty == 0 || ty == 3;
        const uint32_t mdi = (MK && sc.mdesc) ? sc.mdesc[i] : RT_MISS;
        const float4 fr = sc.frame[i];
        float tm[12];
        const float *Tp = sc.T + (size_t)i * 16;
#pragma unroll
        for (int q = 0; q < 12; q++) tm[q] = Tp[q];
        ksum acc[2][RT_GLASS_MAX_CH];
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int c = 0; c < RT_GLASS_MAX_CH; c++) acc[f][c].init();
        const uint32_t nslot = tw != 0 ? 2u : 1u;  // bottom face launched for two-sided primitives only
        for (uint32_t base = 0; base < nslot * nn; base += 64) {
            const uint32_t r = base + lane;
            if (r >= nslot * nn) continue;
            const uint32_t slot = r >= nn ? 1u : 0u, rr = r - slot * nn;
            // a launch of one face (faces bit 0: top, bit 1: bottom; 0: both) skips the other face's rays
            if (F.faces && !(F.faces & (1u << slot))) continue;
            float spx, spy;
            v3f dl;
            if (!diffuse_sample_m(i, rr % n, rr / n, n, seed, patch, sc.rec, mdi, spx, spy, dl)) continue;
            const v3f o = xform3(tm, spx, spy, 0.0f);
            v3f dir = frame_rotate(dl, fr.x, fr.y, fr.z, fr.w);
            if (slot == 1 && tw == 1) dir = mk3(-dir.x, -dir.y, -dir.z);
            float ctau[RT_GLASS_MAX_CH];
            for (uint32_t b = 0; b < G.B; b++) ctau[b] = 1.0f;
            int status;
            const uint32_t h = trace_covers<STK>(
                sc, G, o, dir, 1e-5f, i, -1.0f, ctau, stk, 256u, err, cnt, status,
                [&](uint32_t hc, bool ctop, uint32_t b, float tr, float rho, float alpha) {
                    for (uint32_t c = b; c < C; c += G.B)
                        if (J.active[c]) glass_push(G, hc, c, ctop, J.dflux[c] * strength * tr, rho, alpha);
                });
            if (h != RT_MISS) {
                const float4 hn = sc.nrm[h];
                const bool hit_top = dot3(mk3(hn.x, hn.y, hn.z), dir) < 0.0f;
                const float *src = (hit_top ? J.out_top : J.out_bot) + (size_t)h * C;
#pragma unroll
                for (int c = 0; c < RT_GLASS_MAX_CH; c++)
                    if ((uint32_t)c < C) {
                        if (slot) acc[1][c].add(src[c]);
                        else acc[0][c].add(src[c]);
                    }
            } else {
#pragma unroll
                for (int c = 0; c < RT_GLASS_MAX_CH; c++)
                    if ((uint32_t)c < C) {
                        float v = J.dflux[c];
                        if (J.sky) v = sky_weight(dir, J.sky + 9 * c) * v;
                        v = v * ctau[(uint32_t)c % G.B];
                        if (slot) acc[1][c].add(v);
                        else acc[0][c].add(v);
                    }
            }
        }
#pragma unroll
        for (int c = 0; c < RT_GLASS_MAX_CH; c++) {
            wave_reduce(acc[0][c]);
            wave_reduce(acc[1][c]);
        }
        if (lane < C) {
            float I1 = 0.f, I0 = 0.f;
#pragma unroll
            for (int c = 0; c < RT_GLASS_MAX_CH; c++)
                if ((uint32_t)c == lane) { I1 = acc[0][c].value(); I0 = acc[1][c].value(); }
            const uint32_t c = lane;
            if (J.active[c]) {
                I1 = I1 * strength;
                I0 = I0 * strength;
                const uint32_t b = c % J.B;
                const float rho = J.rho[(size_t)i * J.B + b], tau = J.tau[(size_t)i * J.B + b];
                const size_t q = (size_t)i * C + c;
                J.in[q] += (I1 + I0) * (1.0f - rho - tau);
                J.st[q] += I1 * rho + I0 * tau;
                J.sb[q] += I1 * tau + I0 * rho;
            }
        }
    }
}

// radiation_in / scatter_top / scatter_bottom += the covers' pushed shares (rounded once), which are cleared
__global__ void k_add_push(size_t PC, double *push, float *in, float *st, float *sb) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= PC) return;
    double *p = push + q * 3;
    if (p[0] != 0.0) in[q] += (float)p[0];
    if (p[1] != 0.0) st[q] += (float)p[1];
    if (p[2] != 0.0) sb[q] += (float)p[2];
    p[0] = 0.0;
    p[1] = 0.0;
    p[2] = 0.0;
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
static inline uint32_t grid_for(const LaunchCfg &cfg, int ncu, uint32_t total, uint32_t wpb) {
    uint64_t g = (uint64_t)cfg.blocks_per_cu * (uint64_t)ncu;
    uint64_t per_block = (uint64_t)wpb * std::max(cfg.chunk, 1u);
    uint64_t need = (total + per_block - 1) / per_block;
    if (g > need) g = need;
    if (g < 1) g = 1;
    return (uint32_t)g;
}

#define LCHECK()                                                                        \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            set_error(std::string("kernel launch failed: ") + hipGetErrorString(e_));    \
            return -1;                                                                  \
        }                                                                               \
    } while (0)

// LDS stack depth chosen from the BVH depth: 32 entries (32 KB per 256-lane block, 5 blocks per CU)
// when the tree allows it, else 64.
#define PICK_STK(depth, CALL32, CALL64) \
    do {                                 \
        if ((depth) + 1 <= 32) { CALL32; } else { CALL64; } \
    } while (0)

// Runs f(std::integral_constant<int, FL>) for the traversal flags fl (RT_FL_*): one kernel instance per flag set.
template <typename F>
static void with_flags(int fl, F &&f) {
    switch (fl & (RT_FL_MASK | RT_FL_WT)) {
        case 0: f(std::integral_constant<int, 0>{}); break;
        case RT_FL_MASK: f(std::integral_constant<int, RT_FL_MASK>{}); break;
        case RT_FL_WT: f(std::integral_constant<int, RT_FL_WT>{}); break;
        default: f(std::integral_constant<int, RT_FL_MASK | RT_FL_WT>{}); break;
    }
}
// the counting variants always compile the mask tests in (a scene without masks skips them at run time)
static inline int counting_flags(int fl) { return (fl & RT_FL_WT) | RT_FL_MASK; }

#ifndef RT_DIRECT_QUEUE
#define RT_DIRECT_QUEUE 1
#endif
template <int MK>
static int launch_direct_t(const DirectJob *jobs, uint32_t njobs, uint32_t total, uint32_t depth, uint32_t *counter,
                           int *err, unsigned long long *stats, const LaunchCfg &cfg, int ncu, hipStream_t st,
                           bool grids) {
    if (total == 0) return 0;
    const uint32_t ch = std::max(cfg.chunk, 1u);
    k_zero_u32<<<1, 1, 0, st>>>(counter);
    uint32_t wpb = cfg.waves_per_block;
    uint32_t g = grid_for(cfg, ncu, total, wpb);
    if (RT_DIRECT_QUEUE && grids && !stats) {
        switch (wpb) {
            case 1: k_direct_q<1, MK><<<g, 64, 0, st>>>(jobs, njobs, total, counter, ch); break;
            case 2: k_direct_q<2, MK><<<g, 128, 0, st>>>(jobs, njobs, total, counter, ch); break;
            default: k_direct_q<4, MK><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch); break;
        }
        LCHECK();
        return 0;
    }
    if (stats) {
        // counting variant: per-lane stacks, 4 waves per block
        PICK_STK(depth, (k_direct<4, 32, true, MK><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch, err, stats)),
                 (k_direct<4, 64, true, MK><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch, err, stats)));
        LCHECK();
        return 0;
    }
    switch (wpb) {
        case 1: PICK_STK(depth, (k_direct<1, 32, false, MK><<<g, 64, 0, st>>>(jobs, njobs, total, counter, ch, err, stats)),
                         (k_direct<1, 64, false, MK><<<g, 64, 0, st>>>(jobs, njobs, total, counter, ch, err, stats))); break;
        case 2: PICK_STK(depth, (k_direct<2, 32, false, MK><<<g, 128, 0, st>>>(jobs, njobs, total, counter, ch, err, stats)),
                         (k_direct<2, 64, false, MK><<<g, 128, 0, st>>>(jobs, njobs, total, counter, ch, err, stats))); break;
        default: PICK_STK(depth, (k_direct<4, 32, false, MK><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch, err, stats)),
                          (k_direct<4, 64, false, MK><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch, err, stats))); break;
    }
    LCHECK();
    return 0;
}

int launch_direct(const DirectJob *jobs, uint32_t njobs, uint32_t total, uint32_t depth, uint32_t *counter, int *err,
                  unsigned long long *stats, const LaunchCfg &cfg, int ncu, hipStream_t st, int flags) {
    int rc = 0;
    with_flags(stats ? counting_flags(flags) : flags, [&](auto fl) {
        rc = launch_direct_t<decltype(fl)::value>(jobs, njobs, total, depth, counter, err, stats, cfg, ncu, st,
                                                  (flags & RT_FL_GRIDS) != 0);
    });
    return rc;
}

template <int MK>
static int launch_direct_src_t(const DirectJob *jobs, uint32_t njobs, uint32_t item_lo, uint32_t nitems,
                               uint32_t depth, uint32_t *counter, int *err, unsigned long long *stats, int ncu,
                               hipStream_t st) {
    if (nitems == 0) return 0;
    k_zero_u32<<<1, 1, 0, st>>>(counter);
    uint64_t g64 = std::min<uint64_t>((uint64_t)4 * ncu, ((uint64_t)nitems + 3) / 4);
    uint32_t g = (uint32_t)std::max<uint64_t>(g64, 1);
    if (stats)
        PICK_STK(depth, (k_direct_src<32, true, MK><<<g, 256, 0, st>>>(jobs, njobs, item_lo, nitems, counter, err, stats)),
                 (k_direct_src<64, true, MK><<<g, 256, 0, st>>>(jobs, njobs, item_lo, nitems, counter, err, stats)));
    else
        PICK_STK(depth, (k_direct_src<32, false, MK><<<g, 256, 0, st>>>(jobs, njobs, item_lo, nitems, counter, err, stats)),
                 (k_direct_src<64, false, MK><<<g, 256, 0, st>>>(jobs, njobs, item_lo, nitems, counter, err, stats)));
    LCHECK();
    return 0;
}

int launch_direct_src(const DirectJob *jobs, uint32_t njobs, uint32_t item_lo, uint32_t nitems, uint32_t depth,
                      uint32_t *counter, int *err, unsigned long long *stats, int ncu, hipStream_t st, int flags) {
    int rc = 0;
    with_flags(stats ? counting_flags(flags) : flags, [&](auto fl) {
        rc = launch_direct_src_t<decltype(fl)::value>(jobs, njobs, item_lo, nitems, depth, counter, err, stats, ncu, st);
    });
    return rc;
}

int launch_direct_jobs(const DirectJob *hjobs, const DirectJob *jobs, uint32_t njobs, uint32_t total, uint32_t depth,
                       uint32_t *counter, int *err, unsigned long long *stats, const LaunchCfg &cfg, int ncu,
                       hipStream_t st, int flags) {
    uint32_t ncol = 0;
    while (ncol < njobs && hjobs[ncol].stype == 0) ncol++;
    for (uint32_t k = ncol; k < njobs; k++)
        if (hjobs[k].stype == 0) {
            set_error("internal: collimated direct jobs must precede the other source types");
            return -1;
        }
    // the collimated jobs' launch runs the queue kernel when each of them has an occlusion grid
    bool grids = ncol > 0;
    for (uint32_t k = 0; k < ncol; k++) grids = grids && hjobs[k].grid != nullptr;
    const int cflags = grids ? (flags | RT_FL_GRIDS) : (flags & ~RT_FL_GRIDS);
    if (ncol == njobs) return launch_direct(jobs, njobs, total, depth, counter, err, stats, cfg, ncu, st, cflags);
    if (ncol == 0) return launch_direct_src(jobs, njobs, 0, total, depth, counter, err, stats, ncu, st, flags);
    const uint32_t split = hjobs[ncol].item_begin;
    if (split > total) {
        set_error("internal: direct job items out of order");
        return -1;
    }
    if (launch_direct(jobs, ncol, split, depth, counter, err, stats, cfg, ncu, st, cflags)) return -1;
    return launch_direct_src(jobs + ncol, njobs - ncol, split, total - split, depth, counter, err, stats, ncu, st,
                             flags);
}

int launch_diffuse(const DiffuseJob *job, uint32_t total, uint32_t depth, uint32_t *counter, int *err,
                   unsigned long long *stats, const LaunchCfg &cfg, int ncu, hipStream_t st, int flags) {
    if (total == 0) return 0;
    // the stack's first RT_DIFFUSE_LDS_STACK entries live in LDS and the rest in private memory, so its
    // full depth (RT_STACK) costs no occupancy: one stack size serves the binary and the 4-wide tree
    if (depth + 1 > RT_STACK) {
        set_error("BVH depth " + std::to_string(depth) + " exceeds the traversal stack (" + std::to_string(RT_STACK) + ")");
        return -1;
    }
    const uint32_t ch = std::max(cfg.chunk, 1u);
    k_zero_u32<<<1, 1, 0, st>>>(counter);
    uint32_t wpb = cfg.waves_per_block;
    uint32_t g = grid_for(cfg, ncu, total, wpb);
    if (stats) {
        // counting variant: 4 waves per block
        with_flags(counting_flags(flags), [&](auto fl) {
            k_diffuse<4, RT_STACK, true, decltype(fl)::value><<<g, 256, 0, st>>>(job, total, counter, ch, err, stats);
        });
        LCHECK();
        return 0;
    }
    with_flags(flags, [&](auto fl) {
        constexpr int FL = decltype(fl)::value;
        switch (wpb) {
            case 1: k_diffuse<1, RT_STACK, false, FL><<<g, 64, 0, st>>>(job, total, counter, ch, err, stats); break;
            case 2: k_diffuse<2, RT_STACK, false, FL><<<g, 128, 0, st>>>(job, total, counter, ch, err, stats); break;
            default: k_diffuse<4, RT_STACK, false, FL><<<g, 256, 0, st>>>(job, total, counter, ch, err, stats); break;
        }
    });
    LCHECK();
    return 0;
}

template <bool ACCT>
static int launch_gather_t(const GatherJob *jobs, uint32_t njobs, uint32_t total, uint32_t cmax, uint32_t *counter,
                           uint32_t g, uint32_t ch, hipStream_t st) {
    if (cmax <= 1) k_gather_rays<1, ACCT><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch);
    else if (cmax <= 2) k_gather_rays<2, ACCT><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch);
    else if (cmax <= 4) k_gather_rays<4, ACCT><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch);
    else if (cmax <= 8) k_gather_rays<8, ACCT><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch);
    else if (cmax <= 64) k_gather_chan<1, ACCT><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch);
    else if (cmax <= 128) k_gather_chan<2, ACCT><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch);
    else if (cmax <= 192) k_gather_chan<3, ACCT><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch);
    else if (cmax <= 256) k_gather_chan<4, ACCT><<<g, 256, 0, st>>>(jobs, njobs, total, counter, ch);
    else {
        set_error("more than 256 channels in one gather");
        return -1;
    }
    LCHECK();
    return 0;
}

int launch_gather(const GatherJob *jobs, uint32_t njobs, uint32_t total, uint32_t cmax, uint32_t *counter,
                  const LaunchCfg &cfg, int ncu, hipStream_t st, bool acct) {
    if (total == 0) return 0;
    const uint32_t ch = std::max(cfg.chunk, 1u);
    k_zero_u32<<<1, 1, 0, st>>>(counter);
    uint32_t g = grid_for(cfg, ncu, total, 4);
    return acct ? launch_gather_t<true>(jobs, njobs, total, cmax, counter, g, ch, st)
                : launch_gather_t<false>(jobs, njobs, total, cmax, counter, g, ch, st);
}

int launch_camera(const CameraLaunch &c, uint32_t depth, bool labels, int *err, unsigned long long *stats,
                  hipStream_t st, int flags) {
    if ((uint64_t)c.res_x * c.res_y == 0) return 0;
    // one wave per 8 x 8 pixel block, four waves per workgroup
    const uint64_t blocks = (uint64_t)((c.res_x + 7) / 8) * ((c.res_y + 7) / 8);
    const uint32_t g = (uint32_t)((blocks + 3) / 4);
#if RT_CAMERA_REFILL
#define K_CAM k_camera_r
#else
#define K_CAM k_camera
#endif
    if (labels)
        with_flags(flags, [&](auto fl) {
            constexpr int FL = decltype(fl)::value;
            PICK_STK(depth, (k_pixel_label<32, FL><<<g, 256, 0, st>>>(c, err)),
                     (k_pixel_label<64, FL><<<g, 256, 0, st>>>(c, err)));
        });
    else if (stats)
        with_flags(counting_flags(flags), [&](auto fl) {
            constexpr int FL = decltype(fl)::value;
            if (c.spec)
                PICK_STK(depth, (K_CAM<32, true, FL, true><<<g, 256, 0, st>>>(c, err, stats)),
                         (K_CAM<64, true, FL, true><<<g, 256, 0, st>>>(c, err, stats)));
            else
                PICK_STK(depth, (K_CAM<32, true, FL, false><<<g, 256, 0, st>>>(c, err, stats)),
                         (K_CAM<64, true, FL, false><<<g, 256, 0, st>>>(c, err, stats)));
        });
    else
        // the specular term is a separate instantiation: inlined unconditionally it costs the plain
        // kernel 3 VGPRs and one wave per SIMD (78 -> 81 registers)
        with_flags(flags, [&](auto fl) {
            constexpr int FL = decltype(fl)::value;
            if (c.spec)
                PICK_STK(depth, (K_CAM<32, false, FL, true><<<g, 256, 0, st>>>(c, err, nullptr)),
                         (K_CAM<64, false, FL, true><<<g, 256, 0, st>>>(c, err, nullptr)));
            else
                PICK_STK(depth, (K_CAM<32, false, FL, false><<<g, 256, 0, st>>>(c, err, nullptr)),
                         (K_CAM<64, false, FL, false><<<g, 256, 0, st>>>(c, err, nullptr)));
        });
#undef K_CAM
    LCHECK();
    return 0;
}

int launch_direct_glass(const GlassDirectJob *jobs, uint32_t njobs, uint32_t total, uint32_t depth, uint32_t *counter,
                        const GlassCtx *g, int *err, int ncu, hipStream_t st, int flags) {
    if (total == 0) return 0;
    if (depth + 1 > RT_STACK) {
        set_error("BVH depth " + std::to_string(depth) + " exceeds the traversal stack (" + std::to_string(RT_STACK) + ")");
        return -1;
    }
    k_zero_u32<<<1, 1, 0, st>>>(counter);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)4 * ncu, (total + 3) / 4);
    with_flags(flags, [&](auto fl) {
        constexpr int FL = decltype(fl)::value;
        PICK_STK(depth, (k_direct_glass<32, FL><<<grid, 256, 0, st>>>(jobs, njobs, total, counter, g, err)),
                 (k_direct_glass<RT_STACK, FL><<<grid, 256, 0, st>>>(jobs, njobs, total, counter, g, err)));
    });
    LCHECK();
    return 0;
}

int launch_diffuse_glass(const DiffuseJob *fj, const GatherJob *gj, uint32_t total, uint32_t depth, uint32_t *counter,
                         const GlassCtx *g, int *err, int ncu, hipStream_t st, int flags) {
    if (total == 0) return 0;
    if (depth + 1 > RT_STACK) {
        set_error("BVH depth " + std::to_string(depth) + " exceeds the traversal stack (" + std::to_string(RT_STACK) + ")");
        return -1;
    }
    k_zero_u32<<<1, 1, 0, st>>>(counter);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)4 * ncu, (total + 3) / 4);
    with_flags(flags, [&](auto fl) {
        constexpr int FL = decltype(fl)::value;
        PICK_STK(depth, (k_diffuse_glass<32, FL><<<grid, 256, 0, st>>>(fj, gj, total, counter, g, err)),
                 (k_diffuse_glass<RT_STACK, FL><<<grid, 256, 0, st>>>(fj, gj, total, counter, g, err)));
    });
    LCHECK();
    return 0;
}

int launch_add_push(size_t PC, double *push, float *in, float *stb, float *sbb, hipStream_t st) {
    if (PC == 0) return 0;
    k_add_push<<<(uint32_t)((PC + 255) / 256), 256, 0, st>>>(PC, push, in, stb, sbb);
    LCHECK();
    return 0;
}

int launch_cast(const DevScene *sc, uint32_t depth, uint64_t n, const float *o, const float *d, float maxd, float *t,
                float *nrm, uint32_t *uuid, unsigned long long *counter, int *err, unsigned long long *stats,
                int ncu, hipStream_t st, bool watertight) {
    if (n == 0) return 0;
    if (depth + 1 > RT_STACK) {
        set_error("BVH depth " + std::to_string(depth) + " exceeds the traversal stack (" + std::to_string(RT_STACK) + ")");
        return -1;
    }
    k_zero_u64<<<1, 1, 0, st>>>(counter);
    const uint64_t chunks = (n + 63) / 64;
    uint64_t g = std::min<uint64_t>((uint64_t)8 * ncu, (chunks + 3) / 4);
    if (g < 1) g = 1;
    // castRays ignores texture masks (as CollisionDetection does); only the predicate flag applies
    if (watertight) {
        if (stats) k_cast<RT_STACK, true, RT_FL_WT><<<(uint32_t)g, 256, 0, st>>>(sc, n, o, d, maxd, t, nrm, uuid, counter, err, stats);
        else k_cast<RT_STACK, false, RT_FL_WT><<<(uint32_t)g, 256, 0, st>>>(sc, n, o, d, maxd, t, nrm, uuid, counter, err, stats);
    } else {
        if (stats) k_cast<RT_STACK, true, 0><<<(uint32_t)g, 256, 0, st>>>(sc, n, o, d, maxd, t, nrm, uuid, counter, err, stats);
        else k_cast<RT_STACK, false, 0><<<(uint32_t)g, 256, 0, st>>>(sc, n, o, d, maxd, t, nrm, uuid, counter, err, stats);
    }
    LCHECK();
    return 0;
}

}  // namespace rt
