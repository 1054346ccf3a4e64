// rt_trace.hip -- BVH traversal and the wavefront-persistent ray kernels (pieces b, c, d).
//
//   trace_closest   stack traversal of the 64-byte BVH2 (both child boxes per node), LDS stack,
//                   closest hit with the reference acceptance rule (t > tmin, skip the origin
//                   primitive; R/shaders/common/bvh_traversal.glsl:68-170) and a canonical tie rule
//                   (equal t -> lower primitive index) instead of traversal order.
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

#include <cfloat>
#include <string>

#include "rt_internal.h"

namespace rt {

__device__ __forceinline__ float min3f(float a, float b, float c) { return fminf(fminf(a, b), c); }
__device__ __forceinline__ float max3f(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }

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

template <bool CD>
__device__ __forceinline__ void test_leaf(const DevScene &sc, uint32_t k, v3f o, v3f d, float tmin, uint32_t skip,
                                          float &best_t, uint32_t &best, uint32_t &best_slot) {
    const float4 *r = sc.rec + (size_t)k * 4;
    float4 r0 = r[0];
    uint32_t tag = __float_as_uint(r0.w);
    uint32_t id = tag & ~RT_PATCH_BIT;
    if (id == skip) return;
    bool patch = (tag & RT_PATCH_BIT) != 0;
    float4 r1 = r[1], r2 = r[2];
    float4 r3 = patch ? r[3] : make_float4(0.f, 0.f, 0.f, 0.f);
    float t;
    if (isect_rec<CD>(r0, r1, r2, r3, patch, o, d, t) && t > tmin) {
        if (t < best_t || (t == best_t && best != RT_MISS && id < best)) {
            best_t = t;
            best = id;
            best_slot = k;
        }
    }
}

template <bool CD>
__device__ uint32_t trace_closest(const DevScene &sc, v3f o, v3f d, float tmin, uint32_t skip, float tstart,
                                  float &t_out, uint32_t &slot_out, uint32_t *stk, uint32_t stride, int *err) {
    float best_t = tstart;
    uint32_t best = RT_MISS, best_slot = RT_MISS;
    if (sc.nleaf <= 1) {
        if (sc.nleaf == 1) test_leaf<CD>(sc, 0, o, d, tmin, skip, best_t, best, best_slot);
        t_out = best_t;
        slot_out = best_slot;
        return best;
    }
    RayPre rp = ray_pre(o, d);
    uint32_t ni = 0;
    int sp = 0;
    // every internal node is entered at most once per trace; the cap only guards malformed trees
    uint32_t budget = sc.nleaf + 8;
    while (true) {
        if (budget-- == 0 || ni >= sc.nleaf - 1) {
            atomicOr(err, 2);
            break;
        }
        const float4 *nd = sc.node + (size_t)ni * 4;
        float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
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
        bool h0 = tn0 <= tf0 && tf0 >= 0.f && tn0 <= best_t;
        bool h1 = tn1 <= tf1 && tf1 >= 0.f && tn1 <= best_t;
        uint32_t c0 = __float_as_uint(n3.x), c1 = __float_as_uint(n3.y);
        bool l0 = (c0 & RT_LEAF_BIT) != 0, l1 = (c1 & RT_LEAF_BIT) != 0;
        if (h0 && l0) test_leaf<CD>(sc, c0 & ~RT_LEAF_BIT, o, d, tmin, skip, best_t, best, best_slot);
        if (h1 && l1) test_leaf<CD>(sc, c1 & ~RT_LEAF_BIT, o, d, tmin, skip, best_t, best, best_slot);
        bool g0 = h0 && !l0, g1 = h1 && !l1;
        if (g0 && g1) {
            bool first0 = tn0 <= tn1;
            if (sp >= RT_STACK) {
                atomicOr(err, 1);
                break;
            }
            stk[(uint32_t)sp * stride] = first0 ? c1 : c0;
            sp++;
            ni = first0 ? c0 : c1;
        } else if (g0) {
            ni = c0;
        } else if (g1) {
            ni = c1;
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

// R/shaders/common/periodic_boundary.glsl:73-112
__device__ __forceinline__ bool periodic_wrap(const DevScene &sc, v3f o, v3f d, float bvh_t, v3f &w) {
    float bt = bvh_t;
    int cf = -1;
    for (uint32_t f = 0; f < sc.nbbox; f++) {
        float t = isect_bbox_face(sc.bbox[f], o, d);
        if (t > 0.0f && t < bt) { bt = t; cf = (int)f; }
    }
    if (cf < 0) return false;
    v3f hp = add3(o, mul3(d, bt));
    const float eps = 1e-5f;
    float wx = sc.dom[1] - sc.dom[0], wy = sc.dom[3] - sc.dom[2];
    w = hp;
    if (sc.px && fabsf(hp.x - sc.dom[0]) <= eps) w.x += wx - eps;
    else if (sc.px && fabsf(hp.x - sc.dom[1]) <= eps) w.x -= wx - eps;
    else if (sc.py && fabsf(hp.y - sc.dom[2]) <= eps) w.y += wy - eps;
    else if (sc.py && fabsf(hp.y - sc.dom[3]) <= eps) w.y -= wy - eps;
    return true;
}

// Up to 32 traces with a periodic-face check after each (direct_raygen.comp:603-668,
// diffuse_raygen.comp:498-560, without translucent covers).  Returns the last traversal result.
__device__ __forceinline__ uint32_t trace_loop(const DevScene &sc, v3f o, v3f d, float tmin, uint32_t skip,
                                               bool &exhausted, uint32_t *stk, uint32_t stride, int *err) {
    uint32_t hit = RT_MISS;
    v3f cur = o;
    exhausted = true;
    for (int step = 0; step < 32; ++step) {
        float t;
        uint32_t slot;
        hit = trace_closest<false>(sc, cur, d, tmin, skip, 1e30f, t, slot, stk, stride, err);
        if (sc.nbbox) {
            v3f w;
            if (periodic_wrap(sc, cur, d, t, w)) { cur = w; continue; }
        }
        exhausted = false;
        break;
    }
    return hit;
}

// Next work item of this wave.  No lane-divergent branch around the fetch: every lane issues the
// atomic (lane 0 adds 1, the others 0) and lane 0's old value is broadcast as a scalar.  A
// `if (lane == 0) atomicAdd` + shuffle lets the compiler split the persistent loop into a lane-0
// loop and a lanes-1..63 loop that re-processes item 0 forever.
__device__ __forceinline__ uint32_t wave_item(uint32_t *counter) {
    uint32_t lane = threadIdx.x & 63;
    uint32_t item = atomicAdd(counter, lane == 0 ? 1u : 0u);
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

__device__ __forceinline__ void wave_reduce(ksum &k) {
    for (int off = 32; off >= 1; off >>= 1) {
        float os = __shfl_xor(k.s, off, 64);
        float oc = __shfl_xor(k.c, off, 64);
        k.merge(os, oc);
    }
}

__device__ __forceinline__ bool traced_type(uint32_t t) { return t == 0 || t == 1 || t == 3; }

// ------------------------------------------------------------------------------------------------
// direct rays
// ------------------------------------------------------------------------------------------------
template <int WPB>
__global__ __launch_bounds__(WPB * 64) void k_direct(const DirectJob *__restrict__ jobs, uint32_t njobs, uint32_t total,
                                                     uint32_t *counter, int *err) {
    __shared__ uint32_t stk_all[RT_STACK * WPB * 64];
    uint32_t *stk = stk_all + threadIdx.x;
    const uint32_t stride = WPB * 64;
    const uint32_t lane = threadIdx.x & 63;
    while (true) {
        uint32_t item = wave_item(counter);
        if (item >= total) break;
        uint32_t j = find_job(jobs, njobs, item);
        const DirectJob &J = jobs[j];
        const DevScene &sc = *J.sc;
        uint32_t i = item - J.item_begin;
        uint32_t ty = sc.type[i];
        float sum_out = 0.f;
        bool top = true;
        if (traced_type(ty)) {
            float4 n4 = sc.nrm[i];
            v3f n = mk3(n4.x, n4.y, n4.z);
            uint32_t tw = __float_as_uint(n4.w);
            v3f dir = nrm3(mk3(J.sdir[0], J.sdir[1], J.sdir[2]));
            uint32_t d = J.d, Nrays = d * d;
            float strength = rdiv(fabsf(dot3(n, dir)), (float)Nrays);
            top = dot3(dir, n) > 0.0f;
            if (!(tw == 3 || (!top && tw == 0))) {
                bool patch = ty == 0 || ty == 3;
                float tm[12];
                const float *Tp = sc.T + (size_t)i * 16;
#pragma unroll
                for (int q = 0; q < 12; q++) tm[q] = Tp[q];
                ksum acc;
                acc.init();
                for (uint32_t r = lane; r < Nrays; r += 64) {
                    uint32_t lx = r % d, ly = r / d;
                    float spx, spy;
                    direct_sample(i, lx, ly, d, J.seed, patch, spx, spy);
                    v3f o = xform3(tm, spx, spy, 0.0f);
                    bool ex;
                    uint32_t h = trace_loop(sc, o, dir, 1e-4f, i, ex, stk, stride, err);
                    if (h == RT_MISS || ex) acc.add(strength);
                }
                wave_reduce(acc);
                sum_out = acc.value();
            }
        }
        if (lane == 0) {
            J.out[(size_t)i * 2 + 0] = top ? sum_out : 0.f;
            J.out[(size_t)i * 2 + 1] = top ? 0.f : sum_out;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// diffuse / scattering rays -> hit words
// ------------------------------------------------------------------------------------------------
template <int WPB>
__global__ __launch_bounds__(WPB * 64) void k_diffuse(const DiffuseJob *__restrict__ jobs, uint32_t njobs,
                                                      uint32_t total, uint32_t *counter, int *err) {
    __shared__ uint32_t stk_all[RT_STACK * WPB * 64];
    uint32_t *stk = stk_all + threadIdx.x;
    const uint32_t stride = WPB * 64;
    const uint32_t lane = threadIdx.x & 63;
    while (true) {
        uint32_t item = wave_item(counter);
        if (item >= total) break;
        uint32_t j = find_job(jobs, njobs, item);
        const DiffuseJob &J = jobs[j];
        const DevScene &sc = *J.sc;
        uint32_t i = item - J.item_begin;
        uint32_t n = J.n, nn = n * n;
        uint32_t *w = J.hits + (size_t)i * 2 * nn;
        uint32_t ty = sc.type[i];
        if (!traced_type(ty)) {
            for (uint32_t r = lane; r < 2 * nn; r += 64) w[r] = RT_NOT_LAUNCHED;
            continue;
        }
        uint32_t tw = __float_as_uint(sc.nrm[i].w);
        float4 fr = sc.frame[i];
        bool patch = ty == 0 || ty == 3;
        float tm[12];
        const float *Tp = sc.T + (size_t)i * 16;
#pragma unroll
        for (int q = 0; q < 12; q++) tm[q] = Tp[q];
        for (uint32_t r = lane; r < 2 * nn; r += 64) {
            uint32_t slot = r >= nn ? 1u : 0u;  // slot 0 = top launch (face 1), slot 1 = bottom (face 0)
            uint32_t rr = r - slot * nn;
            if (slot == 1 && tw == 0) { w[r] = RT_NOT_LAUNCHED; continue; }
            float spx, spy;
            v3f dl;
            diffuse_sample(i, rr % n, rr / n, n, J.seed, patch, spx, spy, dl);
            v3f o = xform3(tm, spx, spy, 0.0f);
            v3f dir = frame_rotate(dl, fr.x, fr.y, fr.z, fr.w);
            if (slot == 1 && tw == 1) dir = mk3(-dir.x, -dir.y, -dir.z);
            bool ex;
            uint32_t h = trace_loop(sc, o, dir, 1e-5f, i, ex, stk, stride, err);
            uint32_t word = RT_MISS;
            if (h != RT_MISS) {
                float4 hn = sc.nrm[h];
                bool hit_top = dot3(mk3(hn.x, hn.y, hn.z), dir) < 0.0f;
                word = h | (hit_top ? 0u : RT_HIT_BOTTOM);
            }
            w[r] = word;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// gather: lanes over rays (few channels)
// ------------------------------------------------------------------------------------------------
template <int CM>
__global__ __launch_bounds__(256) void k_gather_rays(const GatherJob *__restrict__ jobs, uint32_t njobs, uint32_t total,
                                                     uint32_t *counter) {
    const uint32_t lane = threadIdx.x & 63;
    while (true) {
        uint32_t item = wave_item(counter);
        if (item >= total) break;
        uint32_t j = find_job(jobs, njobs, item);
        const GatherJob &J = jobs[j];
        uint32_t i = item - J.item_begin;
        uint32_t ty = J.sc->type[i];
        if (!traced_type(ty)) continue;
        uint32_t nn = J.n * J.n, C = J.C;
        const uint32_t *w = J.hits + (size_t)i * 2 * nn;
        ksum acc[2][CM];
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int c = 0; c < CM; c++) acc[f][c].init();
        for (uint32_t r = lane; r < 2 * nn; r += 64) {
            uint32_t word = w[r];
            if (word == RT_NOT_LAUNCHED) continue;
            uint32_t slot = r >= nn ? 1u : 0u;
            if (word == RT_MISS) {
#pragma unroll
                for (int c = 0; c < CM; c++)
                    if ((uint32_t)c < C) {
                        float v = J.dflux[c];
                        if (slot) acc[1][c].add(v); else acc[0][c].add(v);
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
        }
        if (lane < C && lane < CM) {
            float I1 = 0.f, I0 = 0.f;
#pragma unroll
            for (int c = 0; c < CM; c++)
                if ((uint32_t)c == lane) { I1 = acc[0][c].value(); I0 = acc[1][c].value(); }
            uint32_t c = lane;
            if (J.active[c]) {
                I1 = I1 * strength;
                I0 = I0 * strength;
                uint32_t b = c % J.B;
                float rho = J.rho[(size_t)i * J.B + b], tau = J.tau[(size_t)i * J.B + b];
                size_t q = (size_t)i * C + c;
                J.in[q] += (I1 + I0) * (1.0f - rho - tau);
                J.st[q] += I1 * rho + I0 * tau;
                J.sb[q] += I1 * tau + I0 * rho;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// gather: lanes over channels (many channels, e.g. suns x bands of a diurnal batch)
// ------------------------------------------------------------------------------------------------
template <int CB>
__global__ __launch_bounds__(256) void k_gather_chan(const GatherJob *__restrict__ jobs, uint32_t njobs, uint32_t total,
                                                     uint32_t *counter) {
    const uint32_t lane = threadIdx.x & 63;
    while (true) {
        uint32_t item = wave_item(counter);
        if (item >= total) break;
        uint32_t j = find_job(jobs, njobs, item);
        const GatherJob &J = jobs[j];
        uint32_t i = item - J.item_begin;
        uint32_t ty = J.sc->type[i];
        if (!traced_type(ty)) continue;
        uint32_t nn = J.n * J.n, C = J.C;
        const uint32_t *w = J.hits + (size_t)i * 2 * nn;
        ksum acc[2][CB];
        float sky[CB];
#pragma unroll
        for (int k = 0; k < CB; k++) {
            acc[0][k].init();
            acc[1][k].init();
            uint32_t c = lane + 64 * k;
            sky[k] = c < C ? J.dflux[c] : 0.f;
        }
        for (uint32_t base = 0; base < 2 * nn; base += 64) {
            uint32_t mine = base + lane < 2 * nn ? w[base + lane] : RT_NOT_LAUNCHED;
            uint32_t cnt = min(64u, 2 * nn - base);
            for (uint32_t q = 0; q < cnt; q++) {
                uint32_t word = __builtin_amdgcn_readlane(mine, q);
                if (word == RT_NOT_LAUNCHED) continue;
                uint32_t slot = base + q >= nn ? 1u : 0u;
                if (word == RT_MISS) {
#pragma unroll
                    for (int k = 0; k < CB; k++) {
                        if (slot) acc[1][k].add(sky[k]); else acc[0][k].add(sky[k]);
                    }
                } else {
                    uint32_t h = word & ~RT_HIT_BOTTOM;
                    const float *src = ((word & RT_HIT_BOTTOM) ? J.out_bot : J.out_top) + (size_t)h * C;
#pragma unroll
                    for (int k = 0; k < CB; k++) {
                        uint32_t c = lane + 64 * k;
                        float v = c < C ? src[c] : 0.f;
                        if (slot) acc[1][k].add(v); else acc[0][k].add(v);
                    }
                }
            }
        }
        float strength = rdiv(1.0f, (float)nn);
#pragma unroll
        for (int k = 0; k < CB; k++) {
            uint32_t c = lane + 64 * k;
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

// ------------------------------------------------------------------------------------------------
// castRaysSoA
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cast(const DevScene *__restrict__ scp, uint64_t n, const float *__restrict__ O,
                                              const float *__restrict__ D, float maxd, float *__restrict__ T,
                                              float *__restrict__ N, uint32_t *__restrict__ U, int *err) {
    __shared__ uint32_t stk_all[RT_STACK * 256];
    uint32_t *stk = stk_all + threadIdx.x;
    const DevScene &sc = *scp;
    for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (uint64_t)gridDim.x * 256) {
        v3f o = mk3(O[r * 3], O[r * 3 + 1], O[r * 3 + 2]);
        v3f d = mk3(D[r * 3], D[r * 3 + 1], D[r * 3 + 2]);
        float tstart = maxd > 0.f ? maxd : FLT_MAX;
        float t;
        uint32_t slot;
        uint32_t h = trace_closest<true>(sc, o, d, 1e-6f, RT_MISS, tstart, t, slot, stk, 256, err);
        if (h == RT_MISS) { U[r] = RT_MISS; continue; }
        U[r] = sc.uuid[h];
        T[r] = t;
        const float4 *rr = sc.rec + (size_t)slot * 4;
        float4 r0 = rr[0], r1 = rr[1], r2 = rr[2], r3 = rr[3];
        bool patch = (__float_as_uint(r0.w) & RT_PATCH_BIT) != 0;
        // CD/src/CollisionDetection_RayTracing.cpp:1572-1585: normal of cross(v1-v0, v2-v0), facing the origin
        v3f nn;
        bool ok;
        if (patch) {
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
            v3f ip = add3(o, mul3(d, t));
            v3f to = sub3(o, ip);
            if (nn.x * to.x + nn.y * to.y + nn.z * to.z < 0) nn = mul3(nn, -1.0f);
        } else {
            nn = nrm3(mk3(-d.x, -d.y, -d.z));
        }
        N[r * 3] = nn.x;
        N[r * 3 + 1] = nn.y;
        N[r * 3 + 2] = nn.z;
    }
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

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
    diffuse_sample(i, rr % n, rr / n, n, seed, patch, spx, spy, dl);
    v3f o = xform3(sc.T + (size_t)i * 16, spx, spy, 0.0f);
    v3f dir = frame_rotate(dl, fr.x, fr.y, fr.z, fr.w);
    if (slot == 1 && tw == 1) dir = mk3(-dir.x, -dir.y, -dir.z);
    float *w = out + g * 6;
    w[0] = o.x; w[1] = o.y; w[2] = o.z; w[3] = dir.x; w[4] = dir.y; w[5] = dir.z;
}

__global__ __launch_bounds__(256) void k_debug_trace(const DevScene *__restrict__ scp, uint32_t n,
                                                     const float *__restrict__ rays, const uint32_t *__restrict__ skip,
                                                     float tmin, int brute, uint32_t *__restrict__ hit,
                                                     float *__restrict__ tout, int *err) {
    __shared__ uint32_t stk_all[RT_STACK * 256];
    uint32_t *stk = stk_all + threadIdx.x;
    const DevScene &sc = *scp;
    uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    v3f o = mk3(rays[r * 6], rays[r * 6 + 1], rays[r * 6 + 2]);
    v3f d = mk3(rays[r * 6 + 3], rays[r * 6 + 4], rays[r * 6 + 5]);
    float t = 1e30f;
    uint32_t h = RT_MISS, slot;
    if (brute) {
        for (uint32_t k = 0; k < sc.nleaf; k++) test_leaf<false>(sc, k, o, d, tmin, skip[r], t, h, slot);
    } else {
        h = trace_closest<false>(sc, o, d, tmin, skip[r], 1e30f, t, slot, stk, 256, err);
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
                       uint32_t *hit, float *t, int *err, hipStream_t st) {
    if (n == 0) return 0;
    k_debug_trace<<<(n + 255) / 256, 256, 0, st>>>(sc, n, rays, skip, tmin, brute, hit, t, err);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error(std::string("kernel launch failed: ") + hipGetErrorString(e)); return -1; }
    return 0;
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
static inline uint32_t grid_for(const LaunchCfg &cfg, int ncu, uint32_t total, uint32_t wpb) {
    uint64_t g = (uint64_t)cfg.blocks_per_cu * (uint64_t)ncu;
    uint64_t need = (total + wpb - 1) / wpb;
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

int launch_direct(const DirectJob *jobs, uint32_t njobs, uint32_t total, uint32_t *counter, int *err,
                  const LaunchCfg &cfg, int ncu, hipStream_t st) {
    if (total == 0) return 0;
    k_zero_u32<<<1, 1, 0, st>>>(counter);
    uint32_t wpb = cfg.waves_per_block;
    uint32_t g = grid_for(cfg, ncu, total, wpb);
    switch (wpb) {
        case 1: k_direct<1><<<g, 64, 0, st>>>(jobs, njobs, total, counter, err); break;
        case 2: k_direct<2><<<g, 128, 0, st>>>(jobs, njobs, total, counter, err); break;
        default: k_direct<4><<<g, 256, 0, st>>>(jobs, njobs, total, counter, err); break;
    }
    LCHECK();
    return 0;
}

int launch_diffuse(const DiffuseJob *jobs, uint32_t njobs, uint32_t total, uint32_t *counter, int *err,
                   const LaunchCfg &cfg, int ncu, hipStream_t st) {
    if (total == 0) return 0;
    k_zero_u32<<<1, 1, 0, st>>>(counter);
    uint32_t wpb = cfg.waves_per_block;
    uint32_t g = grid_for(cfg, ncu, total, wpb);
    switch (wpb) {
        case 1: k_diffuse<1><<<g, 64, 0, st>>>(jobs, njobs, total, counter, err); break;
        case 2: k_diffuse<2><<<g, 128, 0, st>>>(jobs, njobs, total, counter, err); break;
        default: k_diffuse<4><<<g, 256, 0, st>>>(jobs, njobs, total, counter, err); break;
    }
    LCHECK();
    return 0;
}

int launch_gather(const GatherJob *jobs, uint32_t njobs, uint32_t total, uint32_t cmax, uint32_t *counter,
                  const LaunchCfg &cfg, int ncu, hipStream_t st) {
    if (total == 0) return 0;
    k_zero_u32<<<1, 1, 0, st>>>(counter);
    uint32_t g = grid_for(cfg, ncu, total, 4);
    if (cmax <= 1) k_gather_rays<1><<<g, 256, 0, st>>>(jobs, njobs, total, counter);
    else if (cmax <= 2) k_gather_rays<2><<<g, 256, 0, st>>>(jobs, njobs, total, counter);
    else if (cmax <= 4) k_gather_rays<4><<<g, 256, 0, st>>>(jobs, njobs, total, counter);
    else if (cmax <= 8) k_gather_rays<8><<<g, 256, 0, st>>>(jobs, njobs, total, counter);
    else if (cmax <= 64) k_gather_chan<1><<<g, 256, 0, st>>>(jobs, njobs, total, counter);
    else if (cmax <= 128) k_gather_chan<2><<<g, 256, 0, st>>>(jobs, njobs, total, counter);
    else if (cmax <= 192) k_gather_chan<3><<<g, 256, 0, st>>>(jobs, njobs, total, counter);
    else if (cmax <= 256) k_gather_chan<4><<<g, 256, 0, st>>>(jobs, njobs, total, counter);
    else {
        set_error("more than 256 channels in one gather");
        return -1;
    }
    LCHECK();
    return 0;
}

int launch_cast(const DevScene *sc, uint64_t n, const float *o, const float *d, float maxd, float *t, float *nrm,
                uint32_t *uuid, int *err, hipStream_t st) {
    if (n == 0) return 0;
    uint64_t need = (n + 255) / 256;
    uint32_t g = (uint32_t)(need < 65536ull * 16 ? need : 65536ull * 16);
    k_cast<<<g, 256, 0, st>>>(sc, n, o, d, maxd, t, nrm, uuid, err);
    LCHECK();
    return 0;
}

}  // namespace rt
