// rt_grid.hip -- occlusion grids for collimated sources.
//
// All shadow rays towards one collimated source are parallel, so a ray from origin o can only hit a
// primitive whose projection onto the plane perpendicular to the direction contains the projection of
// o.  For each direction this file bins every traced primitive's projected rectangle (padded) into a
// uniform grid over that plane; the direct kernel then tests only the primitives of the ray's cell, with
// the same intersection predicate as the BVH traversal, and stops at the first accepted hit.
//
// Exactness: a primitive the FP32 predicate accepts is hit at a point whose exact projection lies within
// the primitive's exact projection up to rounding of a few ulps of the scene's coordinates (t is bounded
// by the scene diameter, the basis is orthogonal to the direction to FP32 precision).  The rectangles are
// padded by 2e-4 of the primitive's extent plus 1e-5 of the largest scene coordinate, far above those
// errors, so the rectangle contains the ray origin's FP32 projection.  Cells are found with the same
// monotone formula for rectangles and rays, so the ray's cell lies in the rectangle's cell range.  The
// candidate set is therefore a superset of the primitives the predicate accepts, and every occlusion
// decision equals the BVH traversal's (order within a cell does not matter for an any-hit query).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rt_solve.h"

namespace rt {

namespace {

__device__ __forceinline__ uint32_t f2o_g(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
float o2f_g(uint32_t u) {
    uint32_t v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

// padded projected rectangle of the primitive in leaf slot k, and its depth key: the largest depth of its
// vertices along the source direction w, plus the same padding (an upper bound of any of its points' depth)
__device__ __forceinline__ void proj_rect(const float4 *__restrict__ rec, uint32_t k, const SunGrid &G, float smax,
                                          float &u0, float &u1, float &v0, float &v1, float &key) {
    const float4 r0 = rec[(size_t)k * 4], r1 = rec[(size_t)k * 4 + 1], r2 = rec[(size_t)k * 4 + 2];
    const bool patch = (__float_as_uint(r0.w) & RT_PATCH_BIT) != 0;
    float cx[4], cy[4], cz[4];
    cx[0] = r0.x; cy[0] = r0.y; cz[0] = r0.z;
    cx[1] = r0.x + r1.x; cy[1] = r0.y + r1.y; cz[1] = r0.z + r1.z;
    cx[2] = r0.x + r2.x; cy[2] = r0.y + r2.y; cz[2] = r0.z + r2.z;
    if (patch) { cx[3] = cx[1] + r2.x; cy[3] = cy[1] + r2.y; cz[3] = cz[1] + r2.z; }
    else { cx[3] = cx[0]; cy[3] = cy[0]; cz[3] = cz[0]; }
    u0 = v0 = INFINITY;
    u1 = v1 = -INFINITY;
    float kmax = -INFINITY;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float pu = cx[c] * G.u[0] + cy[c] * G.u[1] + cz[c] * G.u[2];
        float pv = cx[c] * G.v[0] + cy[c] * G.v[1] + cz[c] * G.v[2];
        float pw = cx[c] * G.w[0] + cy[c] * G.w[1] + cz[c] * G.w[2];
        u0 = fminf(u0, pu); u1 = fmaxf(u1, pu);
        v0 = fminf(v0, pv); v1 = fmaxf(v1, pv);
        kmax = fmaxf(kmax, pw);
        lo[0] = fminf(lo[0], cx[c]); hi[0] = fmaxf(hi[0], cx[c]);
        lo[1] = fminf(lo[1], cy[c]); hi[1] = fmaxf(hi[1], cy[c]);
        lo[2] = fminf(lo[2], cz[c]); hi[2] = fmaxf(hi[2], cz[c]);
    }
    const float ext = fmaxf(fmaxf(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    const float pad = ext * 2e-4f + smax * 1e-5f + 1e-30f;
    u0 -= pad; u1 += pad; v0 -= pad; v1 += pad;
    key = kmax + pad;
}

__device__ __forceinline__ int cell_of(float q, float g0, float inv_cs, uint32_t n) {
    float f = floorf((q - g0) * inv_cs);
    int c = f < 0.f ? 0 : (f > (float)(n - 1) ? (int)(n - 1) : (int)f);
    return c;
}

// voxels occlude no radiation ray (the radiation traversal tests patches, tiles and triangles only), so
// the grids leave them out
__device__ __forceinline__ bool leaf_is_voxel(const float4 *__restrict__ rec, uint32_t k) {
    return (__float_as_uint(rec[(size_t)k * 4].w) & RT_VOXEL_BIT) != 0;
}

__global__ void k_grid_bounds(uint32_t nleaf, const float4 *__restrict__ rec, float smax,
                              const SunGrid *__restrict__ grids, uint32_t *__restrict__ bounds /*[n][4]*/) {
    const SunGrid G = grids[blockIdx.y];
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    float u0 = INFINITY, u1 = -INFINITY, v0 = INFINITY, v1 = -INFINITY, key;
    if (k < nleaf && !leaf_is_voxel(rec, k)) proj_rect(rec, k, G, smax, u0, u1, v0, v1, key);
    // block reduction, then one atomic per block
    __shared__ float s[4][256];
    s[0][threadIdx.x] = u0; s[1][threadIdx.x] = -u1; s[2][threadIdx.x] = v0; s[3][threadIdx.x] = -v1;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 4; q++) s[q][threadIdx.x] = fminf(s[q][threadIdx.x], s[q][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        uint32_t *b = bounds + (size_t)blockIdx.y * 4;
        atomicMin(&b[0], f2o_g(s[0][0]));
        atomicMin(&b[1], f2o_g(s[1][0]));
        atomicMin(&b[2], f2o_g(s[2][0]));
        atomicMin(&b[3], f2o_g(s[3][0]));
    }
}

#ifndef RT_FILL_LDS
#define RT_FILL_LDS 8192  // cells of a workgroup's union rectangle that the count and fill kernels keep in LDS
#endif

// Cell counts from a 2D difference array: each primitive adds +1 at (x0, y0), -1 at (x1 + 1, y0) and
// (x0, y1 + 1), +1 at (x1 + 1, y1 + 1) (corners outside the grid dropped), four atomics instead of one per
// covered cell; a prefix sum along the rows and then along the columns turns the differences into the
// number of rectangles covering each cell (uint32 arithmetic, exact modulo 2^32 and so exact).
__global__ __launch_bounds__(256) void k_grid_count_diff(uint32_t nleaf, const float4 *__restrict__ rec, float smax,
                                                         const SunGrid *__restrict__ grids,
                                                         const uint32_t *__restrict__ cell_base,
                                                         uint32_t *__restrict__ diff,
                                                         unsigned long long *__restrict__ entries /*[n]*/) {
    // as in k_grid_fill, a workgroup whose rectangles' union spans at most RT_FILL_LDS cells adds its
    // corners in LDS first and then each nonzero cell once to the global array
    __shared__ uint32_t lc[RT_FILL_LDS];
    __shared__ int bb[4];
    __shared__ unsigned long long nent;
    const SunGrid G = grids[blockIdx.y];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = k < nleaf && !leaf_is_voxel(rec, k);
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1;
    if (valid) {
        float u0, u1, v0, v1, key;
        proj_rect(rec, k, G, smax, u0, u1, v0, v1, key);
        x0 = cell_of(u0, G.g0u, G.inv_cs, G.nx); x1 = cell_of(u1, G.g0u, G.inv_cs, G.nx);
        y0 = cell_of(v0, G.g0v, G.inv_cs, G.ny); y1 = cell_of(v1, G.g0v, G.inv_cs, G.ny);
    }
    if (threadIdx.x < 4) bb[threadIdx.x] = 0x7FFFFFFF;
    if (threadIdx.x == 0) nent = 0ull;
    __syncthreads();
    if (valid) {
        atomicMin(&bb[0], x0); atomicMin(&bb[1], -x1);
        atomicMin(&bb[2], y0); atomicMin(&bb[3], -y1);
        atomicAdd(&nent, (unsigned long long)(x1 - x0 + 1) * (unsigned long long)(y1 - y0 + 1));
    }
    __syncthreads();
    if (threadIdx.x == 0 && nent) atomicAdd(&entries[blockIdx.y], nent);
    // bounds of the corner cells: the rectangles' cells plus one row and column past them
    const int bx0 = bb[0], bx1 = -bb[1] + 1, by0 = bb[2], by1 = -bb[3] + 1;
    const bool any = bx0 <= bx1 - 1;
    const uint32_t bw = any ? (uint32_t)(bx1 - bx0 + 1) : 0u, bh = any ? (uint32_t)(by1 - by0 + 1) : 0u;
    const bool dense = any && (uint64_t)bw * bh <= (uint64_t)RT_FILL_LDS;
    uint32_t *d = diff + cell_base[blockIdx.y];
    const bool xr = (uint32_t)(x1 + 1) < G.nx, yr = (uint32_t)(y1 + 1) < G.ny;
    if (dense) {
        const uint32_t area = bw * bh;
        for (uint32_t q = threadIdx.x; q < area; q += blockDim.x) lc[q] = 0u;
        __syncthreads();
        if (valid) {
            const uint32_t r0 = (uint32_t)(y0 - by0) * bw, r1 = (uint32_t)(y1 + 1 - by0) * bw;
            const uint32_t c0 = (uint32_t)(x0 - bx0), c1 = (uint32_t)(x1 + 1 - bx0);
            atomicAdd(&lc[r0 + c0], 1u);
            if (xr) atomicAdd(&lc[r0 + c1], 0xFFFFFFFFu);
            if (yr) atomicAdd(&lc[r1 + c0], 0xFFFFFFFFu);
            if (xr && yr) atomicAdd(&lc[r1 + c1], 1u);
        }
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < area; q += blockDim.x) {
            const uint32_t v = lc[q];
            if (v == 0u) continue;
            atomicAdd(&d[(size_t)(by0 + (int)(q / bw)) * G.nx + (uint32_t)bx0 + q % bw], v);
        }
    } else if (valid) {
        atomicAdd(&d[(size_t)y0 * G.nx + x0], 1u);
        if (xr) atomicAdd(&d[(size_t)y0 * G.nx + x1 + 1], 0xFFFFFFFFu);
        if (yr) atomicAdd(&d[(size_t)(y1 + 1) * G.nx + x0], 0xFFFFFFFFu);
        if (xr && yr) atomicAdd(&d[(size_t)(y1 + 1) * G.nx + x1 + 1], 1u);
    }
}

// inclusive prefix sums along each row (one wave per row: 64 cells per step with a running carry)
__global__ void k_grid_scan_rows(const SunGrid *__restrict__ grids, const uint32_t *__restrict__ cell_base,
                                 uint32_t *__restrict__ diff) {
    const SunGrid &G = grids[blockIdx.y];
    const uint32_t nx = G.nx, ny = G.ny;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= ny) return;
    uint32_t *d = diff + cell_base[blockIdx.y] + (size_t)row * nx;
    uint32_t carry = 0;
    for (uint32_t x0 = 0; x0 < nx; x0 += 64) {
        const uint32_t x = x0 + lane;
        uint32_t v = x < nx ? d[x] : 0u;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = (uint32_t)__shfl_up((int)v, off, 64);
            if (lane >= (uint32_t)off) v += t;
        }
        if (x < nx) d[x] = v + carry;
        carry += (uint32_t)__shfl((int)v, 63, 64);
    }
}

// inclusive prefix sums down each column (one thread per column; neighbouring threads read neighbouring
// cells of a row)
__global__ void k_grid_scan_cols(const SunGrid *__restrict__ grids, const uint32_t *__restrict__ cell_base,
                                 uint32_t *__restrict__ diff) {
    const SunGrid &G = grids[blockIdx.y];
    const uint32_t nx = G.nx, ny = G.ny;
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    uint32_t *d = diff + cell_base[blockIdx.y] + x;
    uint32_t run = 0;
    for (uint32_t y = 0; y < ny; y++) {
        run += d[(size_t)y * nx];
        d[(size_t)y * nx] = run;
    }
}

// Each cell's list in descending key order (ties: ascending leaf slot), so the list order is deterministic
// and a scan can stop at the first key below the ray origin's depth.  One thread per entry counts the entries of its cell that precede it and
// moves the entry to that rank: the work is the sum of squared list lengths, spread over all entries.
#define RT_RANK_MARGIN 256
__global__ __launch_bounds__(256) void k_grid_rank(const SunGrid *__restrict__ grids,
                                                   const uint32_t *__restrict__ cell_base,
                                                   const unsigned long long *__restrict__ item_base,
                                                   const unsigned long long *__restrict__ item_count,
                                                   const uint32_t *__restrict__ start,
                                                   const uint2 *__restrict__ tmp, const uint32_t *__restrict__ tcell,
                                                   uint2 *__restrict__ ents) {
    // The unsorted lists lie cell after cell in the scratch array, so the lists of a workgroup's entries
    // mostly fall in a window of its 256 entries and RT_RANK_MARGIN on either side: that window is read
    // into LDS once, and an entry whose cell's list fits it scans the list there instead of in memory.
    __shared__ uint2 win[256 + 2 * RT_RANK_MARGIN];
    const unsigned long long ni = item_count[blockIdx.y];
    const unsigned long long ib = item_base[blockIdx.y];
    const long long w0 = (long long)blockIdx.x * 256 - RT_RANK_MARGIN;
    for (uint32_t t = threadIdx.x; t < 256 + 2 * RT_RANK_MARGIN; t += 256) {
        const long long y = w0 + (long long)t;
        win[t] = (y >= 0 && (unsigned long long)y < ni) ? tmp[ib + (unsigned long long)y] : make_uint2(RT_MISS, 0u);
    }
    __syncthreads();
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ni) return;
    const uint32_t sb = (uint32_t)ib;
    const size_t e0 = ib + x;
    const uint2 me = win[(long long)x - w0];  // (entry, key)
    const uint32_t cl = tcell[e0];
    // a slot the fill left empty (none when the fill matches the count, as it does) has no cell: it is not
    // placed, and its NaN key never counts as preceding anything; the final list keeps its trailing slots
    // as "no primitive" with a NaN key, which ends a scan
    if (cl >= grids[blockIdx.y].nx * grids[blockIdx.y].ny) return;
    const size_t cc = cell_base[blockIdx.y] + cl;
    const uint32_t b = start[cc] - sb, e = start[cc + 1] - sb;
    const float kx = __uint_as_float(me.y);
    const uint32_t ix = me.x;
    uint32_t rank = 0;
    const bool inwin = (long long)b >= w0 && (long long)e <= w0 + 256 + 2 * RT_RANK_MARGIN;
    for (uint32_t y = b; y < e; y++) {
        const uint2 oy = inwin ? win[(long long)y - w0] : tmp[ib + y];
        const float ky = __uint_as_float(oy.y);
        const uint32_t iy = oy.x;
        rank += (ky > kx || (ky == kx && ((iy & RT_GRID_SLOT) < (ix & RT_GRID_SLOT) ||
                                          (iy == ix && y < (uint32_t)x)))) ? 1u : 0u;
    }
    if (b + rank < e) ents[ib + b + rank] = make_uint2(ix, __float_as_uint(kx));
}

// start[] is one exclusive scan over the cells of all directions in uint32 arithmetic, so it holds the
// true offsets modulo 2^32; a direction's own list starts at the 64-bit offset item_base[j] and its
// offsets count from sbase = item_base[j] mod 2^32 (each direction has fewer than 2^32 entries).
// The fill takes one list slot per entry from its cell's count (atomicSub, counting down from the cell's
// total).  A workgroup's 256 leaves are consecutive in leaf order, so their projected rectangles overlap:
// when the union of the workgroup's rectangles spans at most RT_FILL_LDS cells, the workgroup first counts
// its entries per cell in LDS, takes each cell's block of slots with one global atomic, and hands the slots
// out from LDS; otherwise every entry takes its slot with a global atomic.
__global__ __launch_bounds__(256) void k_grid_fill(uint32_t nleaf, const float4 *__restrict__ rec, float smax,
                                                   const SunGrid *__restrict__ grids,
                                                   const uint32_t *__restrict__ cell_base,
                                                   const unsigned long long *__restrict__ item_base,
                                                   const unsigned long long *__restrict__ item_count,
                                                   const uint32_t *__restrict__ start, uint32_t *__restrict__ counts,
                                                   uint2 *__restrict__ tmp, uint32_t *__restrict__ tcell) {
    __shared__ uint32_t lc[RT_FILL_LDS];
    __shared__ int bb[4];  // union of the workgroup's cell ranges: x0, -x1, y0, -y1 (minima)
    const SunGrid G = grids[blockIdx.y];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = k < nleaf && !leaf_is_voxel(rec, k);
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1;
    float key = 0.f;
    if (valid) {
        float u0, u1, v0, v1;
        proj_rect(rec, k, G, smax, u0, u1, v0, v1, key);
        x0 = cell_of(u0, G.g0u, G.inv_cs, G.nx); x1 = cell_of(u1, G.g0u, G.inv_cs, G.nx);
        y0 = cell_of(v0, G.g0v, G.inv_cs, G.ny); y1 = cell_of(v1, G.g0v, G.inv_cs, G.ny);
    }
    if (threadIdx.x < 4) bb[threadIdx.x] = 0x7FFFFFFF;
    __syncthreads();
    if (valid) {
        atomicMin(&bb[0], x0); atomicMin(&bb[1], -x1);
        atomicMin(&bb[2], y0); atomicMin(&bb[3], -y1);
    }
    __syncthreads();
    const int bx0 = bb[0], bx1 = -bb[1], by0 = bb[2], by1 = -bb[3];
    const size_t base = cell_base[blockIdx.y];
    const unsigned long long ib = item_base[blockIdx.y], ni = item_count[blockIdx.y];
    const uint32_t sbase = (uint32_t)ib;
    uint32_t item = 0;
    if (valid) {
        const uint32_t tag = __float_as_uint(rec[(size_t)k * 4].w);
        item = k | ((tag & (RT_PATCH_BIT | RT_MASK_BIT)) ? RT_GRID_REC3 : 0u);
    }
    const bool any = bx0 <= bx1;  // the workgroup has a traced leaf
    const uint32_t bw = any ? (uint32_t)(bx1 - bx0 + 1) : 0u, bh = any ? (uint32_t)(by1 - by0 + 1) : 0u;
    const bool dense = any && (uint64_t)bw * bh <= (uint64_t)RT_FILL_LDS;
    // the ranges equal the counting kernel's (same code, same FP32 operations); the guards below only keep
    // a mismatch from writing outside the direction's list
    if (dense) {
        const uint32_t area = bw * bh;
        for (uint32_t q = threadIdx.x; q < area; q += blockDim.x) lc[q] = 0u;
        __syncthreads();
        for (int y = y0; y <= y1; y++)
            for (int x = x0; x <= x1; x++) atomicAdd(&lc[(uint32_t)(y - by0) * bw + (uint32_t)(x - bx0)], 1u);
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < area; q += blockDim.x) {
            const uint32_t cnt = lc[q];
            if (cnt == 0u) continue;
            const size_t c = base + (size_t)(by0 + (int)(q / bw)) * G.nx + (uint32_t)bx0 + q % bw;
            lc[q] = atomicSub(&counts[c], cnt);  // the slots left + 1 - cnt .. left are this workgroup's
        }
        __syncthreads();
    }
    for (int y = y0; y <= y1; y++)
        for (int x = x0; x <= x1; x++) {
            const size_t c = base + (size_t)y * G.nx + x;
            const uint32_t left = dense ? atomicSub(&lc[(uint32_t)(y - by0) * bw + (uint32_t)(x - bx0)], 1u)
                                        : atomicSub(&counts[c], 1u);
            if (left == 0u) continue;
            const unsigned long long at = (unsigned long long)(uint32_t)(start[c] - sbase) + left - 1u;
            if (at < ni) {
                // the entry and its key in one 8-byte store (the list slots are scattered: each store is a
                // partial-line write), the cell in a second
                tmp[ib + at] = make_uint2(item, __float_as_uint(key));
                tcell[ib + at] = (uint32_t)(c - base);
            }
        }
}

inline uint32_t nblk(size_t n) { return (uint32_t)((n + 255) / 256); }

#define GK(x)                                                                     \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #x); \
            return -1;                                                            \
        }                                                                         \
    } while (0)

}  // namespace

int build_sun_grids(GridSet &gs, const Scene &S, const float *dirs, uint32_t n, hipStream_t st) {
    gs.n = 0;
    const uint32_t nleaf = S.h.nleaf;
    if (n == 0 || nleaf == 0 || S.h.nbbox != 0) return 1;
    // host: basis per direction (double precision, then rounded)
    std::vector<SunGrid> hg(n);
    for (uint32_t j = 0; j < n; j++) {
        double d[3] = {dirs[j * 3], dirs[j * 3 + 1], dirs[j * 3 + 2]};
        double m = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!(m > 0.0)) return 1;
        for (double &x : d) x /= m;
        int a = 0;
        for (int q = 1; q < 3; q++)
            if (std::fabs(d[q]) < std::fabs(d[a])) a = q;
        double e[3] = {0, 0, 0};
        e[a] = 1.0;
        double u[3] = {d[1] * e[2] - d[2] * e[1], d[2] * e[0] - d[0] * e[2], d[0] * e[1] - d[1] * e[0]};
        double um = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (double &x : u) x /= um;
        double v[3] = {d[1] * u[2] - d[2] * u[1], d[2] * u[0] - d[0] * u[2], d[0] * u[1] - d[1] * u[0]};
        SunGrid &G = hg[j];
        for (int q = 0; q < 3; q++) { G.u[q] = (float)u[q]; G.v[q] = (float)v[q]; G.w[q] = (float)d[q]; }
        G.g0u = G.g0v = 0.f;
        G.inv_cs = 1.f;
        G.nx = G.ny = 1;
        G.sbase = 0;
        G.sorted = 0;
        G.start = nullptr;
        G.ents = nullptr;
    }
    // 1. projected extents
    if (gs.small.reserve(n * sizeof(SunGrid) + n * 16 + n * 4 + n * 16 + 2048)) return -1;
    char *p = (char *)gs.small.base;
    SunGrid *dgrids = (SunGrid *)p;
    uint32_t *bounds = (uint32_t *)(p + ((n * sizeof(SunGrid) + 255) & ~(size_t)255));
    uint32_t *dbase = bounds + 4 * n;
    unsigned long long *dentries = (unsigned long long *)(dbase + ((n + 63) & ~(size_t)63));  // [n]
    unsigned long long *dibase = dentries + n;                                                // [n]
    GK(hipMemcpyAsync(dgrids, hg.data(), n * sizeof(SunGrid), hipMemcpyHostToDevice, st));
    GK(hipMemsetAsync(bounds, 0xFF, n * 16, st));
    dim3 grid(nblk(nleaf), n);
    k_grid_bounds<<<grid, 256, 0, st>>>(nleaf, S.h.rec, S.h.smax, dgrids, bounds);
    GK(hipGetLastError());
    std::vector<uint32_t> hb((size_t)n * 4);
    GK(hipMemcpyAsync(hb.data(), bounds, n * 16, hipMemcpyDeviceToHost, st));
    GK(hipStreamSynchronize(st));
    // 2. grid shapes: `density` cells per primitive (default 2) over the projected extent
    std::vector<uint32_t> base(n + 1, 0);
    size_t total_cells = 0;
    const double density = gs.density > 0.f ? (double)gs.density : 2.0;
    const double target = std::min(std::max(density * nleaf, 1024.0), (double)(1u << 24));
    for (uint32_t j = 0; j < n; j++) {
        SunGrid &G = hg[j];
        float u0 = o2f_g(hb[j * 4]), u1 = -o2f_g(hb[j * 4 + 1]), v0 = o2f_g(hb[j * 4 + 2]), v1 = -o2f_g(hb[j * 4 + 3]);
        if (!(u1 >= u0) || !(v1 >= v0) || !std::isfinite(u0) || !std::isfinite(v1)) return 1;
        double W = (double)u1 - u0, H = (double)v1 - v0;
        double cs = std::sqrt(W * H / target);
        if (!(cs > 0.0) || !std::isfinite(cs)) cs = std::max(std::max(W, H), 1e-6) / std::sqrt(target);
        uint32_t nx = (uint32_t)std::min(16384.0, std::max(1.0, std::ceil(W / cs) + 1.0));
        uint32_t ny = (uint32_t)std::min(16384.0, std::max(1.0, std::ceil(H / cs) + 1.0));
        G.g0u = u0;
        G.g0v = v0;
        G.inv_cs = (float)(1.0 / cs);
        G.nx = nx;
        G.ny = ny;
        base[j] = (uint32_t)total_cells;
        total_cells += (size_t)nx * ny;
        if (total_cells + 1 > (1ull << 30)) return 1;  // scans index with int
    }
    // 3. counts -> starts -> items
    size_t scan_bytes = 0;
    GK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                        (int)(total_cells + 1), st));
    const size_t cell_bytes = 2 * (total_cells + 1) * 4 + scan_bytes + 1024;
    if (gs.cap && cell_bytes > gs.cap) return 1;
    if (gs.cells.reserve(cell_bytes)) return -1;
    uint32_t *counts = (uint32_t *)gs.cells.base;
    uint32_t *starts = counts + ((total_cells + 1 + 63) & ~(size_t)63);
    void *tmp = starts + ((total_cells + 1 + 63) & ~(size_t)63);
    GK(hipMemcpyAsync(dbase, base.data(), n * 4, hipMemcpyHostToDevice, st));
    GK(hipMemcpyAsync(dgrids, hg.data(), n * sizeof(SunGrid), hipMemcpyHostToDevice, st));
    GK(hipMemsetAsync(counts, 0, (total_cells + 1) * 4, st));
    GK(hipMemsetAsync(dentries, 0, (size_t)n * 8, st));
    {
        uint32_t maxnx = 1, maxny = 1;
        for (uint32_t j = 0; j < n; j++) { maxnx = std::max(maxnx, hg[j].nx); maxny = std::max(maxny, hg[j].ny); }
        k_grid_count_diff<<<grid, 256, 0, st>>>(nleaf, S.h.rec, S.h.smax, dgrids, dbase, counts, dentries);
        GK(hipGetLastError());
        k_grid_scan_rows<<<dim3((maxny + 3) / 4, n), 256, 0, st>>>(dgrids, dbase, counts);
        GK(hipGetLastError());
        k_grid_scan_cols<<<dim3((maxnx + 255) / 256, n), 256, 0, st>>>(dgrids, dbase, counts);
        GK(hipGetLastError());
    }
    std::vector<unsigned long long> ent(n), ibase(n);
    GK(hipMemcpyAsync(ent.data(), dentries, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    GK(hipStreamSynchronize(st));
    unsigned long long total = 0;
    for (uint32_t j = 0; j < n; j++) {
        if (ent[j] >= (1ull << 32) - 1) return 1;  // a direction's offsets are uint32: traverse the BVH
        ibase[j] = total;
        total += ent[j];
    }
    if (gs.cap && cell_bytes + total * 20 + 1024 > gs.cap) return 1;  // lists (8 B) + sort scratch (12 B)
    GK(hipcub::DeviceScan::ExclusiveSum(tmp, scan_bytes, counts, starts, (int)(total_cells + 1), st));
    if (gs.items.reserve((size_t)total * 8 + 256)) return -1;
    uint2 *ents = (uint2 *)gs.items.base;
    // the fill writes unsorted lists of (entry, key, cell) to scratch; k_grid_rank moves each entry to its
    // place.  Slots nothing is written to read as "no primitive" with a NaN key and no cell.
    const size_t tb = ((size_t)total * 4 + 255) & ~(size_t)255;
    if (gs.sortbuf.reserve(3 * tb + 256)) return -1;
    uint2 *unsorted = (uint2 *)gs.sortbuf.base;
    uint32_t *ucell = (uint32_t *)((char *)gs.sortbuf.base + 2 * tb);
    GK(hipMemsetAsync(gs.sortbuf.base, 0xFF, 3 * tb, st));
    GK(hipMemsetAsync(ents, 0xFF, (size_t)total * 8, st));
    GK(hipMemcpyAsync(dibase, ibase.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    k_grid_fill<<<grid, 256, 0, st>>>(nleaf, S.h.rec, S.h.smax, dgrids, dbase, dibase, dentries, starts, counts,
                                      unsorted, ucell);
    GK(hipGetLastError());
    {
        unsigned long long maxent = 0;
        for (uint32_t j = 0; j < n; j++) maxent = std::max(maxent, ent[j]);
        if (maxent) {
            k_grid_rank<<<dim3((uint32_t)((maxent + 255) / 256), n), 256, 0, st>>>(dgrids, dbase, dibase, dentries,
                                                                                   starts, unsorted, ucell, ents);
            GK(hipGetLastError());
        }
    }
    const bool sorted = true;
    for (uint32_t j = 0; j < n; j++) {
        hg[j].start = starts + base[j];
        hg[j].ents = ents + ibase[j];
        hg[j].sbase = (uint32_t)ibase[j];
        hg[j].sorted = sorted ? 1u : 0u;
    }
    GK(hipMemcpyAsync(dgrids, hg.data(), n * sizeof(SunGrid), hipMemcpyHostToDevice, st));
    GK(hipStreamSynchronize(st));
    gs.dev = dgrids;
    gs.n = n;
    gs.entries = total;
    gs.cells_total = total_cells;
    return 0;
}

}  // namespace rt
