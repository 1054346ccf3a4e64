// rt_api.hip -- extern "C" entry points declared in include/helios_rt.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rt_ctx.h"

namespace rt {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
}  // namespace rt

using namespace rt;

void backend_release(BackendState &be);

// ------------------------------------------------------------------------------------------------
// Scratch of the allocation-free LiDAR cast (rt_lidar_cast_rays_noalloc): one array per device, in the
// library's own device data, holding the work counter and the traversal error flag in its first 256 bytes,
// then the castRays normals [n][3] and UUIDs [n] of up to RT_NOALLOC_MAX_RAYS rays.  Host code never uses
// the array's name as an address: rt_create looks up each device's copy with hipGetSymbolAddress once and
// the kernels get that address.
// ------------------------------------------------------------------------------------------------
namespace {
constexpr size_t NOALLOC_HEAD = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_HEAD + (size_t)RT_NOALLOC_MAX_RAYS * 16;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;                       // guards the table below
void *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {}; // device address of each device's scratch
std::mutex g_noalloc_run[NOALLOC_MAX_DEVICES]; // one call at a time per device in this process

// Looks up this device's scratch (the device must be current).  Called by rt_create, once per device.
int noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return 0;  // such a device gets no allocation-free path
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return 0;
    void *p = nullptr;
    CKR(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) { set_error("hipGetSymbolAddress returned no address for the scratch array"); return -1; }
    g_noalloc_ptr[device] = p;
    return 0;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
}  // namespace

extern "C" {

const char *rt_last_error(void) { return g_err.c_str(); }

int rt_create(int device, rt_ctx **out) {
    NN(out, "out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_error("no HIP device available (this backend has no CPU path)");
        return -1;
    }
    if (device < 0 || device >= n) { set_error("device index out of range"); return -1; }
    CKR(hipSetDevice(device));
    hipDeviceProp_t prop;
    CKR(hipGetDeviceProperties(&prop, device));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0) {
        set_error("device " + std::to_string(device) + " is " + arch + "; this library is built for gfx950 only");
        return -1;
    }
    if (noalloc_scratch_init(device)) return -1;
    rt_ctx *c = new rt_ctx();
    c->device = device;
    c->name = std::string("hip-gfx950 (") + prop.name + ")";
    c->sx.ncu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->sx.stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("hipStreamCreate failed");
        return -1;
    }
    *out = c;
    return 0;
}

int rt_destroy(rt_ctx *ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->sx.stream);
    for (auto &e : ctx->sx.ev) { if (e.a) (void)hipEventDestroy(e.a); if (e.b) (void)hipEventDestroy(e.b); }
    ctx->sx.ws.release();
    ctx->sx.grids.release();
    backend_release(ctx->be);
    if (ctx->sx.stats) (void)hipFree(ctx->sx.stats);
    (void)hipStreamDestroy(ctx->sx.stream);
    delete ctx;
    return 0;
}

const char *rt_backend_name(const rt_ctx *ctx) { return ctx ? ctx->name.c_str() : "none"; }

int rt_query_memory(rt_ctx *ctx, size_t *free_bytes, size_t *total_bytes) {
    NN(ctx, "ctx");
    CKR(hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    CKR(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return 0;
}

int rt_scene_create(rt_ctx *ctx, uint32_t nprims, const float *transforms, const uint32_t *types,
                    const uint32_t *uuids, const uint32_t *twosided, int periodic_x, int periodic_y, rt_scene **out) {
    NN(ctx, "ctx");
    NN(out, "out");
    if (nprims) { NN(transforms, "transforms"); NN(types, "types"); NN(twosided, "twosided"); }
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    rt_scene *s = new rt_scene();
    if (scene_build(s->s, nprims, transforms, types, uuids, twosided, periodic_x ? 1 : 0, periodic_y ? 1 : 0,
                    ctx->sx.stream)) {
        delete s;
        return -1;
    }
    *out = s;
    return 0;
}

int rt_scene_destroy(rt_ctx *ctx, rt_scene *scene) {
    if (!scene) return 0;
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->sx.stream);
    }
    scene_free(scene->s);
    delete scene;
    return 0;
}

int rt_scene_set_masks(rt_ctx *ctx, rt_scene *scene, uint32_t nmasks, const uint8_t *texels, const uint32_t *sizes,
                       const int32_t *mask_ids, const float *uv, const int32_t *uv_ids) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    return scene_set_masks(scene->s, nmasks, texels, sizes, mask_ids, uv, uv_ids, ctx->sx.stream);
}

int rt_scene_get_info(const rt_scene *scene, rt_scene_info *info) {
    NN(scene, "scene");
    NN(info, "info");
    info->nprims = scene->s.h.P;
    info->nleaves = scene->s.h.nleaf;
    info->depth = scene->s.h.depth;
    info->nbbox = scene->s.h.nbbox;
    for (int k = 0; k < 4; k++) info->domain[k] = scene->s.h.dom[k];
    info->device_bytes = scene->s.bytes;
    info->builder = (uint32_t)scene->s.builder;
    info->depth4 = scene->s.h.node4 ? scene->s.h.depth4 : 0;
    return 0;
}

int rt_cast_rays_soa_device(rt_ctx *ctx, const rt_scene *scene, const float *o, const float *d, size_t n,
                            float max_dist, float *t, float *nrm, uint32_t *uuid) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    if (n == 0) return 0;
    NN(o, "origins"); NN(d, "dirs"); NN(t, "out_t"); NN(nrm, "out_normal"); NN(uuid, "out_uuid");
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    hipStream_t st = ctx->sx.stream;
    if (ctx->sx.ws.reserve(64)) return -1;
    int *err = (int *)ctx->sx.ws.base;
    unsigned long long *counter = (unsigned long long *)((char *)ctx->sx.ws.base + 8);
    CKR(hipMemsetAsync(err, 0, sizeof(int), st));
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->sx.timing) { CKR(hipEventCreate(&a)); CKR(hipEventCreate(&b)); CKR(hipEventRecord(a, st)); }
    if (launch_cast(scene->s.d, scene->s.h.depth, n, o, d, max_dist, t, nrm, uuid, counter, err, ctx->sx.stat_ptr(2),
                    ctx->sx.ncu, st, ctx->sx.watertight))
        return -1;
    if (b) CKR(hipEventRecord(b, st));
    int herr = 0;
    CKR(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
    CKR(hipStreamSynchronize(st));
    if (a) {
        float ms = 0.f;
        CKR(hipEventElapsedTime(&ms, a, b));
        ctx->sx.timing_info = rt_timing{};
        ctx->sx.timing_info.direct_ms = ms;  // castRays time reported in the first slot
        ctx->sx.timing_info.direct_rays = n;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    if (herr) { set_error("traversal error code " + std::to_string(herr) + " (1 = stack overflow, 2 = malformed tree)"); return -1; }
    return 0;
}

int rt_cast_rays_soa(rt_ctx *ctx, const rt_scene *scene, const float *o, const float *d, size_t n, float max_dist,
                     float *t, float *nrm, uint32_t *uuid) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    if (n == 0) return 0;
    NN(o, "origins"); NN(d, "dirs"); NN(t, "out_t"); NN(nrm, "out_normal"); NN(uuid, "out_uuid");
    CKR(hipSetDevice(ctx->device));
    float *dbuf = nullptr;
    size_t per = 3 + 3 + 1 + 3 + 1;  // o, d, t, n, uuid (4 bytes each)
    CKR(hipMalloc(&dbuf, n * per * 4));
    float *dO = dbuf, *dD = dbuf + n * 3, *dT = dbuf + n * 6, *dN = dbuf + n * 7;
    uint32_t *dU = reinterpret_cast<uint32_t *>(dbuf + n * 10);
    int rc = 0;
    do {
        if (hipMemcpy(dO, o, n * 12, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dD, d, n * 12, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("castRays upload failed");
            rc = -1;
            break;
        }
        if (rt_cast_rays_soa_device(ctx, scene, dO, dD, n, max_dist, dT, dN, dU)) { rc = -1; break; }
        // misses leave the caller's out_t / out_normal untouched (CD castRaysSoA contract)
        std::vector<float> ht(n), hn(n * 3);
        if (hipMemcpy(ht.data(), dT, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hn.data(), dN, n * 12, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(uuid, dU, n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("castRays download failed");
            rc = -1;
            break;
        }
        for (size_t r = 0; r < n; r++) {
            if (uuid[r] == RT_MISS_UUID) continue;
            t[r] = ht[r];
            nrm[r * 3] = hn[r * 3]; nrm[r * 3 + 1] = hn[r * 3 + 1]; nrm[r * 3 + 2] = hn[r * 3 + 2];
        }
    } while (0);
    (void)hipFree(dbuf);
    return rc;
}

int rt_cast_rays_soa_packets(rt_ctx *ctx, const rt_scene *scene, const float *o, const float *d, size_t n,
                             size_t packet_size, float max_dist, float *t, float *nrm, uint32_t *uuid) {
    // Results equal the per-ray path (CollisionDetection.h:404-408).  The reference's packet traversal is
    // its CPU path; with a GPU-resident scene it defers to the per-ray GPU kernel, whose massively parallel
    // traversal the packet sharing does not map onto (CD/src/CollisionDetection_RayTracing.cpp:714-720).
    // The same holds here: a wave takes 64 consecutive rays, so the sub-rays of a pulse traverse together
    // (the nodes they share come from the same cache lines) whatever packet_size is.
    (void)packet_size;
    return rt_cast_rays_soa(ctx, scene, o, d, n, max_dist, t, nrm, uuid);
}

namespace {
// LiDARcloud::castRaysUnified's per-ray conversion (plugins/lidar/src/LiDAR.cpp:585-597)
__global__ void k_lidar_convert(size_t n, const float *__restrict__ dirs, const float *__restrict__ nrm,
                                const uint32_t *__restrict__ uuid, float miss_t, float *__restrict__ hit_t,
                                float *__restrict__ fnorm, int32_t *__restrict__ id) {
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t u = uuid[r];
    if (u != RT_MISS_UUID) {
        id[r] = (int32_t)u;
        fnorm[r] = dirs[3 * r] * nrm[3 * r] + dirs[3 * r + 1] * nrm[3 * r + 1] + dirs[3 * r + 2] * nrm[3 * r + 2];
    } else {
        hit_t[r] = miss_t;
        id[r] = -1;
        fnorm[r] = 1e6f;
    }
}
}  // namespace

int rt_lidar_cast_rays(rt_ctx *ctx, const rt_scene *scene, const float *o, const float *d, size_t n,
                       size_t packet_size, float miss_t, float *hit_t, float *hit_fnorm, int32_t *hit_id) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    (void)packet_size;
    if (n == 0) return 0;
    NN(o, "origins"); NN(d, "dirs"); NN(hit_t, "hit_t"); NN(hit_fnorm, "hit_fnorm"); NN(hit_id, "hit_id");
    CKR(hipSetDevice(ctx->device));
    float *dbuf = nullptr;
    const size_t per = 3 + 3 + 1 + 3 + 1 + 1 + 1;  // o, d, t, normal, uuid, fnorm, id (4 bytes each)
    CKR(hipMalloc(&dbuf, n * per * 4));
    float *dO = dbuf, *dD = dbuf + n * 3, *dT = dbuf + n * 6, *dN = dbuf + n * 7, *dF = dbuf + n * 11;
    uint32_t *dU = reinterpret_cast<uint32_t *>(dbuf + n * 10);
    int32_t *dI = reinterpret_cast<int32_t *>(dbuf + n * 12);
    int rc = 0;
    do {
        if (hipMemcpy(dO, o, n * 12, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dD, d, n * 12, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("LiDAR cast upload failed");
            rc = -1;
            break;
        }
        // castRaysSoA with the miss distance as the maximum distance (LiDAR.cpp:577-579)
        if (rt_cast_rays_soa_device(ctx, scene, dO, dD, n, miss_t, dT, dN, dU)) { rc = -1; break; }
        k_lidar_convert<<<(unsigned)((n + 255) / 256), 256, 0, ctx->sx.stream>>>(n, dD, dN, dU, miss_t, dT, dF, dI);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->sx.stream) != hipSuccess) {
            set_error("LiDAR conversion kernel failed");
            rc = -1;
            break;
        }
        if (hipMemcpy(hit_t, dT, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hit_fnorm, dF, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hit_id, dI, n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("LiDAR cast download failed");
            rc = -1;
            break;
        }
    } while (0);
    (void)hipFree(dbuf);
    return rc;
}

int rt_lidar_cast_rays_noalloc(rt_ctx *ctx, const rt_scene *scene, const float *d_origins, const float *d_dirs,
                               size_t n, float miss_t, float *d_hit_t, float *d_hit_fnorm, int32_t *d_hit_id) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    if (n == 0) return 0;
    NN(d_origins, "origins"); NN(d_dirs, "dirs"); NN(d_hit_t, "hit_t"); NN(d_hit_fnorm, "hit_fnorm");
    NN(d_hit_id, "hit_id");
    if (n > RT_NOALLOC_MAX_RAYS) {
        set_error("rt_lidar_cast_rays_noalloc: " + std::to_string(n) + " rays in one call (at most " +
                  std::to_string(RT_NOALLOC_MAX_RAYS) + "); cast larger scans in slices");
        return -1;
    }
    const int dev = ctx->device;
    if (dev >= NOALLOC_MAX_DEVICES) { set_error("rt_lidar_cast_rays_noalloc: device index above 63"); return -1; }
    void *base;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        base = g_noalloc_ptr[dev];
    }
    if (!base) { set_error("rt_lidar_cast_rays_noalloc: no scratch for this device (rt_create did not set it up)"); return -1; }
    std::lock_guard<std::mutex> run(g_noalloc_run[dev]);
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(dev));
    hipStream_t st = ctx->sx.stream;
    int *err = reinterpret_cast<int *>(base);
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(static_cast<char *>(base) + 8);
    float *nrm = reinterpret_cast<float *>(static_cast<char *>(base) + NOALLOC_HEAD);
    uint32_t *uuid = reinterpret_cast<uint32_t *>(nrm + (size_t)RT_NOALLOC_MAX_RAYS * 3);
    // the error flag keeps the last call's value: clear it with a kernel (no memset node if captured)
    k_fill_u32<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t *>(err), 1, 0u);
    CKR(hipGetLastError());
    // castRaysSoA with the miss distance as the maximum distance (LiDAR.cpp:577-579); hits write hit_t
    // directly, misses leave it for the conversion
    if (launch_cast(scene->s.d, scene->s.h.depth, n, d_origins, d_dirs, miss_t, d_hit_t, nrm, uuid, counter, err,
                    nullptr, ctx->sx.ncu, st, ctx->sx.watertight))
        return -1;
    k_lidar_convert<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, d_dirs, nrm, uuid, miss_t, d_hit_t, d_hit_fnorm,
                                                                 d_hit_id);
    CKR(hipGetLastError());
    int herr = 0;
    CKR(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
    CKR(hipStreamSynchronize(st));
    if (herr) { set_error("traversal error code " + std::to_string(herr) + " (1 = stack overflow, 2 = malformed tree)"); return -1; }
    return 0;
}

static int check_request(const rt_scene *scene, const rt_band_request *q) {
    NN(scene, "scene");
    NN(q, "request");
    if (q->nbands == 0) { set_error("request has no bands"); return -1; }
    NN(q->band_depth, "band_depth");
    NN(q->emission, "emission");
    if (scene->s.h.P) { NN(q->rho, "rho"); NN(q->tau, "tau"); }
    uint32_t maxd = 0;
    for (uint32_t b = 0; b < q->nbands; b++) maxd = std::max(maxd, q->band_depth[b]);
    NN(q->diffuse_seed, "diffuse_seed");
    (void)maxd;
    return 0;
}

int rt_run_band(rt_ctx *ctx, const rt_scene *scene, const rt_band_request *q, float *R_out) {
    NN(ctx, "ctx");
    if (check_request(scene, q)) return -1;
    NN(R_out, "R_out");
    if (q->nsources) { NN(q->source_dir, "source_dir"); NN(q->source_flux, "source_flux"); }
    NN(q->diffuse_flux, "diffuse_flux");
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    static const float zero3[3] = {0.f, 0.f, 1.f};
    SunSet s{1, q->nsources ? q->source_dir : zero3, q->source_flux, q->diffuse_flux, &q->direct_seed};
    return run_solve(ctx->sx, scene->s, *q, s, R_out, 0);
}

int rt_run_band_batch(rt_ctx *ctx, const rt_scene *scene, const rt_band_request *base, const rt_sun_batch *suns,
                      float *R_out, int out_on_device) {
    NN(ctx, "ctx");
    if (check_request(scene, base)) return -1;
    NN(suns, "suns");
    NN(R_out, "R_out");
    if (suns->nsuns == 0) { set_error("batch has no suns"); return -1; }
    if (base->nsources) { NN(suns->source_dir, "suns.source_dir"); NN(suns->source_flux, "suns.source_flux"); }
    NN(suns->diffuse_flux, "suns.diffuse_flux");
    NN(suns->direct_seed, "suns.direct_seed");
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    SunSet s{suns->nsuns, suns->source_dir, suns->source_flux, suns->diffuse_flux, suns->direct_seed,
             suns->diffuse_peak_dir};
    return run_solve(ctx->sx, scene->s, *base, s, R_out, out_on_device);
}

int rt_set_predicates(rt_ctx *ctx, int mode) {
    NN(ctx, "ctx");
    if (mode != RT_PREDICATES_REFERENCE && mode != RT_PREDICATES_WATERTIGHT) {
        set_error("rt_set_predicates: mode must be RT_PREDICATES_REFERENCE (0) or RT_PREDICATES_WATERTIGHT (1)");
        return -1;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->sx.watertight = mode == RT_PREDICATES_WATERTIGHT;
    return 0;
}

int rt_set_energy_accounting(rt_ctx *ctx, int on) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->sx.energy_acct = on != 0;
    return 0;
}

static int energy_ready(rt_ctx *ctx, uint32_t nchannels) {
    if (!ctx->sx.energy_valid && ctx->sx.energy_glass) {
        set_error("no energy bookkeeping: the last solve had translucent covers, which the bookkeeping does not follow");
        return -1;
    }
    if (!ctx->sx.energy_valid) {
        set_error("no energy bookkeeping: call rt_set_energy_accounting(ctx, 1) before the solve");
        return -1;
    }
    if (nchannels != ctx->sx.energy_C) {
        set_error("channel count " + std::to_string(nchannels) + " does not match the last solve's " +
                  std::to_string(ctx->sx.energy_C));
        return -1;
    }
    return 0;
}

int rt_energy_residual_fp64(rt_ctx *ctx, double *out, uint32_t nchannels) {
    NN(ctx, "ctx");
    NN(out, "out");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (energy_ready(ctx, nchannels)) return -1;
    std::copy(ctx->sx.energy.begin(), ctx->sx.energy.end(), out);
    return 0;
}

int rt_energy_passes_fp64(rt_ctx *ctx, double *out, uint32_t nchannels, uint32_t max_passes, uint32_t *npasses) {
    NN(ctx, "ctx");
    NN(npasses, "npasses");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (energy_ready(ctx, nchannels)) return -1;
    const uint32_t np = ctx->sx.energy_npass;
    *npasses = np;
    if (out) {
        const size_t n = (size_t)std::min(np, max_passes) * RT_ENERGY_PASS_ROWS * nchannels;
        std::copy(ctx->sx.energy_passes.begin(), ctx->sx.energy_passes.begin() + n, out);
    }
    return 0;
}

int rt_set_launch_cfg(rt_ctx *ctx, int kernel, const rt_launch_cfg *cfg) {
    NN(ctx, "ctx");
    NN(cfg, "cfg");
    if (kernel < 0 || kernel > 2) { set_error("unknown kernel id"); return -1; }
    if (!(cfg->waves_per_block == 1 || cfg->waves_per_block == 2 || cfg->waves_per_block == 4)) {
        set_error("waves_per_block must be 1, 2 or 4");
        return -1;
    }
    if (cfg->blocks_per_cu < 1 || cfg->blocks_per_cu > 64) { set_error("blocks_per_cu must be in [1, 64]"); return -1; }
    if (cfg->chunk > 256) { set_error("chunk must be in [0, 256] (0 = 1)"); return -1; }
    if (cfg->grid_density > 64) { set_error("grid_density must be in [0, 64] (0 = 2)"); return -1; }
    ctx->sx.cfg[kernel].waves_per_block = cfg->waves_per_block;
    ctx->sx.cfg[kernel].blocks_per_cu = cfg->blocks_per_cu;
    ctx->sx.cfg[kernel].chunk = cfg->chunk ? cfg->chunk : 1;
    ctx->sx.cfg[kernel].grid_density = cfg->grid_density ? cfg->grid_density : 2;
    return 0;
}

int rt_get_launch_cfg(rt_ctx *ctx, int kernel, rt_launch_cfg *cfg) {
    NN(ctx, "ctx");
    NN(cfg, "cfg");
    if (kernel < 0 || kernel > 2) { set_error("unknown kernel id"); return -1; }
    cfg->waves_per_block = ctx->sx.cfg[kernel].waves_per_block;
    cfg->blocks_per_cu = ctx->sx.cfg[kernel].blocks_per_cu;
    cfg->chunk = ctx->sx.cfg[kernel].chunk;
    cfg->grid_density = ctx->sx.cfg[kernel].grid_density;
    return 0;
}

int rt_set_counting(rt_ctx *ctx, int on) {
    NN(ctx, "ctx");
    CKR(hipSetDevice(ctx->device));
    if (!ctx->sx.stats) CKR(hipMalloc(&ctx->sx.stats, 12 * sizeof(unsigned long long)));
    CKR(hipMemset(ctx->sx.stats, 0, 12 * sizeof(unsigned long long)));
    ctx->sx.counting = on != 0;
    return 0;
}

int rt_get_counters(rt_ctx *ctx, uint64_t *out) {
    NN(ctx, "ctx");
    NN(out, "out");
    CKR(hipSetDevice(ctx->device));
    if (!ctx->sx.stats) { for (int k = 0; k < 12; k++) out[k] = 0; return 0; }
    CKR(hipStreamSynchronize(ctx->sx.stream));
    unsigned long long h[12];
    CKR(hipMemcpy(h, ctx->sx.stats, sizeof(h), hipMemcpyDeviceToHost));
    for (int k = 0; k < 12; k++) out[k] = h[k];
    return 0;
}

int rt_enable_timing(rt_ctx *ctx, int on) {
    NN(ctx, "ctx");
    ctx->sx.timing = on != 0;
    return 0;
}

int rt_set_direct_grids(rt_ctx *ctx, int on) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->sx.use_grid = on != 0;
    return 0;
}

int rt_get_timing(rt_ctx *ctx, rt_timing *out) {
    NN(ctx, "ctx");
    NN(out, "out");
    *out = ctx->sx.timing_info;
    out->direct_grid = ctx->sx.last_direct_grid ? 1u : 0u;
    out->grid_entries = ctx->sx.last_direct_grid ? ctx->sx.last_grid_entries : 0u;
    return 0;
}

int rt_debug_diffuse_hits(rt_ctx *ctx, const rt_scene *scene, uint32_t n, uint32_t seed, uint32_t *out) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    NN(out, "out");
    if (n == 0) { set_error("n must be positive"); return -1; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    return debug_diffuse_hits(ctx->sx, scene->s, n, seed, out);
}

int rt_debug_diffuse_rays(rt_ctx *ctx, const rt_scene *scene, uint32_t n, uint32_t seed, float *out) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    NN(out, "out");
    if (n == 0) { set_error("n must be positive"); return -1; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    size_t bytes = (size_t)scene->s.h.P * 2 * n * n * 6 * 4;
    float *d = nullptr;
    CKR(hipMalloc(&d, bytes + 4));
    int rc = launch_debug_rays(scene->s.d, scene->s.h.P, n, seed, d, ctx->sx.stream);
    if (!rc && (hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, ctx->sx.stream) != hipSuccess ||
                hipStreamSynchronize(ctx->sx.stream) != hipSuccess)) {
        set_error("debug ray download failed");
        rc = -1;
    }
    (void)hipFree(d);
    return rc;
}

int rt_debug_trace(rt_ctx *ctx, const rt_scene *scene, uint32_t n, const float *rays, const uint32_t *skip, float tmin,
                   int brute, uint32_t *hit, float *t) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    if (n == 0) return 0;
    NN(rays, "rays"); NN(skip, "skip"); NN(hit, "hit"); NN(t, "t");
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    char *d = nullptr;
    size_t b_r = (size_t)n * 24, b_s = (size_t)n * 4;
    CKR(hipMalloc(&d, b_r + 3 * b_s + 64));
    float *dr = (float *)d;
    uint32_t *ds = (uint32_t *)(d + b_r), *dh = (uint32_t *)(d + b_r + b_s);
    float *dt = (float *)(d + b_r + 2 * b_s);
    int *err = (int *)(d + b_r + 3 * b_s);
    hipStream_t st = ctx->sx.stream;
    int rc = 0, herr = 0;
    if (hipMemcpyAsync(dr, rays, b_r, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(ds, skip, b_s, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemsetAsync(err, 0, 4, st) != hipSuccess) {
        set_error("debug trace upload failed");
        rc = -1;
    }
    if (!rc) rc = launch_debug_trace(scene->s.d, n, dr, ds, tmin, brute, dh, dt, err, st, scene_flags(scene->s, ctx->sx));
    if (!rc && (hipMemcpyAsync(hit, dh, b_s, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(t, dt, b_s, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)) {
        set_error("debug trace download failed");
        rc = -1;
    }
    (void)hipFree(d);
    if (!rc && herr) { set_error("traversal error code " + std::to_string(herr)); rc = -1; }
    return rc;
}

int rt_debug_direct_sums(rt_ctx *ctx, const rt_scene *scene, uint32_t d, uint32_t seed, uint32_t nsrc,
                         const float *src_dir, float *out) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    NN(out, "out");
    NN(src_dir, "src_dir");
    if (d == 0 || nsrc == 0) { set_error("d and nsrc must be positive"); return -1; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    return debug_direct_sums(ctx->sx, scene->s, d, seed, nsrc, src_dir, nullptr, nullptr, nullptr, out);
}

int rt_debug_direct_sums_src(rt_ctx *ctx, const rt_scene *scene, uint32_t d, uint32_t seed, uint32_t nsrc,
                             const float *src_pos, const uint32_t *types, const float *widths,
                             const float *rotations, float *out) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    NN(out, "out");
    NN(src_pos, "src_pos");
    if (d == 0 || nsrc == 0) { set_error("d and nsrc must be positive"); return -1; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    return debug_direct_sums(ctx->sx, scene->s, d, seed, nsrc, src_pos, types, widths, rotations, out);
}

}  // extern "C"
