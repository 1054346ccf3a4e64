// rt_internal.h -- host-side runtime objects behind the C ABI (not part of the public interface).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "rt_types.h"

namespace rt {

void set_error(const std::string &msg);

// A resident scene: device arrays plus the DevScene descriptor (host copy h, device copy d).
struct Scene {
    DevScene h{};
    DevScene *d = nullptr;
    uint32_t *twos = nullptr;
    std::vector<void *> own;
    size_t bytes = 0;
    int builder = 0;  // 1 = PLOC, 0 = Karras LBVH
};

// Largest primitive count of a scene: leaf tags and hit words carry a 30-bit primitive index.
constexpr uint64_t RT_MAX_PRIMS = (uint64_t)RT_ID_BITS + 1u;
// Largest work-item count of one persistent launch: the 32-bit work counter overshoots the item count by at
// most one chunk per wave (<= 256 items x 65,536 waves) before every wave has seen the end.
constexpr uint64_t RT_MAX_WORK_ITEMS = 0xFFFFFFFFull - (1ull << 25);
// Checks a primitive set before it is staged or built: at most RT_MAX_PRIMS primitives, every type one this
// backend takes (0 patch, 1 triangle, 3 tile, 4 voxel); disks (2) and anything else are rejected with an
// error naming the type, so no primitive silently drops out of the scene.  `who` prefixes the message.
int validate_geometry(uint64_t P, const uint32_t *types, const char *who);
// Rejects a launch of more than RT_MAX_WORK_ITEMS work items (e.g. suns x sources x primitives of a direct pass).
int check_work_items(uint64_t items, const char *what);

int scene_build(Scene &S, uint32_t P, const float *T, const uint32_t *type, const uint32_t *uuid,
                const uint32_t *twos, int px, int py, hipStream_t st);
void scene_free(Scene &S);
// Texture masks (RayTracingGeometry mask_data / mask_sizes / mask_IDs / uv_data / uv_IDs,
// R/include/RayTracingTypes.h:72-77): nmasks masks of sizes[m] = (width, height) texels, one byte per texel
// (0 = transparent) concatenated row-major in texels; mask_id[P] (-1 = none); uv[P][4][2] (null = zeros)
// and uv_id[P] (null or < 0: default patch UVs).  Replaces the masks of the scene.
int scene_set_masks(Scene &S, uint32_t nmasks, const uint8_t *texels, const uint32_t *sizes, const int32_t *mask_id,
                    const float *uv, const int32_t *uv_id, hipStream_t st);

// Launch configuration chosen per kernel (autotuned or default).
struct LaunchCfg {
    uint32_t waves_per_block = 4;  // 64-lane waves per workgroup
    uint32_t blocks_per_cu = 4;    // persistent workgroups per CU
    uint32_t chunk = 1;            // work items taken per work-queue fetch
    uint32_t grid_density = 2;     // direct pass: occlusion-grid cells per primitive
};
// Default shape of each kernel class (RT_KERNEL_*): the autotuner's pick on both the 1 M-primitive orchard
// and the 825 k-primitive masked canopy (two-wave workgroups, 16 per CU, for the traversal kernels;
// 8 four-wave workgroups per CU for the gather).
inline LaunchCfg default_launch_cfg(int kernel) {
    LaunchCfg c;
    if (kernel == 2) { c.waves_per_block = 4; c.blocks_per_cu = 8; }
    else { c.waves_per_block = 2; c.blocks_per_cu = 16; }
    return c;
}

// kernel launchers (rt_trace.hip)
// stats (may be null): [0] BVH node fetches, [1] primitive tests, [2] rays traced
// flags: traversal flags -- 1 = the scene carries texture masks (the kernels without mask tests run
// otherwise), 2 = watertight intersection predicates (rt_set_predicates); see scene_flags()
int launch_direct(const DirectJob *jobs_dev, uint32_t njobs, uint32_t total_items, uint32_t depth, uint32_t *counter,
                  int *err, unsigned long long *stats, const LaunchCfg &cfg, int ncu, hipStream_t st, int flags);
// Non-collimated sources (k_direct_src): items [item_lo, item_lo + nitems) of jobs with ascending item_begin.
int launch_direct_src(const DirectJob *jobs_dev, uint32_t njobs, uint32_t item_lo, uint32_t nitems, uint32_t depth,
                      uint32_t *counter, int *err, unsigned long long *stats, int ncu, hipStream_t st, int flags);
// Any mix of source types: hjobs is the host copy of jobs_dev, collimated jobs first (their items precede
// the others'); collimated jobs go to k_direct, the rest to k_direct_src.
int launch_direct_jobs(const DirectJob *hjobs, const DirectJob *jobs_dev, uint32_t njobs, uint32_t total_items,
                       uint32_t depth, uint32_t *counter, int *err, unsigned long long *stats, const LaunchCfg &cfg,
                       int ncu, hipStream_t st, int flags);
// one diffuse job (device pointer) per launch: items [0, total_items) of that job
int launch_diffuse(const DiffuseJob *job_dev, uint32_t total_items, uint32_t depth, uint32_t *counter, int *err,
                   unsigned long long *stats, const LaunchCfg &cfg, int ncu, hipStream_t st, int flags);
// acct: the jobs carry energy-bookkeeping outputs (GatherJob.acct / .miss)
int launch_gather(const GatherJob *jobs_dev, uint32_t njobs, uint32_t total_items, uint32_t cmax, uint32_t *counter,
                  const LaunchCfg &cfg, int ncu, hipStream_t st, bool acct = false);
// Solves with translucent covers (GlassCtx g, device): the direct jobs (items ascending by dj.item_begin),
// one diffuse/scattering pass traced and gathered together, and the covers' pushed shares added to the
// radiation buffers (PC = P * C).
int launch_direct_glass(const GlassDirectJob *jobs_dev, uint32_t njobs, uint32_t total_items, uint32_t depth,
                        uint32_t *counter, const GlassCtx *g, int *err, int ncu, hipStream_t st, int flags);
int launch_diffuse_glass(const DiffuseJob *fj_dev, const GatherJob *gj_dev, uint32_t total_items, uint32_t depth,
                         uint32_t *counter, const GlassCtx *g, int *err, int ncu, hipStream_t st, int flags);
int launch_add_push(size_t PC, double *push, float *in, float *st, float *sb, hipStream_t s);
// Camera pixels (labels = false) or pixel labels/depths (labels = true) of one tile.
int launch_camera(const CameraLaunch &c, uint32_t depth, bool labels, int *err, unsigned long long *stats,
                  hipStream_t st, int flags);
// counter: device scratch for the work queue (zeroed by the launch)
int launch_cast(const DevScene *sc, uint32_t depth, uint64_t n, const float *o, const float *d, float maxd, float *t,
                float *nrm, uint32_t *uuid, unsigned long long *counter, int *err, unsigned long long *stats,
                int ncu, hipStream_t st, bool watertight = false);
int launch_debug_rays(const DevScene *sc, uint32_t P, uint32_t n, uint32_t seed, float *out, hipStream_t st);
int launch_debug_trace(const DevScene *sc, uint32_t n, const float *rays, const uint32_t *skip, float tmin, int brute,
                       uint32_t *hit, float *t, int *err, hipStream_t st, int flags = 0);

}  // namespace rt
