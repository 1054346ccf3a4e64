// rt_solve.h -- solve orchestration shared by the C ABI, the batcher and the request server.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/helios_rt.h"
#include "rt_internal.h"

namespace rt {

// grow-only device workspace (one per stream that runs solves)
struct Workspace {
    void *base = nullptr;
    size_t cap = 0;
    int reserve(size_t bytes);
    void release();
    static size_t capacity_for(size_t bytes);  // what reserve(bytes) allocates when it has to grow
};

struct EvRec {
    int cls;
    hipEvent_t a, b;
};

// Occlusion grids of the collimated-source directions of one solve (rt_grid.hip).
struct GridSet {
    Workspace small, cells, items, keys, sortbuf;
    SunGrid *dev = nullptr;  // [n] device headers
    uint32_t n = 0;
    size_t entries = 0;      // list entries over all directions
    size_t cells_total = 0;
    size_t cap = 0;          // byte limit for cells + lists (0 = none); above it the BVH is traversed
    float density = 2.f;     // grid cells per traced primitive
    // bytes a server reserves for the grids of one solve of `njobs` directions over P primitives
    static size_t budget(uint32_t P, uint32_t njobs) { return (size_t)512 * P * njobs + (1u << 20); }
    void release() {
        small.release(); cells.release(); items.release(); keys.release(); sortbuf.release();
        dev = nullptr; n = 0;
    }
};
// Builds grids for n directions (host [n][3]).  Returns 0 on success, 1 when grids do not apply (periodic
// scene, degenerate input, too large) and the BVH is to be traversed instead, -1 on a HIP error.
int build_sun_grids(GridSet &gs, const Scene &S, const float *dirs, uint32_t n, hipStream_t st);

// Everything one stream needs to run solves.
struct SolveCtx {
    hipStream_t stream = nullptr;
    int ncu = 256;
    LaunchCfg cfg[3] = {default_launch_cfg(0), default_launch_cfg(1), default_launch_cfg(2)};
    Workspace ws;
    bool timing = false;
    std::vector<EvRec> ev;
    rt_timing timing_info{};
    // FP64 energy bookkeeping (rt_set_energy_accounting): off by default; when on, each solve records the
    // per-channel totals (RT_ENERGY_ROWS rows of C) and per-pass rows (RT_ENERGY_PASS_ROWS rows of C per
    // diffuse/scattering pass) of its channels
    bool energy_acct = false;
    bool energy_valid = false;
    bool energy_glass = false;          // the last solve had translucent covers, which the bookkeeping skips
    uint32_t energy_C = 0, energy_npass = 0;
    std::vector<double> energy;         // [RT_ENERGY_ROWS][C]
    std::vector<double> energy_passes;  // [npass][RT_ENERGY_PASS_ROWS][C]
    // traversal counters: [0..2] direct, [3..5] diffuse, [6..8] castRays, [9..11] camera
    // (node fetches, prim tests, rays)
    bool counting = false;
    unsigned long long *stats = nullptr;  // device, 12 counters
    bool use_grid = true;                 // occlusion grids for direct passes of non-periodic scenes
    bool last_direct_grid = false;        // the last solve's direct pass used them
    uint32_t last_grid_entries = 0;       // their list entries per direction per primitive, x100
    GridSet grids;
    bool watertight = false;              // watertight intersection predicates (rt_set_predicates)
    unsigned long long *stat_ptr(int k) { return counting ? stats + 3 * k : nullptr; }
};
// traversal flags of launches on scene S from this context (RT_FL_*)
inline int scene_flags(const Scene &S, const SolveCtx &sx) {
    return (S.h.mdesc ? RT_FL_MASK : 0) | (sx.watertight ? RT_FL_WT : 0);
}

// The per-sun part of a solve (K suns share everything else).
// Shape of one radiation source as the direct kernels take it.
struct SourceShape {
    uint32_t type = 0;     // RT_SOURCE_*
    float width[2] = {0.f, 0.f};
    float rot[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
};
// Validates S sources (types may be null = all collimated; widths [S][2], rotations [S][3]) and builds
// their shapes, rotation matrices from ZYX Euler angles (transforms.glsl:49-64) evaluated in double and
// rounded.  Returns -1 with the error set on invalid input.
int source_shapes(uint32_t S, const uint32_t *types, const float *widths, const float *rotations,
                  std::vector<SourceShape> &out, const char *who);
// Fills a direct job's source fields from shapes[s]; skip_draws counts the earlier non-collimated sources.
void set_job_source(DirectJob &j, const std::vector<SourceShape> &shapes, uint32_t s);

struct SunSet {
    uint32_t K;
    const float *dir;     // [K][S][3]
    const float *flux;    // [K][S][B]
    const float *dflux;   // [K][B]
    const uint32_t *seed; // [K]
    const float *peak = nullptr;  // [K][B][3] anisotropic-sky peak directions, null = the request's
};

// Optical properties already resident in device memory ([P][B] rho, tau, eps or null, [P] temp or null).
struct DevMaterials {
    const float *rho, *tau, *eps, *temp;
};

// Device bytes a solve of K suns needs in its workspace.
// acct: with the energy bookkeeping's buffers
size_t solve_workspace_bytes(uint32_t P, const rt_band_request &q, uint32_t K, bool resident_materials,
                             bool acct = false);

int run_solve(SolveCtx &sx, const Scene &S, const rt_band_request &q, const SunSet &suns, float *R_out,
              int out_on_device, const DevMaterials *dm = nullptr);
// radiation_in / scatter_top / scatter_bottom += direct strength sums x source flux x optical properties;
// gsum (per-band sums [S][P][B][2] of a launch with translucent covers) replaces dsum when given
int launch_deposit_direct(uint32_t P, uint32_t C, uint32_t B, uint32_t S, const float *dsum, const float *flux,
                          const float *rho, const float *tau, const uint8_t *mask, float *in, float *st, float *sb,
                          hipStream_t s, const float *gsum = nullptr);
int debug_diffuse_hits(SolveCtx &sx, const Scene &S, uint32_t n, uint32_t seed, uint32_t *out);
int debug_direct_sums(SolveCtx &sx, const Scene &S, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_dir,
                      const uint32_t *types, const float *widths, const float *rotations, float *out);

}  // namespace rt
