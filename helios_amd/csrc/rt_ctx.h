// rt_ctx.h -- definitions of the opaque C-ABI handles (internal).
#pragma once

#include <mutex>
#include <string>
#include <vector>

#include "rt_solve.h"

// State of the call-by-call backend interface (the helios::RayTracingBackend counterpart,
// R/include/RayTracingBackend.h:43-312): one resident scene, materials, sources and the five
// [P][B_launch] radiation buffers that VulkanComputeBackend keeps (radiation_in, scatter_top,
// scatter_bottom, radiation_out_top, radiation_out_bottom).
struct BackendState {
    // geometry staged by updateGeometry until buildAccelerationStructure
    uint32_t gP = 0;
    std::vector<float> gT;
    std::vector<uint32_t> gtype, guuid, gtwos;
    int gpx = 0, gpy = 0;
    bool geometry_staged = false;
    // texture masks staged by rt_update_texture_masks, applied by the next build
    uint32_t nmasks = 0;
    std::vector<uint8_t> mtex;
    std::vector<uint32_t> msize;
    std::vector<int32_t> mid, muvid;
    std::vector<float> muv;

    rt::Scene scene;
    bool has_scene = false;
    uint32_t P = 0;   // primitives of the built scene
    uint32_t Bg = 0;  // global bands of the materials
    uint32_t Bl = 0;  // launch bands (set by zeroRadiationBuffers)
    uint32_t S = 0;   // sources
    bool has_materials = false, has_fluxes = false;
    std::vector<float> src_dir;  // [S][3] direction (collimated) or position
    std::vector<rt::SourceShape> shapes;  // [S]
    void *mat = nullptr;         // rho [P][Bg] then tau [P][Bg] (source-0 slice)
    size_t mat_cap = 0;
    void *flux = nullptr;        // [S][Bl]
    size_t flux_cap = 0;
    void *rad = nullptr;         // in, st, sb, ot, ob, each [P][Bl]
    size_t rad_cap = 0;
    rt::Workspace ws;            // per-launch scratch (strength sums, hit words, band tables, jobs)
    // camera state (updateSkyModel, launchCameraRays, launchPixelLabelRays, getCameraResults)
    uint32_t sky_B = 0;
    std::vector<float> sky_params, sky_rad, solar_rad, cam_dflux;  // [B][4], [B], [B], [B]
    std::vector<uint32_t> emis_flag;                                 // [B]
    float sun_dir[3] = {0.f, 0.f, 1.f};
    float solar_cos = 0.f;
    bool has_sky = false;
    void *cam = nullptr;         // camera_radiation [B_launch][full pixels] (band-major)
    size_t cam_cap = 0;
    uint64_t cam_pix = 0;        // full pixels the radiation image holds (0: none yet)
    uint32_t cam_B = 0;          // launch bands of that image
    void *lab = nullptr;         // pixel labels [pixels] then depths [pixels]
    size_t lab_cap = 0;
    uint64_t lab_pix = 0;        // pixels the label/depth buffers hold
    // specular reflection: per-primitive exponent and scale, per-camera source weights, and the incident
    // radiation of each source that direct launches with specular on accumulate
    void *spm = nullptr;         // specular_exponent [P] then specular_scale [P]
    size_t spm_cap = 0;
    bool has_specular = false;
    void *fcam = nullptr;        // source_fluxes_cam [S][Bl]
    size_t fcam_cap = 0;
    bool has_fcam = false;
    void *spec = nullptr;        // radiation_specular [S][P][Bl]
    size_t spec_cap = 0;
    bool spec_live = false;      // spec holds this launch set's sums (false: all zero)
    uint32_t spec_S = 0;         // sources of those sums
    // translucent covers (updateMaterials' glass fields, source-0 slice): [P][Bg] each, host side; a launch
    // selects its bands and uploads the covers' parameters with its other tables
    std::vector<uint8_t> g_is;
    std::vector<float> g_n, g_kl;
    float *rho() const { return (float *)mat; }
    float *tau() const { return (float *)mat + (size_t)P * Bg; }
    float *rbuf(int k) const { return (float *)rad + (size_t)k * P * Bl; }  // 0 in 1 st 2 sb 3 ot 4 ob
};

struct rt_ctx {
    int device = 0;
    std::string name;
    rt::SolveCtx sx;
    std::mutex mu;
    BackendState be;
};

struct rt_scene {
    rt::Scene s;
};

#define CKR(x)                                                                         \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            rt::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #x); \
            return -1;                                                                 \
        }                                                                              \
    } while (0)

#define NN(p, what)                                            \
    do {                                                       \
        if (!(p)) {                                            \
            rt::set_error(std::string(what) + " is NULL");     \
            return -1;                                         \
        }                                                      \
    } while (0)
