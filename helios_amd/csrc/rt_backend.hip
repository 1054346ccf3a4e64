// rt_backend.hip -- call-by-call backend entry points: the helios::RayTracingBackend interface
// (R/include/RayTracingBackend.h:43-312) as C functions over one context-owned resident scene.
//
// A RadiationModel that keeps its own pass loop (R/src/RadiationModel.cpp:3804-4356) drives these
// in the order it drives VulkanComputeBackend.  Every buffer stays in HBM between calls; only the
// explicit upload/download calls move data over PCIe.  Launch calls run the same persistent
// traversal kernels as rt_run_band: a direct launch is one direct trace job per source followed by
// the deposit kernel, a diffuse launch is one trace pass of the requested face followed by the
// gather kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rt_ctx.h"

using namespace rt;

namespace {

__global__ void k_select_bands(uint32_t P, uint32_t Bg, uint32_t Bl, const uint32_t *__restrict__ map,
                               const float *__restrict__ rho, const float *__restrict__ tau, float *__restrict__ rl,
                               float *__restrict__ tl) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)P * Bl) return;
    uint32_t i = (uint32_t)(q / Bl), b = (uint32_t)(q % Bl);
    size_t g = (size_t)i * Bg + map[b];
    rl[q] = rho[g];
    tl[q] = tau[g];
}

inline uint32_t nblk(size_t n) { return (uint32_t)((n + 255) / 256); }

// radiation_specular[s][i][b] += flux[s][b] * (top + bottom strength sum of source s on primitive i): the
// radiation each source delivers to each primitive in a direct launch, which the camera's specular term
// reads (R/shaders/direct_raygen.comp:717-726 adds flux * strength for every unoccluded ray).  Bands
// that take no energy in the launch add nothing.
__global__ void k_deposit_specular(uint32_t P, uint32_t B, uint32_t S, const float *__restrict__ dsum,
                                   const float *__restrict__ flux, const uint8_t *__restrict__ active,
                                   float *__restrict__ spec) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)S * P * B) return;
    const uint32_t b = (uint32_t)(q % B);
    const size_t si = q / B;  // s * P + i
    const uint32_t s = (uint32_t)(si / P);
    if (!active[b]) return;
    spec[q] += flux[(size_t)s * B + b] * (dsum[si * 2] + dsum[si * 2 + 1]);
}

struct Carve {
    char *p;
    size_t off = 0;
    template <typename T>
    T *take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T *r = reinterpret_cast<T *>(p + off);
        off += n * sizeof(T);
        return r;
    }
};

int grow(void *&p, size_t &cap, size_t bytes) {
    if (bytes <= cap) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    CKR(hipMalloc(&p, bytes));
    cap = bytes;
    return 0;
}

int check_err(rt_ctx *ctx, int *err) {
    int herr = 0;
    CKR(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->sx.stream));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    if (herr) {
        set_error("traversal error code " + std::to_string(herr) + " (1 = stack overflow, 2 = malformed tree)");
        return -1;
    }
    return 0;
}

int need_scene(const BackendState &be, const char *who) {
    if (!be.has_scene) {
        set_error(std::string(who) + ": no acceleration structure (call rt_update_geometry and "
                                     "rt_build_acceleration_structure first)");
        return -1;
    }
    return 0;
}

int need_launch_buffers(const BackendState &be, const char *who) {
    if (need_scene(be, who)) return -1;
    if (be.Bl == 0 || !be.rad) {
        set_error(std::string(who) + ": radiation buffers not sized (call rt_zero_radiation_buffers first)");
        return -1;
    }
    return 0;
}

// Validates a launch's band table and writes the launch -> global band map and the active flags.
int launch_bands(const BackendState &be, const rt_launch_params *p, std::vector<uint32_t> &map,
                 std::vector<uint8_t> &active, const char *who) {
    if (p->num_bands_launch != be.Bl) {
        set_error(std::string(who) + ": num_bands_launch " + std::to_string(p->num_bands_launch) +
                  " differs from the launch band count " + std::to_string(be.Bl) + " set by rt_zero_radiation_buffers");
        return -1;
    }
    if (!be.has_materials) { set_error(std::string(who) + ": no materials (call rt_update_materials)"); return -1; }
    if (p->num_bands_global != be.Bg) {
        set_error(std::string(who) + ": num_bands_global " + std::to_string(p->num_bands_global) +
                  " differs from the materials' band count " + std::to_string(be.Bg));
        return -1;
    }
    map.assign(be.Bl, 0);
    active.assign(be.Bl, 1);
    for (uint32_t b = 0; b < be.Bl; b++) {
        map[b] = p->band_map ? p->band_map[b] : b;
        if (map[b] >= be.Bg) {
            set_error(std::string(who) + ": band_map[" + std::to_string(b) + "] = " + std::to_string(map[b]) +
                      " is not a global band");
            return -1;
        }
        if (p->band_active) active[b] = p->band_active[b] ? 1 : 0;
    }
    if (p->launch_offset > be.P || p->launch_count > be.P - p->launch_offset) {
        set_error(std::string(who) + ": primitive range [" + std::to_string(p->launch_offset) + ", +" +
                  std::to_string(p->launch_count) + ") exceeds the " + std::to_string(be.P) + " primitives");
        return -1;
    }
    return 0;
}

// Work items of a launch are 0..launch_count-1 and a kernel's primitive is item - item_begin
// (uint32 arithmetic).  With one job, item_begin = -launch_offset maps item k to primitive
// launch_offset + k, so a launch of a primitive sub-range seeds, traces and writes exactly the
// primitives (and the per-primitive rows of the output buffers) that a full launch would.
inline uint32_t range_begin(uint32_t offset) { return 0u - offset; }

// Translucent covers of a launch: primitives glass in every launch band (R/shaders/direct_raygen.comp:
// 629-637), with their refractive index and absorption product per launch band.  Returns 1 when the launch
// has covers, 0 when not, -1 on error (more launch bands than the glass kernels carry).
int glass_launch(const BackendState &be, const std::vector<uint32_t> &map, std::vector<uint8_t> &cover,
                 std::vector<float> &gn, std::vector<float> &gkl, const char *who) {
    if (be.g_is.empty()) return 0;
    const uint32_t P = be.P, Bl = (uint32_t)map.size(), Bg = be.Bg;
    cover.assign(P, 0);
    bool any = false;
    for (uint32_t i = 0; i < P; i++) {
        bool all = Bl > 0;
        for (uint32_t b = 0; b < Bl && all; b++) all = be.g_is[(size_t)i * Bg + map[b]] != 0;
        cover[i] = all ? 1 : 0;
        any = any || all;
    }
    if (!any) return 0;
    if (Bl > RT_GLASS_MAX_CH) {
        set_error(std::string(who) + ": a launch with translucent covers takes at most " +
                  std::to_string(RT_GLASS_MAX_CH) + " bands, this one has " + std::to_string(Bl));
        return -1;
    }
    gn.assign((size_t)P * Bl, 1.0f);
    gkl.assign((size_t)P * Bl, 0.0f);
    for (uint32_t i = 0; i < P; i++)
        for (uint32_t b = 0; b < Bl; b++) {
            gn[(size_t)i * Bl + b] = be.g_n[(size_t)i * Bg + map[b]];
            gkl[(size_t)i * Bl + b] = be.g_kl[(size_t)i * Bg + map[b]];
        }
    return 1;
}

// Device tables of a launch with covers, carved from the launch's workspace and filled on `st`.
struct GlassTables {
    uint8_t *cover = nullptr;
    float *gn = nullptr, *gkl = nullptr;
    double *push = nullptr;  // [P][B_launch][3] shares pushed into the covers
    GlassCtx *ctx = nullptr;
};

size_t glass_bytes(uint32_t P, uint32_t Bl) {
    return (size_t)P + 2 * (size_t)P * Bl * 4 + (size_t)P * Bl * 3 * 8 + sizeof(GlassCtx) + 6 * 256;
}

int glass_upload(GlassTables &g, const std::vector<uint8_t> &cover, const std::vector<float> &gn,
                 const std::vector<float> &gkl, uint32_t P, uint32_t Bl, hipStream_t st) {
    CKR(hipMemcpyAsync(g.cover, cover.data(), P, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(g.gn, gn.data(), (size_t)P * Bl * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(g.gkl, gkl.data(), (size_t)P * Bl * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemsetAsync(g.push, 0, (size_t)P * Bl * 3 * 8, st));
    GlassCtx h{};
    h.cover = g.cover; h.gn = g.gn; h.gkl = g.gkl; h.push = g.push; h.B = Bl; h.C = Bl;
    CKR(hipMemcpyAsync(g.ctx, &h, sizeof(GlassCtx), hipMemcpyHostToDevice, st));
    // the host tables must outlive the copies
    CKR(hipStreamSynchronize(st));
    return 0;
}

// radiation_specular from the per-band sums of a launch with covers (gsum [S][P][B][2], the flux behind
// the covers' transmittance, as the reference's specular sum takes it)
__global__ void k_deposit_specular_bands(uint32_t P, uint32_t B, uint32_t S, const float *__restrict__ gsum,
                                         const float *__restrict__ flux, const uint8_t *__restrict__ active,
                                         float *__restrict__ spec) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)S * P * B) return;
    const uint32_t b = (uint32_t)(q % B);
    const uint32_t s = (uint32_t)(q / B / P);
    if (!active[b]) return;
    spec[q] += flux[(size_t)s * B + b] * (gsum[q * 2] + gsum[q * 2 + 1]);
}

}  // namespace

void backend_release(BackendState &be) {
    if (be.has_scene) scene_free(be.scene);
    be.has_scene = false;
    if (be.mat) (void)hipFree(be.mat);
    if (be.flux) (void)hipFree(be.flux);
    if (be.rad) (void)hipFree(be.rad);
    if (be.cam) (void)hipFree(be.cam);
    if (be.lab) (void)hipFree(be.lab);
    if (be.spm) (void)hipFree(be.spm);
    if (be.fcam) (void)hipFree(be.fcam);
    if (be.spec) (void)hipFree(be.spec);
    be.mat = be.flux = be.rad = be.cam = be.lab = be.spm = be.fcam = be.spec = nullptr;
    be.mat_cap = be.flux_cap = be.rad_cap = be.cam_cap = be.lab_cap = be.spm_cap = be.fcam_cap = be.spec_cap = 0;
    be.cam_pix = be.lab_pix = 0;
    be.has_specular = be.has_fcam = be.spec_live = false;
    be.ws.release();
}

extern "C" {

int rt_update_geometry(rt_ctx *ctx, uint32_t nprims, const float *transforms, const uint32_t *types,
                       const uint32_t *uuids, const uint32_t *twosided, int periodic_x, int periodic_y) {
    NN(ctx, "ctx");
    if (nprims) { NN(transforms, "transforms"); NN(types, "types"); NN(twosided, "twosided"); }
    if (validate_geometry(nprims, types, "updateGeometry")) return -1;
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    be.gP = nprims;
    be.gT.assign(transforms, transforms + (size_t)nprims * 16);
    be.gtype.assign(types, types + nprims);
    be.gtwos.assign(twosided, twosided + nprims);
    if (uuids) be.guuid.assign(uuids, uuids + nprims);
    else {
        be.guuid.resize(nprims);
        for (uint32_t i = 0; i < nprims; i++) be.guuid[i] = i;
    }
    be.gpx = periodic_x ? 1 : 0;
    be.gpy = periodic_y ? 1 : 0;
    be.geometry_staged = true;
    // new geometry carries no texture masks until rt_update_texture_masks stages them
    be.nmasks = 0;
    be.mtex.clear(); be.msize.clear(); be.mid.clear(); be.muv.clear(); be.muvid.clear();
    return 0;
}

int rt_update_texture_masks(rt_ctx *ctx, uint32_t nprims, uint32_t nmasks, const uint8_t *texels,
                            const uint32_t *sizes, const int32_t *mask_ids, const float *uv, const int32_t *uv_ids) {
    NN(ctx, "ctx");
    if (nmasks) { NN(texels, "texels"); NN(sizes, "sizes"); NN(mask_ids, "mask_ids"); }
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    be.nmasks = nmasks;
    size_t ntex = 0;
    for (uint32_t m = 0; m < nmasks; m++) ntex += (size_t)sizes[2 * m] * sizes[2 * m + 1];
    be.mtex.assign(texels, texels + (nmasks ? ntex : 0));
    be.msize.assign(sizes, sizes + 2 * (size_t)nmasks);
    be.mid.assign(mask_ids, mask_ids + (nmasks ? nprims : 0));
    be.muv.clear();
    be.muvid.clear();
    if (nmasks && uv) be.muv.assign(uv, uv + (size_t)nprims * 8);
    if (nmasks && uv_ids) be.muvid.assign(uv_ids, uv_ids + nprims);
    return 0;
}

int rt_build_acceleration_structure(rt_ctx *ctx) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (!be.geometry_staged) { set_error("rt_build_acceleration_structure: no geometry (call rt_update_geometry)"); return -1; }
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    if (be.has_scene) { scene_free(be.scene); be.has_scene = false; }
    be.scene = Scene{};
    if (scene_build(be.scene, be.gP, be.gT.data(), be.gtype.data(), be.guuid.data(), be.gtwos.data(), be.gpx, be.gpy,
                    ctx->sx.stream))
        return -1;
    be.has_scene = true;
    if (be.nmasks) {
        if (be.mid.size() != be.gP) {
            set_error("rt_build_acceleration_structure: texture masks were staged for " + std::to_string(be.mid.size()) +
                      " primitives, the geometry has " + std::to_string(be.gP));
            return -1;
        }
        if (scene_set_masks(be.scene, be.nmasks, be.mtex.data(), be.msize.data(), be.mid.data(),
                            be.muv.empty() ? nullptr : be.muv.data(), be.muvid.empty() ? nullptr : be.muvid.data(),
                            ctx->sx.stream))
            return -1;
    }
    if (be.P != be.gP) {
        // per-primitive state of the previous geometry no longer applies
        be.has_materials = false;
        be.has_fluxes = false;
        be.has_specular = false;
        be.has_fcam = false;
        be.spec_live = false;
        be.g_is.clear(); be.g_n.clear(); be.g_kl.clear();
        be.Bl = 0;
    }
    be.P = be.gP;
    return 0;
}

int rt_update_materials(rt_ctx *ctx, uint32_t nsources, uint32_t nprims, uint32_t nbands_global, const float *rho,
                        const float *tau) {
    NN(ctx, "ctx");
    NN(rho, "rho");
    NN(tau, "tau");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_scene(be, "rt_update_materials")) return -1;
    if (nprims != be.P) {
        set_error("rt_update_materials: " + std::to_string(nprims) + " primitives, the scene has " + std::to_string(be.P));
        return -1;
    }
    if (nsources == 0 || nbands_global == 0) { set_error("rt_update_materials: no sources or no bands"); return -1; }
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    size_t PB = (size_t)be.P * nbands_global;
    if (grow(be.mat, be.mat_cap, 2 * PB * sizeof(float) + 256)) return -1;
    // new materials carry no covers until rt_update_glass_materials sets them
    be.g_is.clear(); be.g_n.clear(); be.g_kl.clear();
    be.Bg = nbands_global;
    // source-0 slice, as the Vulkan backend's shaders index rho/tau (R/shaders/direct_raygen.comp:684-687)
    CKR(hipMemcpy(be.rho(), rho, PB * sizeof(float), hipMemcpyHostToDevice));
    CKR(hipMemcpy(be.tau(), tau, PB * sizeof(float), hipMemcpyHostToDevice));
    be.has_materials = true;
    return 0;
}

int rt_update_glass_materials(rt_ctx *ctx, uint32_t nprims, uint32_t nbands_global, const uint8_t *is_glass,
                              const float *glass_n, const float *glass_KL) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_scene(be, "rt_update_glass_materials")) return -1;
    if (nprims != be.P || nbands_global != be.Bg || !be.has_materials) {
        set_error("rt_update_glass_materials: " + std::to_string(nprims) + " primitives x " +
                  std::to_string(nbands_global) + " bands, the materials have " + std::to_string(be.P) + " x " +
                  std::to_string(be.Bg) + " (call rt_update_materials first)");
        return -1;
    }
    if (!is_glass) {  // no covers
        be.g_is.clear(); be.g_n.clear(); be.g_kl.clear();
        return 0;
    }
    NN(glass_n, "glass_n");
    const size_t n = (size_t)nprims * nbands_global;
    for (size_t k = 0; k < n; k++)
        if (is_glass[k] && (!(glass_n[k] >= 1.0f) || (glass_KL && !(glass_KL[k] >= 0.0f)))) {
            set_error("rt_update_glass_materials: primitive " + std::to_string(k / nbands_global) + " band " +
                      std::to_string(k % nbands_global) + ": glass needs a refractive index >= 1 and an absorption K*L >= 0");
            return -1;
        }
    be.g_is.assign(is_glass, is_glass + n);
    be.g_n.assign(glass_n, glass_n + n);
    if (glass_KL) be.g_kl.assign(glass_KL, glass_KL + n);
    else be.g_kl.assign(n, 0.0f);
    return 0;
}

int rt_update_sources(rt_ctx *ctx, uint32_t nsources, const uint32_t *types, const float *positions,
                      const float *widths, const float *rotations) {
    NN(ctx, "ctx");
    if (nsources) { NN(types, "types"); NN(positions, "positions"); }
    std::vector<SourceShape> shapes;
    if (source_shapes(nsources, types, widths, rotations, shapes, "rt_update_sources")) return -1;
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (nsources != be.S) { be.has_fluxes = false; be.has_fcam = false; }
    be.S = nsources;
    be.src_dir.assign(positions, positions + (size_t)nsources * 3);
    be.shapes = std::move(shapes);
    return 0;
}

int rt_zero_radiation_buffers(rt_ctx *ctx, uint32_t launch_band_count) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_scene(be, "rt_zero_radiation_buffers")) return -1;
    if (launch_band_count == 0 || launch_band_count > 256) {
        set_error("rt_zero_radiation_buffers: launch_band_count must be 1..256");
        return -1;
    }
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    if (launch_band_count != be.Bl) { be.has_fluxes = false; be.has_fcam = false; }
    be.Bl = launch_band_count;
    be.spec_live = false;  // radiation_specular reads as zero until a direct launch with specular on
    size_t PB = (size_t)be.P * be.Bl;
    if (grow(be.rad, be.rad_cap, 5 * PB * sizeof(float) + 256)) return -1;
    // radiation_in and radiation_out; the scatter buffers are zeroed as well so that a new runBand
    // starts from a clean state (VulkanComputeBackend.cpp:1935-1988 re-creates them when resized)
    CKR(hipMemsetAsync(be.rad, 0, 5 * PB * sizeof(float), ctx->sx.stream));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    return 0;
}

int rt_zero_scatter_buffers(rt_ctx *ctx) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_launch_buffers(be, "rt_zero_scatter_buffers")) return -1;
    CKR(hipSetDevice(ctx->device));
    size_t PB = (size_t)be.P * be.Bl;
    CKR(hipMemsetAsync(be.rbuf(1), 0, 2 * PB * sizeof(float), ctx->sx.stream));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    return 0;
}

int rt_copy_scatter_to_radiation(rt_ctx *ctx) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_launch_buffers(be, "rt_copy_scatter_to_radiation")) return -1;
    CKR(hipSetDevice(ctx->device));
    size_t PB = (size_t)be.P * be.Bl;
    CKR(hipMemcpyAsync(be.rbuf(3), be.rbuf(1), PB * sizeof(float), hipMemcpyDeviceToDevice, ctx->sx.stream));
    CKR(hipMemcpyAsync(be.rbuf(4), be.rbuf(2), PB * sizeof(float), hipMemcpyDeviceToDevice, ctx->sx.stream));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    return 0;
}

int rt_upload_radiation_out(rt_ctx *ctx, const float *top, const float *bottom, size_t count) {
    NN(ctx, "ctx");
    NN(top, "top");
    NN(bottom, "bottom");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_launch_buffers(be, "rt_upload_radiation_out")) return -1;
    size_t PB = (size_t)be.P * be.Bl;
    if (count != PB) {
        set_error("rt_upload_radiation_out: " + std::to_string(count) + " values, expected P*B_launch = " + std::to_string(PB));
        return -1;
    }
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    CKR(hipMemcpy(be.rbuf(3), top, PB * sizeof(float), hipMemcpyHostToDevice));
    CKR(hipMemcpy(be.rbuf(4), bottom, PB * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int rt_upload_source_fluxes(rt_ctx *ctx, const float *fluxes, size_t count) {
    NN(ctx, "ctx");
    NN(fluxes, "fluxes");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_launch_buffers(be, "rt_upload_source_fluxes")) return -1;
    if (count != (size_t)be.S * be.Bl) {
        set_error("rt_upload_source_fluxes: " + std::to_string(count) + " values, expected S*B_launch = " +
                  std::to_string((size_t)be.S * be.Bl));
        return -1;
    }
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    if (grow(be.flux, be.flux_cap, count * sizeof(float) + 256)) return -1;
    if (count) CKR(hipMemcpy(be.flux, fluxes, count * sizeof(float), hipMemcpyHostToDevice));
    be.has_fluxes = true;
    return 0;
}

int rt_update_specular_materials(rt_ctx *ctx, uint32_t nprims, const float *exponent, const float *scale) {
    NN(ctx, "ctx");
    NN(exponent, "exponent");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_scene(be, "rt_update_specular_materials")) return -1;
    if (nprims != be.P) {
        set_error("rt_update_specular_materials: " + std::to_string(nprims) + " primitives, the scene has " +
                  std::to_string(be.P));
        return -1;
    }
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    if (grow(be.spm, be.spm_cap, 2 * (size_t)be.P * sizeof(float) + 256)) return -1;
    float *dexp = (float *)be.spm, *dscale = dexp + be.P;
    CKR(hipMemcpy(dexp, exponent, (size_t)be.P * sizeof(float), hipMemcpyHostToDevice));
    if (scale) CKR(hipMemcpy(dscale, scale, (size_t)be.P * sizeof(float), hipMemcpyHostToDevice));
    else CKR(hipMemset(dscale, 0, (size_t)be.P * sizeof(float)));
    be.has_specular = true;
    return 0;
}

int rt_upload_source_fluxes_cam(rt_ctx *ctx, const float *fluxes_cam, size_t count) {
    NN(ctx, "ctx");
    NN(fluxes_cam, "fluxes_cam");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_launch_buffers(be, "rt_upload_source_fluxes_cam")) return -1;
    if (count != (size_t)be.S * be.Bl) {
        set_error("rt_upload_source_fluxes_cam: " + std::to_string(count) + " values, expected S*B_launch = " +
                  std::to_string((size_t)be.S * be.Bl));
        return -1;
    }
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    if (grow(be.fcam, be.fcam_cap, count * sizeof(float) + 256)) return -1;
    if (count) CKR(hipMemcpy(be.fcam, fluxes_cam, count * sizeof(float), hipMemcpyHostToDevice));
    be.has_fcam = true;
    return 0;
}

int rt_get_radiation_specular(rt_ctx *ctx, float *out, size_t count) {
    NN(ctx, "ctx");
    NN(out, "out");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_launch_buffers(be, "rt_get_radiation_specular")) return -1;
    const size_t n = (size_t)be.S * be.P * be.Bl;
    if (count != n) {
        set_error("rt_get_radiation_specular: " + std::to_string(count) + " values, expected S*P*B_launch = " +
                  std::to_string(n));
        return -1;
    }
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    if (be.spec_live && be.spec_S == be.S) {
        if (n) CKR(hipMemcpy(out, be.spec, n * sizeof(float), hipMemcpyDeviceToHost));
    } else {
        std::fill(out, out + n, 0.0f);
    }
    return 0;
}

int rt_launch_direct_rays(rt_ctx *ctx, const rt_launch_params *p) {
    NN(ctx, "ctx");
    NN(p, "params");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_launch_buffers(be, "rt_launch_direct_rays")) return -1;
    std::vector<uint32_t> map;
    std::vector<uint8_t> active;
    if (launch_bands(be, p, map, active, "rt_launch_direct_rays")) return -1;
    if (be.S == 0 || p->launch_count == 0 || p->rays_per_primitive == 0) return 0;
    if (!be.has_fluxes) { set_error("rt_launch_direct_rays: no source fluxes (call rt_upload_source_fluxes)"); return -1; }
    std::vector<uint8_t> gcover;
    std::vector<float> ggn, ggkl;
    const int gl = glass_launch(be, map, gcover, ggn, ggkl, "rt_launch_direct_rays");
    if (gl < 0) return -1;
    const bool glass = gl == 1;
    CKR(hipSetDevice(ctx->device));
    hipStream_t st = ctx->sx.stream;
    const uint32_t P = be.P, Bl = be.Bl, S = be.S;
    const uint32_t d = (uint32_t)std::ceil(std::sqrt((double)p->rays_per_primitive));
    size_t bytes = (size_t)S * P * 2 * 4 + 2 * (size_t)P * Bl * 4 + Bl * 8 + S * sizeof(DirectJob) + 16 * 256;
    if (glass) bytes += glass_bytes(P, Bl) + (size_t)S * P * Bl * 2 * 4 + S * sizeof(GlassDirectJob) + 2 * 256;
    if (be.ws.reserve(bytes)) return -1;
    Carve cv{(char *)be.ws.base};
    float *dsum = cv.take<float>((size_t)S * P * 2);
    float *rl = cv.take<float>((size_t)P * Bl), *tl = cv.take<float>((size_t)P * Bl);
    uint32_t *dmap = cv.take<uint32_t>(Bl);
    uint8_t *dact = cv.take<uint8_t>(Bl);
    DirectJob *jobs = cv.take<DirectJob>(S);
    uint32_t *counter = cv.take<uint32_t>(16);
    int *err = reinterpret_cast<int *>(counter + 8);
    GlassTables gt;
    float *gsum = nullptr;            // [S][P][B_launch][2] strength x cover transmittance per band and face
    GlassDirectJob *gjobs = nullptr;  // [S]
    if (glass) {
        gt.cover = cv.take<uint8_t>(P);
        gt.gn = cv.take<float>((size_t)P * Bl);
        gt.gkl = cv.take<float>((size_t)P * Bl);
        gt.push = cv.take<double>((size_t)P * Bl * 3);
        gt.ctx = cv.take<GlassCtx>(1);
        gsum = cv.take<float>((size_t)S * P * Bl * 2);
        gjobs = cv.take<GlassDirectJob>(S);
    }
    if (cv.off > be.ws.cap) { set_error("rt_launch_direct_rays: workspace carve overflow"); return -1; }
    // one job per source, collimated sources first (k_direct), then the other types (k_direct_src)
    std::vector<DirectJob> hj;
    std::vector<float> coldir;
    for (int pass = 0; pass < 2; pass++)
        for (uint32_t s = 0; s < S; s++) {
            if ((be.shapes[s].type == 0u) != (pass == 0)) continue;
            DirectJob j{};
            j.sc = be.scene.d;
            j.sdir[0] = be.src_dir[s * 3]; j.sdir[1] = be.src_dir[s * 3 + 1]; j.sdir[2] = be.src_dir[s * 3 + 2];
            j.d = d;
            j.seed = p->random_seed;
            j.out = dsum + (size_t)s * P * 2;
            j.item_begin = (uint32_t)hj.size() * p->launch_count + range_begin(p->launch_offset);
            j.P = P;
            set_job_source(j, be.shapes, s);
            if (pass == 0) coldir.insert(coldir.end(), j.sdir, j.sdir + 3);
            hj.push_back(j);
        }
    const uint32_t ncol = (uint32_t)(coldir.size() / 3);
    // occlusion grids per collimated source direction (non-periodic scenes), as in rt_run_band; a launch with
    // translucent covers traces its shadow rays through them instead
    if (ctx->sx.use_grid && ncol && !glass) {
        ctx->sx.grids.density = (float)ctx->sx.cfg[0].grid_density;
        int g = build_sun_grids(ctx->sx.grids, be.scene, coldir.data(), ncol, st);
        if (g < 0) return -1;
        if (g == 0)
            for (uint32_t s = 0; s < ncol; s++) hj[s].grid = ctx->sx.grids.dev + s;
    }
    CKR(hipMemsetAsync(dsum, 0, (size_t)S * P * 2 * 4, st));
    CKR(hipMemsetAsync(err, 0, sizeof(int), st));
    CKR(hipMemcpyAsync(dmap, map.data(), Bl * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(dact, active.data(), Bl, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(jobs, hj.data(), S * sizeof(DirectJob), hipMemcpyHostToDevice, st));
    CKR(hipStreamSynchronize(st));
    k_select_bands<<<nblk((size_t)P * Bl), 256, 0, st>>>(P, be.Bg, Bl, dmap, be.rho(), be.tau(), rl, tl);
    CKR(hipGetLastError());
    // Sources are jobs over items [s*count, (s+1)*count).  find_job needs ascending item_begin, which
    // the modular offset of a sub-range breaks for S > 1: such launches go one source at a time.
    std::vector<GlassDirectJob> hg;
    if (glass) {
        if (glass_upload(gt, gcover, ggn, ggkl, P, Bl, st)) return -1;
        CKR(hipMemsetAsync(gsum, 0, (size_t)S * P * Bl * 2 * 4, st));
        for (const DirectJob &j : hj) {
            GlassDirectJob g{};
            g.dj = j;
            const size_t s = (size_t)(j.out - dsum) / ((size_t)P * 2);  // the job's source (its output slot)
            g.flux = (const float *)be.flux + s * Bl;
            g.gsum = gsum + s * P * Bl * 2;
            g.chan0 = 0;
            hg.push_back(g);
        }
        for (uint32_t s = 0; s < (p->launch_offset == 0 ? 1u : S); s++) {
            const uint32_t nj = p->launch_offset == 0 ? S : 1u;
            if (p->launch_offset != 0) hg[s].dj.item_begin = range_begin(p->launch_offset);
            CKR(hipMemcpyAsync(gjobs + s, hg.data() + s, nj * sizeof(GlassDirectJob), hipMemcpyHostToDevice, st));
            CKR(hipStreamSynchronize(st));
            if (launch_direct_glass(gjobs + s, nj, nj * p->launch_count, be.scene.h.depth, counter, gt.ctx, err,
                                    ctx->sx.ncu, st, scene_flags(be.scene, ctx->sx)))
                return -1;
        }
    } else if (p->launch_offset == 0) {
        if (launch_direct_jobs(hj.data(), jobs, S, S * p->launch_count, be.scene.h.depth, counter, err,
                               ctx->sx.stat_ptr(0), ctx->sx.cfg[0], ctx->sx.ncu, st, scene_flags(be.scene, ctx->sx)))
            return -1;
    } else {
        for (uint32_t s = 0; s < S; s++) {
            DirectJob j = hj[s];
            j.item_begin = range_begin(p->launch_offset);
            CKR(hipMemcpyAsync(jobs + s, &j, sizeof(DirectJob), hipMemcpyHostToDevice, st));
            CKR(hipStreamSynchronize(st));
            if (launch_direct_jobs(&j, jobs + s, 1, p->launch_count, be.scene.h.depth, counter, err,
                                   ctx->sx.stat_ptr(0), ctx->sx.cfg[0], ctx->sx.ncu, st, scene_flags(be.scene, ctx->sx)))
                return -1;
        }
    }
    if (launch_deposit_direct(P, Bl, Bl, S, dsum, (const float *)be.flux, rl, tl, dact, be.rbuf(0), be.rbuf(1),
                              be.rbuf(2), st, gsum))
        return -1;
    if (glass && launch_add_push((size_t)P * Bl, gt.push, be.rbuf(0), be.rbuf(1), be.rbuf(2), st)) return -1;
    if (p->specular_reflection_enabled) {
        const size_t n = (size_t)S * P * Bl;
        if (!be.spec_live || be.spec_S != S) {
            CKR(hipStreamSynchronize(st));
            if (grow(be.spec, be.spec_cap, n * sizeof(float))) return -1;
            CKR(hipMemsetAsync(be.spec, 0, n * sizeof(float), st));
            be.spec_live = true;
            be.spec_S = S;
        }
        if (glass)
            k_deposit_specular_bands<<<nblk(n), 256, 0, st>>>(P, Bl, S, gsum, (const float *)be.flux, dact,
                                                              (float *)be.spec);
        else
            k_deposit_specular<<<nblk(n), 256, 0, st>>>(P, Bl, S, dsum, (const float *)be.flux, dact,
                                                        (float *)be.spec);
        CKR(hipGetLastError());
    }
    return check_err(ctx, err);
}

int rt_launch_diffuse_rays(rt_ctx *ctx, const rt_launch_params *p) {
    NN(ctx, "ctx");
    NN(p, "params");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_launch_buffers(be, "rt_launch_diffuse_rays")) return -1;
    std::vector<uint32_t> map;
    std::vector<uint8_t> active;
    if (launch_bands(be, p, map, active, "rt_launch_diffuse_rays")) return -1;
    if (p->launch_face > 1) { set_error("rt_launch_diffuse_rays: launch_face must be 1 (top) or 0 (bottom)"); return -1; }
    if ((p->radiation_out_top == nullptr) != (p->radiation_out_bottom == nullptr)) {
        set_error("rt_launch_diffuse_rays: radiation_out_top and radiation_out_bottom must both be given or both NULL");
        return -1;
    }
    std::vector<uint8_t> gcover;
    std::vector<float> ggn, ggkl;
    const int gl = glass_launch(be, map, gcover, ggn, ggkl, "rt_launch_diffuse_rays");
    if (gl < 0) return -1;
    const bool glass = gl == 1;
    CKR(hipSetDevice(ctx->device));
    hipStream_t st = ctx->sx.stream;
    const uint32_t P = be.P, Bl = be.Bl;
    size_t PB = (size_t)P * Bl;
    if (p->radiation_out_top) {
        // launch parameters carry radiation_out (VulkanComputeBackend.cpp:1131-1134)
        CKR(hipStreamSynchronize(st));
        CKR(hipMemcpy(be.rbuf(3), p->radiation_out_top, PB * sizeof(float), hipMemcpyHostToDevice));
        CKR(hipMemcpy(be.rbuf(4), p->radiation_out_bottom, PB * sizeof(float), hipMemcpyHostToDevice));
    }
    if (p->launch_count == 0 || p->rays_per_primitive == 0) return 0;
    const uint32_t n = (uint32_t)std::ceil(std::sqrt((double)p->rays_per_primitive));
    const size_t nn = (size_t)n * n;
    // anisotropic sky (diffuse_raygen.comp:613-625): power law from the launch, Prague from the sky model
    std::vector<float> hsky;
    {
        bool any = false;
        for (uint32_t b = 0; b < Bl; b++) {
            // the sky weight only scales sky flux (diffuse_raygen.comp:616): bands without any need none
            if (!(p->diffuse_flux && p->diffuse_flux[b] > 0.f)) continue;
            const bool pl = p->diffuse_extinction && p->diffuse_extinction[b] > 0.f;
            const bool pr = be.has_sky && be.sky_B == Bl && be.sky_params[4 * b + 3] > 0.f;
            if (pl && !p->diffuse_peak_dir) {
                set_error("rt_launch_diffuse_rays: a power-law sky needs diffuse_peak_dir");
                return -1;
            }
            if (pl && !p->diffuse_dist_norm) {
                set_error("rt_launch_diffuse_rays: a power-law sky needs diffuse_dist_norm");
                return -1;
            }
            any = any || pl || pr;
        }
        if (any) {
            hsky.assign((size_t)Bl * 9, 0.f);
            for (uint32_t b = 0; b < Bl; b++) {
                float *q = hsky.data() + (size_t)b * 9;
                q[0] = p->diffuse_extinction ? p->diffuse_extinction[b] : 0.f;
                q[1] = p->diffuse_dist_norm ? p->diffuse_dist_norm[b] : 0.f;
                // peak: the launch's, else the sky model's sun direction
                for (int a = 0; a < 3; a++) q[2 + a] = p->diffuse_peak_dir ? p->diffuse_peak_dir[3 * b + a] : be.sun_dir[a];
                if (be.has_sky && be.sky_B == Bl)
                    for (int a = 0; a < 4; a++) q[5 + a] = be.sky_params[4 * b + a];
            }
        }
    }
    // a launch with covers traces and gathers in one kernel and keeps no hit table
    size_t bytes = (glass ? 0 : (size_t)P * 2 * nn * 4) + 2 * PB * 4 + Bl * 9 + sizeof(DiffuseJob) +
                   sizeof(GatherJob) + hsky.size() * 4 + 17 * 256 + (glass ? glass_bytes(P, Bl) : 0);
    if (be.ws.reserve(bytes)) return -1;
    Carve cv{(char *)be.ws.base};
    uint32_t *hits = glass ? nullptr : cv.take<uint32_t>((size_t)P * 2 * nn);
    float *rl = cv.take<float>(PB), *tl = cv.take<float>(PB);
    uint32_t *dmap = cv.take<uint32_t>(Bl);
    float *dfl = cv.take<float>(Bl);
    uint8_t *dact = cv.take<uint8_t>(Bl);
    DiffuseJob *fjob = cv.take<DiffuseJob>(1);
    GatherJob *gjob = cv.take<GatherJob>(1);
    float *dsky = hsky.empty() ? nullptr : cv.take<float>(hsky.size());
    uint32_t *counter = cv.take<uint32_t>(16);
    int *err = reinterpret_cast<int *>(counter + 8);
    GlassTables gt;
    if (glass) {
        gt.cover = cv.take<uint8_t>(P);
        gt.gn = cv.take<float>(PB);
        gt.gkl = cv.take<float>(PB);
        gt.push = cv.take<double>(PB * 3);
        gt.ctx = cv.take<GlassCtx>(1);
    }
    if (cv.off > be.ws.cap) { set_error("rt_launch_diffuse_rays: workspace carve overflow"); return -1; }
    std::vector<float> hf(Bl, 0.f);
    if (p->diffuse_flux)
        for (uint32_t b = 0; b < Bl; b++) hf[b] = p->diffuse_flux[b];
    DiffuseJob fj{};
    fj.sc = be.scene.d;
    fj.n = n;
    fj.seed = p->random_seed;
    fj.hits = hits;
    fj.item_begin = range_begin(p->launch_offset);
    fj.P = P;
    fj.faces = p->launch_face == 1 ? 1u : 2u;  // hit-table slot 0 = top launch, slot 1 = bottom launch
    GatherJob gj{};
    gj.sc = be.scene.d;
    gj.hits = hits;
    gj.n = n;
    gj.C = Bl;
    gj.B = Bl;
    gj.out_top = be.rbuf(3);
    gj.out_bot = be.rbuf(4);
    gj.dflux = dfl;
    gj.active = dact;
    gj.rho = rl;
    gj.tau = tl;
    gj.in = be.rbuf(0);
    gj.st = be.rbuf(1);
    gj.sb = be.rbuf(2);
    gj.item_begin = range_begin(p->launch_offset);
    gj.P = P;
    gj.sky = dsky;
    gj.seed = p->random_seed;
    CKR(hipMemsetAsync(err, 0, sizeof(int), st));
    if (dsky) CKR(hipMemcpyAsync(dsky, hsky.data(), hsky.size() * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(dmap, map.data(), Bl * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(dfl, hf.data(), Bl * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(dact, active.data(), Bl, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(fjob, &fj, sizeof(fj), hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(gjob, &gj, sizeof(gj), hipMemcpyHostToDevice, st));
    CKR(hipStreamSynchronize(st));
    k_select_bands<<<nblk(PB), 256, 0, st>>>(P, be.Bg, Bl, dmap, be.rho(), be.tau(), rl, tl);
    CKR(hipGetLastError());
    if (glass) {
        // rays pass through covers, which take their shares of the sky flux (diffuse_raygen.comp:484-560)
        if (glass_upload(gt, gcover, ggn, ggkl, P, Bl, st)) return -1;
        if (launch_diffuse_glass(fjob, gjob, p->launch_count, be.scene.h.depth, counter, gt.ctx, err, ctx->sx.ncu, st,
                                 scene_flags(be.scene, ctx->sx)))
            return -1;
        if (launch_add_push(PB, gt.push, be.rbuf(0), be.rbuf(1), be.rbuf(2), st)) return -1;
        return check_err(ctx, err);
    }
    if (launch_diffuse(fjob, p->launch_count, be.scene.h.depth, counter, err, ctx->sx.stat_ptr(1), ctx->sx.cfg[1],
                       ctx->sx.ncu, st, scene_flags(be.scene, ctx->sx)))
        return -1;
    if (launch_gather(gjob, 1, p->launch_count, Bl, counter, ctx->sx.cfg[2], ctx->sx.ncu, st)) return -1;
    return check_err(ctx, err);
}

int rt_get_radiation_results(rt_ctx *ctx, float *radiation_in, float *radiation_out_top, float *radiation_out_bottom,
                             float *scatter_top, float *scatter_bottom) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    if (need_launch_buffers(be, "rt_get_radiation_results")) return -1;
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    size_t bytes = (size_t)be.P * be.Bl * sizeof(float);
    float *dst[5] = {radiation_in, scatter_top, scatter_bottom, radiation_out_top, radiation_out_bottom};
    for (int k = 0; k < 5; k++)
        if (dst[k]) CKR(hipMemcpy(dst[k], be.rbuf(k), bytes, hipMemcpyDeviceToHost));
    return 0;
}

int rt_get_backend_info(rt_ctx *ctx, uint32_t *out) {
    NN(ctx, "ctx");
    NN(out, "out");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const BackendState &be = ctx->be;
    out[0] = be.has_scene ? be.P : 0;
    out[1] = be.Bg;
    out[2] = be.Bl;
    out[3] = be.S;
    return 0;
}

// ---- cameras ----------------------------------------------------------------------------------

int rt_update_sky_model(rt_ctx *ctx, uint32_t nbands, const float *sky_radiance_params,
                        const float *camera_sky_radiance, const float *sun_direction, const float *solar_disk_radiance,
                        float solar_disk_cos_angle, const float *camera_diffuse_flux,
                        const uint32_t *band_emission_flag) {
    NN(ctx, "ctx");
    NN(sun_direction, "sun_direction");
    if (nbands) {
        NN(sky_radiance_params, "sky_radiance_params");
        NN(camera_sky_radiance, "camera_sky_radiance");
        NN(solar_disk_radiance, "solar_disk_radiance");
    }
    if (nbands > 256) { set_error("rt_update_sky_model: more than 256 bands"); return -1; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    be.sky_B = nbands;
    be.sky_params.assign(sky_radiance_params, sky_radiance_params + (size_t)nbands * 4);
    be.sky_rad.assign(camera_sky_radiance, camera_sky_radiance + nbands);
    be.solar_rad.assign(solar_disk_radiance, solar_disk_radiance + nbands);
    be.cam_dflux.assign(nbands, 0.0f);
    be.emis_flag.assign(nbands, 0u);
    if (camera_diffuse_flux) be.cam_dflux.assign(camera_diffuse_flux, camera_diffuse_flux + nbands);
    if (band_emission_flag) be.emis_flag.assign(band_emission_flag, band_emission_flag + nbands);
    for (int k = 0; k < 3; k++) be.sun_dir[k] = sun_direction[k];
    be.solar_cos = solar_disk_cos_angle;
    be.has_sky = true;
    return 0;
}

namespace {

// Checks a camera launch and fills the kernel arguments that do not depend on the image buffers.
int camera_setup(rt_ctx *ctx, const rt_camera_params *p, CameraLaunch &c, const char *who) {
    BackendState &be = ctx->be;
    if (need_scene(be, who)) return -1;
    if (p->specular_reflection_enabled > 2) {
        set_error(std::string(who) + ": specular_reflection_enabled must be 0, 1 (scale 1) or 2 (specular_scale)");
        return -1;
    }
    const uint64_t rx = p->resolution[0], ry = p->resolution[1], fx = p->resolution_full[0], fy = p->resolution_full[1];
    if (fx == 0 || fy == 0 || fx * fy > (1ull << 31)) {
        set_error(std::string(who) + ": full resolution must be 1..2^31 pixels");
        return -1;
    }
    if ((uint64_t)p->pixel_offset[0] + rx > fx || (uint64_t)p->pixel_offset[1] + ry > fy) {
        set_error(std::string(who) + ": tile at offset (" + std::to_string(p->pixel_offset[0]) + ", " +
                  std::to_string(p->pixel_offset[1]) + ") of " + std::to_string(rx) + "x" + std::to_string(ry) +
                  " exceeds the full resolution " + std::to_string(fx) + "x" + std::to_string(fy));
        return -1;
    }
    if (!(p->viewplane_length > 0.0f) || !(p->fov_aspect > 0.0f)) {
        set_error(std::string(who) + ": viewplane_length and fov_aspect must be positive");
        return -1;
    }
    c = CameraLaunch{};
    c.sc = be.scene.d;
    for (int k = 0; k < 3; k++) c.pos[k] = p->position[k];
    // camera_raygen.comp:258-259: theta = elevation - pi/2, phi = pi/2 - azimuth; sin/cos evaluated in
    // double from the float angles and rounded.
    const double th = (double)(-0.5f * RT_PI_F + p->direction[0]);
    const double ph = (double)(0.5f * RT_PI_F - p->direction[1]);
    c.st = (float)std::sin(th); c.ct = (float)std::cos(th);
    c.sp = (float)std::sin(ph); c.cp = (float)std::cos(ph);
    c.focal = p->focal_length;
    c.lens = p->lens_diameter;
    c.aspect = p->fov_aspect;
    c.viewplane = p->viewplane_length;
    c.res_x = p->resolution[0]; c.res_y = p->resolution[1];
    c.full_x = p->resolution_full[0]; c.full_y = p->resolution_full[1];
    c.off_x = p->pixel_offset[0]; c.off_y = p->pixel_offset[1];
    c.spp = p->antialiasing_samples ? p->antialiasing_samples : 1u;
    c.seed = p->random_seed;
    return 0;
}

}  // namespace

int rt_launch_camera_rays(rt_ctx *ctx, const rt_camera_params *p) {
    NN(ctx, "ctx");
    NN(p, "params");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    const char *who = "rt_launch_camera_rays";
    CameraLaunch c;
    if (camera_setup(ctx, p, c, who)) return -1;
    if (be.Bl == 0 || !be.rad) {
        set_error(std::string(who) + ": radiation buffers not sized (call rt_zero_radiation_buffers first)");
        return -1;
    }
    const uint32_t B = be.Bl, S = be.S;
    if (be.has_sky && be.sky_B != B) {
        set_error(std::string(who) + ": sky model has " + std::to_string(be.sky_B) + " bands, launch has " +
                  std::to_string(B));
        return -1;
    }
    if (S && !be.has_fluxes) { set_error(std::string(who) + ": no source fluxes (call rt_upload_source_fluxes)"); return -1; }
    const bool specular = p->specular_reflection_enabled != 0 && S != 0;
    if (specular) {
        if (!be.has_specular) {
            set_error(std::string(who) + ": specular reflection needs the primitives' specular exponents (call "
                                         "rt_update_specular_materials)");
            return -1;
        }
        if (!be.has_fcam) {
            set_error(std::string(who) + ": specular reflection needs the camera's source weights (call "
                                         "rt_upload_source_fluxes_cam)");
            return -1;
        }
        if (be.spec_live && be.spec_S != S) {
            set_error(std::string(who) + ": radiation_specular holds " + std::to_string(be.spec_S) +
                      " sources, the backend has " + std::to_string(S) + " (call rt_zero_radiation_buffers)");
            return -1;
        }
    }
    CKR(hipSetDevice(ctx->device));
    hipStream_t st = ctx->sx.stream;
    const uint64_t npix = (uint64_t)c.full_x * c.full_y;
    // camera_radiation at full resolution, zeroed by the first tile (VulkanComputeBackend.cpp:1469-1486)
    if (grow(be.cam, be.cam_cap, npix * B * sizeof(float))) return -1;
    const bool first = p->pixel_offset[0] == 0 && p->pixel_offset[1] == 0;
    if (first || be.cam_pix != npix || be.cam_B != B) {
        CKR(hipMemsetAsync(be.cam, 0, npix * B * sizeof(float), st));
        be.cam_pix = npix;
        be.cam_B = B;
    }
    // per-launch tables: sky model [B] and source shapes [S]
    size_t bytes = 9 * 256 * (size_t)B + 32 * 256 + (size_t)S * 64 + 16 * 256;
    if (be.ws.reserve(bytes)) return -1;
    Carve cv{(char *)be.ws.base};
    float *d_sky = cv.take<float>((size_t)B * 4);
    float *d_srad = cv.take<float>(B), *d_solar = cv.take<float>(B), *d_dflux = cv.take<float>(B);
    uint32_t *d_emis = cv.take<uint32_t>(B);
    uint32_t *d_type = cv.take<uint32_t>(S ? S : 1);
    float *d_width = cv.take<float>(2 * (size_t)(S ? S : 1)), *d_rot = cv.take<float>(9 * (size_t)(S ? S : 1));
    int *err = cv.take<int>(64);
    std::vector<float> hsky(4 * (size_t)B, 0.f), hsrad(B, 0.f), hsolar(B, 0.f), hdflux(B, 0.f);
    std::vector<uint32_t> hemis(B, 0u), htype(S);
    std::vector<float> hwidth(2 * (size_t)S), hrot(9 * (size_t)S);
    if (be.has_sky) {
        hsky = be.sky_params; hsrad = be.sky_rad; hsolar = be.solar_rad; hdflux = be.cam_dflux; hemis = be.emis_flag;
    }
    for (uint32_t s = 0; s < S; s++) {
        htype[s] = be.shapes[s].type;
        hwidth[2 * s] = be.shapes[s].width[0]; hwidth[2 * s + 1] = be.shapes[s].width[1];
        std::memcpy(&hrot[9 * (size_t)s], be.shapes[s].rot, 9 * sizeof(float));
    }
    CKR(hipMemcpyAsync(d_sky, hsky.data(), hsky.size() * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(d_srad, hsrad.data(), B * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(d_solar, hsolar.data(), B * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(d_dflux, hdflux.data(), B * 4, hipMemcpyHostToDevice, st));
    CKR(hipMemcpyAsync(d_emis, hemis.data(), B * 4, hipMemcpyHostToDevice, st));
    if (S) {
        CKR(hipMemcpyAsync(d_type, htype.data(), S * 4, hipMemcpyHostToDevice, st));
        CKR(hipMemcpyAsync(d_width, hwidth.data(), hwidth.size() * 4, hipMemcpyHostToDevice, st));
        CKR(hipMemcpyAsync(d_rot, hrot.data(), hrot.size() * 4, hipMemcpyHostToDevice, st));
    }
    CKR(hipMemsetAsync(err, 0, sizeof(int), st));
    c.B = B;
    c.S = S;
    c.out_top = be.rbuf(3);
    c.out_bot = be.rbuf(4);
    c.flux = (const float *)be.flux;
    c.src_pos = nullptr;
    c.src_type = d_type;
    c.src_width = d_width;
    c.src_rot = d_rot;
    c.sky_params = d_sky;
    c.sky_rad = d_srad;
    c.solar_rad = d_solar;
    c.cam_dflux = d_dflux;
    c.emis = d_emis;
    const float *sun = be.has_sky ? be.sun_dir : nullptr;
    for (int k = 0; k < 3; k++) c.sun[k] = sun ? sun[k] : (k == 2 ? 1.f : 0.f);
    c.solar_cos = be.has_sky ? be.solar_cos : 2.0f;
    c.pixels = (float *)be.cam;
    // specular term: off without sums to read (radiation_specular all zero adds nothing)
    if (specular && be.spec_live) {
        c.spec = (const float *)be.spec;
        c.fcam = (const float *)be.fcam;
        c.spec_exp = (const float *)be.spm;
        c.spec_scale = (const float *)be.spm + be.P;
        c.spec_P = be.P;
        c.spec_mode = p->specular_reflection_enabled;
    }
    // source positions travel with the shapes table
    float *d_pos = cv.take<float>(3 * (size_t)(S ? S : 1));
    if (cv.off > be.ws.cap) { set_error(std::string(who) + ": workspace carve overflow"); return -1; }
    if (S) CKR(hipMemcpyAsync(d_pos, be.src_dir.data(), (size_t)S * 12, hipMemcpyHostToDevice, st));
    c.src_pos = d_pos;
    CKR(hipStreamSynchronize(st));
    hipEvent_t ea = nullptr, eb = nullptr;
    if (ctx->sx.timing) { CKR(hipEventCreate(&ea)); CKR(hipEventCreate(&eb)); CKR(hipEventRecord(ea, st)); }
    if (launch_camera(c, be.scene.h.depth, false, err, ctx->sx.stat_ptr(3), st, scene_flags(be.scene, ctx->sx))) return -1;
    if (eb) {
        CKR(hipEventRecord(eb, st));
        CKR(hipEventSynchronize(eb));
        float ms = 0.f;
        CKR(hipEventElapsedTime(&ms, ea, eb));
        (void)hipEventDestroy(ea);
        (void)hipEventDestroy(eb);
        // camera kernel time reported in the first slot, as for castRays
        ctx->sx.timing_info = rt_timing{};
        ctx->sx.timing_info.direct_ms = ms;
        ctx->sx.timing_info.direct_rays = (uint64_t)c.res_x * c.res_y * c.spp;
        ctx->sx.timing_info.direct_launches = 1;
    }
    return check_err(ctx, err);
}

int rt_zero_camera_pixel_buffers(rt_ctx *ctx, uint32_t width, uint32_t height) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    const uint64_t n = (uint64_t)width * height;
    if (n == 0) return 0;
    if (n > (1ull << 31)) { set_error("rt_zero_camera_pixel_buffers: more than 2^31 pixels"); return -1; }
    CKR(hipSetDevice(ctx->device));
    if (grow(be.lab, be.lab_cap, n * 8)) return -1;
    CKR(hipMemsetAsync(be.lab, 0, n * 8, ctx->sx.stream));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    be.lab_pix = n;
    return 0;
}

int rt_launch_pixel_label_rays(rt_ctx *ctx, const rt_camera_params *p) {
    NN(ctx, "ctx");
    NN(p, "params");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    const char *who = "rt_launch_pixel_label_rays";
    CameraLaunch c;
    if (camera_setup(ctx, p, c, who)) return -1;
    const uint64_t npix = (uint64_t)c.full_x * c.full_y;
    if (be.lab_pix != npix) {
        set_error(std::string(who) + ": label buffers hold " + std::to_string(be.lab_pix) + " pixels, the image has " +
                  std::to_string(npix) + " (call rt_zero_camera_pixel_buffers with the full resolution)");
        return -1;
    }
    CKR(hipSetDevice(ctx->device));
    hipStream_t st = ctx->sx.stream;
    if (be.ws.reserve(256)) return -1;
    int *err = (int *)be.ws.base;
    CKR(hipMemsetAsync(err, 0, sizeof(int), st));
    c.labels = (uint32_t *)be.lab;
    c.depths = (float *)((uint32_t *)be.lab + npix);
    if (launch_camera(c, be.scene.h.depth, true, err, nullptr, st, scene_flags(be.scene, ctx->sx))) return -1;
    return check_err(ctx, err);
}

int rt_get_camera_results(rt_ctx *ctx, float *pixel_data, uint32_t *pixel_labels, float *pixel_depths, uint32_t width,
                          uint32_t height) {
    NN(ctx, "ctx");
    std::lock_guard<std::mutex> lk(ctx->mu);
    BackendState &be = ctx->be;
    const uint64_t n = (uint64_t)width * height;
    if (n == 0) return 0;
    CKR(hipSetDevice(ctx->device));
    CKR(hipStreamSynchronize(ctx->sx.stream));
    // camera_radiation is band-major on the device; the caller receives [pixel][band]
    // (VulkanComputeBackend.cpp:1891-1911).  Buffers of another size read as zeros.
    if (pixel_data) {
        const uint32_t B = be.cam_B;
        if (be.cam_pix == n && B && B == be.Bl) {
            std::vector<float> bm(n * B);
            CKR(hipMemcpy(bm.data(), be.cam, n * B * sizeof(float), hipMemcpyDeviceToHost));
            for (uint32_t b = 0; b < B; b++)
                for (uint64_t q = 0; q < n; q++) pixel_data[q * B + b] = bm[(size_t)b * n + q];
        } else {
            std::memset(pixel_data, 0, n * be.Bl * sizeof(float));
        }
    }
    const bool lab = be.lab_pix == n;
    if (pixel_labels) {
        if (lab) CKR(hipMemcpy(pixel_labels, be.lab, n * 4, hipMemcpyDeviceToHost));
        else std::memset(pixel_labels, 0, n * 4);
    }
    if (pixel_depths) {
        if (lab) CKR(hipMemcpy(pixel_depths, (uint32_t *)be.lab + n, n * 4, hipMemcpyDeviceToHost));
        else std::memset(pixel_depths, 0, n * 4);
    }
    return 0;
}

}  // extern "C"
