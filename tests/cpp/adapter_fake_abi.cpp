// adapter_fake_abi.cpp -- recording test double of the C ABI entry points the RayTracingBackend adapter
// calls (see adapter_fake_abi.h).  Results are deterministic functions of the index so the adapter's
// copies can be checked.
#include "adapter_fake_abi.h"

#include <cstring>

namespace {
FakeAbi g_fake;
std::string g_err;
struct rt_ctx_fake {
    int device;
};

int status(const char *name) {
    g_fake.calls.push_back(name);
    if (g_fake.fail_next) {
        g_fake.fail_next = 0;
        g_err = g_fake.fail_msg;
        return -1;
    }
    return 0;
}

template <typename T>
std::vector<T> copy(const T *p, size_t n) {
    return p ? std::vector<T>(p, p + n) : std::vector<T>();
}

FakeLaunch record(const rt_launch_params *p) {
    FakeLaunch l;
    l.p = *p;
    const size_t B = p->num_bands_launch;
    l.band_map = copy(p->band_map, B);
    l.band_active = copy(p->band_active, B);
    l.diffuse_flux = copy(p->diffuse_flux, B);
    l.extinction = copy(p->diffuse_extinction, B);
    l.dist_norm = copy(p->diffuse_dist_norm, B);
    l.peak_dir = copy(p->diffuse_peak_dir, 3 * B);
    l.has_out = p->radiation_out_top != nullptr && p->radiation_out_bottom != nullptr;
    return l;
}
}  // namespace

FakeAbi &fake() { return g_fake; }

extern "C" {

const char *rt_last_error(void) { return g_err.c_str(); }

int rt_create(int device, rt_ctx **out) {
    g_fake.device = device;
    *out = reinterpret_cast<rt_ctx *>(new rt_ctx_fake{device});
    return status("rt_create");
}

int rt_destroy(rt_ctx *ctx) {
    delete reinterpret_cast<rt_ctx_fake *>(ctx);
    return status("rt_destroy");
}

const char *rt_backend_name(const rt_ctx *) { return "fake"; }

int rt_query_memory(rt_ctx *, size_t *free_bytes, size_t *total_bytes) {
    *free_bytes = 1u << 20;
    *total_bytes = 2u << 20;
    return status("rt_query_memory");
}

int rt_update_geometry(rt_ctx *, uint32_t nprims, const float *transforms, const uint32_t *types, const uint32_t *uuids,
                       const uint32_t *twosided, int periodic_x, int periodic_y) {
    g_fake.P = nprims;
    g_fake.transforms = copy(transforms, 16 * (size_t)nprims);
    g_fake.types = copy(types, nprims);
    g_fake.uuids = copy(uuids, nprims);
    g_fake.twosided = copy(twosided, nprims);
    g_fake.px = periodic_x;
    g_fake.py = periodic_y;
    return status("rt_update_geometry");
}

int rt_update_texture_masks(rt_ctx *, uint32_t nprims, uint32_t nmasks, const uint8_t *texels, const uint32_t *sizes,
                            const int32_t *mask_ids, const float *uv, const int32_t *uv_ids) {
    g_fake.nmasks = nmasks;
    g_fake.sizes = copy(sizes, 2 * (size_t)nmasks);
    size_t n = 0;
    for (uint32_t m = 0; m < nmasks; m++) n += (size_t)sizes[2 * m] * sizes[2 * m + 1];
    g_fake.texels = copy(texels, n);
    g_fake.mask_ids = copy(mask_ids, nprims);
    g_fake.uv = copy(uv, 8 * (size_t)nprims);
    g_fake.uv_ids = copy(uv_ids, nprims);
    return status("rt_update_texture_masks");
}

int rt_build_acceleration_structure(rt_ctx *) { return status("rt_build_acceleration_structure"); }

int rt_update_materials(rt_ctx *, uint32_t nsources, uint32_t nprims, uint32_t nbands_global, const float *rho,
                        const float *tau) {
    g_fake.mat_S = nsources;
    g_fake.mat_P = nprims;
    g_fake.mat_B = nbands_global;
    const size_t n = (size_t)nsources * nprims * nbands_global;
    g_fake.rho = copy(rho, n);
    g_fake.tau = copy(tau, n);
    return status("rt_update_materials");
}

int rt_update_sources(rt_ctx *, uint32_t nsources, const uint32_t *types, const float *positions, const float *widths,
                      const float *rotations) {
    g_fake.src_types = copy(types, nsources);
    g_fake.src_pos = copy(positions, 3 * (size_t)nsources);
    g_fake.src_width = copy(widths, 2 * (size_t)nsources);
    g_fake.src_rot = copy(rotations, 3 * (size_t)nsources);
    return status("rt_update_sources");
}

int rt_zero_radiation_buffers(rt_ctx *, uint32_t launch_band_count) {
    g_fake.launch_bands = launch_band_count;
    return status("rt_zero_radiation_buffers");
}

int rt_zero_scatter_buffers(rt_ctx *) { return status("rt_zero_scatter_buffers"); }
int rt_copy_scatter_to_radiation(rt_ctx *) { return status("rt_copy_scatter_to_radiation"); }

int rt_upload_radiation_out(rt_ctx *, const float *top, const float *bottom, size_t count) {
    g_fake.out_top = copy(top, count);
    g_fake.out_bottom = copy(bottom, count);
    return status("rt_upload_radiation_out");
}

int rt_upload_source_fluxes(rt_ctx *, const float *fluxes, size_t count) {
    g_fake.fluxes = copy(fluxes, count);
    return status("rt_upload_source_fluxes");
}

int rt_update_glass_materials(rt_ctx *, uint32_t nprims, uint32_t nbands_global, const uint8_t *is_glass,
                              const float *glass_n, const float *glass_KL) {
    const size_t n = (size_t)nprims * nbands_global;
    g_fake.is_glass = copy(is_glass, n);
    g_fake.glass_n = copy(glass_n, n);
    g_fake.glass_KL = copy(glass_KL, n);
    return status("rt_update_glass_materials");
}

int rt_update_specular_materials(rt_ctx *, uint32_t nprims, const float *exponent, const float *scale) {
    g_fake.spec_exp = copy(exponent, nprims);
    g_fake.spec_scale = copy(scale, nprims);
    return status("rt_update_specular_materials");
}

int rt_upload_source_fluxes_cam(rt_ctx *, const float *fluxes_cam, size_t count) {
    g_fake.fluxes_cam = copy(fluxes_cam, count);
    return status("rt_upload_source_fluxes_cam");
}

int rt_launch_direct_rays(rt_ctx *, const rt_launch_params *params) {
    g_fake.direct.push_back(record(params));
    return status("rt_launch_direct_rays");
}

int rt_launch_diffuse_rays(rt_ctx *, const rt_launch_params *params) {
    g_fake.diffuse.push_back(record(params));
    return status("rt_launch_diffuse_rays");
}

// buffer k of [P][B]: value = 1000 k + index
int rt_get_radiation_results(rt_ctx *, float *radiation_in, float *radiation_out_top, float *radiation_out_bottom,
                             float *scatter_top, float *scatter_bottom) {
    float *bufs[5] = {radiation_in, radiation_out_top, radiation_out_bottom, scatter_top, scatter_bottom};
    const size_t n = (size_t)g_fake.P * g_fake.launch_bands;
    for (int k = 0; k < 5; k++)
        if (bufs[k])
            for (size_t i = 0; i < n; i++) bufs[k][i] = 1000.f * (float)k + (float)i;
    return status("rt_get_radiation_results");
}

int rt_update_sky_model(rt_ctx *, uint32_t, const float *, const float *, const float *, const float *, float,
                        const float *, const uint32_t *) {
    return status("rt_update_sky_model");
}

int rt_launch_camera_rays(rt_ctx *, const rt_camera_params *params) {
    g_fake.camera.push_back(*params);
    return status("rt_launch_camera_rays");
}

int rt_zero_camera_pixel_buffers(rt_ctx *, uint32_t, uint32_t) { return status("rt_zero_camera_pixel_buffers"); }

int rt_launch_pixel_label_rays(rt_ctx *, const rt_camera_params *params) {
    g_fake.label.push_back(*params);
    return status("rt_launch_pixel_label_rays");
}

int rt_get_camera_results(rt_ctx *, float *pixel_data, uint32_t *pixel_labels, float *pixel_depths, uint32_t width,
                          uint32_t height) {
    const size_t n = (size_t)width * height;
    for (size_t i = 0; i < n; i++) {
        if (pixel_labels) pixel_labels[i] = (uint32_t)i + 1u;
        if (pixel_depths) pixel_depths[i] = 0.5f * (float)i;
        if (pixel_data)
            for (uint32_t b = 0; b < g_fake.launch_bands; b++) pixel_data[i * g_fake.launch_bands + b] = (float)b;
    }
    return status("rt_get_camera_results");
}

}  // extern "C"
