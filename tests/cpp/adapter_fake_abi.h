// adapter_fake_abi.h -- a recording test double of the call-by-call C ABI (include/helios_rt.h), so the
// helios::RayTracingBackend adapter can run on a CPU: every entry point the adapter calls stores a copy of
// its arguments here and returns success (or the status in `fail_next`).  Test code only.
#ifndef ADAPTER_FAKE_ABI_H
#define ADAPTER_FAKE_ABI_H

#include <cstdint>
#include <string>
#include <vector>

#include "helios_rt.h"

struct FakeLaunch {
    rt_launch_params p{};
    std::vector<uint32_t> band_map;
    std::vector<uint8_t> band_active;
    std::vector<float> diffuse_flux, peak_dir, extinction, dist_norm;
    bool has_out = false;
};

struct FakeAbi {
    std::vector<std::string> calls;  // entry point names in call order
    int device = -1;
    // rt_update_geometry
    uint32_t P = 0;
    std::vector<float> transforms;
    std::vector<uint32_t> types, uuids, twosided;
    int px = 0, py = 0;
    // rt_update_texture_masks
    uint32_t nmasks = 0;
    std::vector<uint8_t> texels;
    std::vector<uint32_t> sizes;
    std::vector<int32_t> mask_ids, uv_ids;
    std::vector<float> uv;
    // rt_update_materials
    uint32_t mat_S = 0, mat_P = 0, mat_B = 0;
    std::vector<float> rho, tau;
    // rt_update_sources
    std::vector<uint32_t> src_types;
    std::vector<float> src_pos, src_width, src_rot;
    // rt_zero_radiation_buffers
    uint32_t launch_bands = 0;
    // launches
    std::vector<FakeLaunch> direct, diffuse;
    std::vector<rt_camera_params> camera, label;
    // rt_upload_source_fluxes / rt_upload_radiation_out
    std::vector<float> fluxes, out_top, out_bottom;
    // rt_update_specular_materials / rt_upload_source_fluxes_cam
    std::vector<float> spec_exp, spec_scale, fluxes_cam;
    // rt_update_glass_materials
    std::vector<uint8_t> is_glass;
    std::vector<float> glass_n, glass_KL;
    int fail_next = 0;  // a nonzero status makes the next call fail with this message
    std::string fail_msg = "injected failure";
};

FakeAbi &fake();

#endif
