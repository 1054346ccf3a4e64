// HipRayTracingBackend.cpp -- helios::RayTracingBackend on the gfx950 C ABI.  Builds inside a Helios tree
// (plugins/radiation/src/backends/), linked with libhelios_rt.so.
#include "HipRayTracingBackend.h"

#include <iostream>
#include <string>

namespace helios {

    namespace {
        [[noreturn]] void unsupported(const char *who, const std::string &what) {
            helios_runtime_error(std::string("ERROR (HipRayTracingBackend::") + who + "): " + what + " is not supported by the HIP backend.");
            throw std::runtime_error(what);  // helios_runtime_error throws; this keeps [[noreturn]] honest
        }
    } // namespace

    HipRayTracingBackend::HipRayTracingBackend(int device) : device_(device) {
    }

    HipRayTracingBackend::~HipRayTracingBackend() {
        if (ctx_) {
            rt_destroy(ctx_);
            ctx_ = nullptr;
        }
    }

    void HipRayTracingBackend::check(int rc, const char *who) const {
        if (rc != 0) {
            helios_runtime_error(std::string("ERROR (HipRayTracingBackend::") + who + "): " + rt_last_error());
        }
    }

    void HipRayTracingBackend::initialize() {
        if (!ctx_) {
            check(rt_create(device_, &ctx_), "initialize");
        }
    }

    void HipRayTracingBackend::shutdown() {
        if (ctx_) {
            check(rt_destroy(ctx_), "shutdown");
            ctx_ = nullptr;
        }
    }

    void HipRayTracingBackend::updateGeometry(const RayTracingGeometry &geometry) {
        validateGeometryBeforeUpload(geometry);
        const size_t P = geometry.primitive_count;
        std::vector<uint32_t> types(geometry.primitive_types.begin(), geometry.primitive_types.end());
        std::vector<uint32_t> uuids(geometry.primitive_UUIDs.begin(), geometry.primitive_UUIDs.end());
        std::vector<uint32_t> twosided(P);
        for (size_t i = 0; i < P; i++) {
            twosided[i] = static_cast<uint32_t>(static_cast<unsigned char>(geometry.twosided_flags.at(i)));
        }
        check(rt_update_geometry(ctx_, static_cast<uint32_t>(P), geometry.transform_matrices.data(), types.data(), uuids.empty() ? nullptr : uuids.data(), twosided.data(),
                                 geometry.periodic_flag.x != 0.f ? 1 : 0, geometry.periodic_flag.y != 0.f ? 1 : 0),
              "updateGeometry");
        primitive_count_ = P;
        bool any_mask = false;
        for (int id: geometry.mask_IDs) {
            any_mask = any_mask || id >= 0;
        }
        if (!any_mask) {
            return;
        }
        // texture masks: bool texels -> bytes, compact uv_data (4 entries per primitive with uv_IDs >= 0, in
        // primitive order) -> [P][4][2] (VulkanComputeBackend.cpp:439-470)
        const size_t M = geometry.mask_sizes.size();
        std::vector<uint32_t> sizes(2 * M);
        size_t ntex = 0;
        for (size_t m = 0; m < M; m++) {
            sizes[2 * m] = static_cast<uint32_t>(geometry.mask_sizes[m].x);
            sizes[2 * m + 1] = static_cast<uint32_t>(geometry.mask_sizes[m].y);
            ntex += size_t(geometry.mask_sizes[m].x) * size_t(geometry.mask_sizes[m].y);
        }
        if (geometry.mask_data.size() < ntex) {
            helios_runtime_error("ERROR (HipRayTracingBackend::updateGeometry): mask_data holds " + std::to_string(geometry.mask_data.size()) + " texels, the masks need " + std::to_string(ntex) + ".");
        }
        std::vector<uint8_t> texels(ntex);
        for (size_t k = 0; k < ntex; k++) {
            texels[k] = geometry.mask_data[k] ? 1 : 0;
        }
        std::vector<int32_t> mask_ids(P, -1), uv_ids(P, -1);
        std::vector<float> uv(8 * P, 0.f);
        size_t next = 0;
        for (size_t p = 0; p < P; p++) {
            if (p < geometry.mask_IDs.size()) {
                mask_ids[p] = geometry.mask_IDs[p];
            }
            if (p < geometry.uv_IDs.size() && geometry.uv_IDs[p] >= 0) {
                uv_ids[p] = geometry.uv_IDs[p];
                for (int v = 0; v < 4 && next < geometry.uv_data.size(); v++, next++) {
                    uv[8 * p + 2 * v] = geometry.uv_data[next].x;
                    uv[8 * p + 2 * v + 1] = geometry.uv_data[next].y;
                }
            }
        }
        check(rt_update_texture_masks(ctx_, static_cast<uint32_t>(P), static_cast<uint32_t>(M), texels.data(), sizes.data(), mask_ids.data(), uv.data(), uv_ids.data()), "updateGeometry");
    }

    void HipRayTracingBackend::buildAccelerationStructure() {
        check(rt_build_acceleration_structure(ctx_), "buildAccelerationStructure");
    }

    void HipRayTracingBackend::updateMaterials(const RayTracingMaterial &materials) {
        if (primitive_count_ == 0) {
            return; // no geometry yet (VulkanComputeBackend.cpp:592-598)
        }
        check(rt_update_materials(ctx_, static_cast<uint32_t>(materials.num_sources), static_cast<uint32_t>(materials.num_primitives), static_cast<uint32_t>(materials.num_bands), materials.reflectivity.data(),
                                  materials.transmissivity.data()),
              "updateMaterials");
        // translucent covers: the source-0 slice of is_glass / glass_n / glass_KL (VulkanComputeBackend.cpp:656-690)
        const size_t PB = materials.num_primitives * materials.num_bands;
        bool any_glass = false;
        for (char g: materials.is_glass) {
            any_glass = any_glass || g != 0;
        }
        if (any_glass) {
            if (materials.is_glass.size() < PB || materials.glass_n.size() < PB || (!materials.glass_KL.empty() && materials.glass_KL.size() < PB)) {
                helios_runtime_error("ERROR (HipRayTracingBackend::updateMaterials): is_glass / glass_n / glass_KL need Nsources x Nprims x Nbands entries.");
            }
            std::vector<uint8_t> is_glass(materials.is_glass.begin(), materials.is_glass.begin() + static_cast<std::ptrdiff_t>(PB));
            check(rt_update_glass_materials(ctx_, static_cast<uint32_t>(materials.num_primitives), static_cast<uint32_t>(materials.num_bands), is_glass.data(), materials.glass_n.data(),
                                            materials.glass_KL.empty() ? nullptr : materials.glass_KL.data()),
                  "updateMaterials");
        } else {
            check(rt_update_glass_materials(ctx_, static_cast<uint32_t>(materials.num_primitives), static_cast<uint32_t>(materials.num_bands), nullptr, nullptr, nullptr), "updateMaterials");
        }
        // specular exponent and scale per primitive (VulkanComputeBackend.cpp:631-655)
        if (!materials.specular_exponent.empty()) {
            const size_t P = materials.num_primitives;
            if (materials.specular_exponent.size() != P || (!materials.specular_scale.empty() && materials.specular_scale.size() != P)) {
                helios_runtime_error("ERROR (HipRayTracingBackend::updateMaterials): specular_exponent / specular_scale need " + std::to_string(P) + " entries (Nprimitives).");
            }
            check(rt_update_specular_materials(ctx_, static_cast<uint32_t>(P), materials.specular_exponent.data(), materials.specular_scale.empty() ? nullptr : materials.specular_scale.data()), "updateMaterials");
        }
    }

    void HipRayTracingBackend::updateSources(const std::vector<RayTracingSource> &sources) {
        std::vector<uint32_t> types(sources.size());
        std::vector<float> positions(3 * sources.size()), widths(2 * sources.size()), rotations(3 * sources.size());
        for (size_t s = 0; s < sources.size(); s++) {
            types[s] = sources[s].type;
            positions[3 * s] = sources[s].position.x;
            positions[3 * s + 1] = sources[s].position.y;
            positions[3 * s + 2] = sources[s].position.z;
            widths[2 * s] = sources[s].width.x;
            widths[2 * s + 1] = sources[s].width.y;
            rotations[3 * s] = sources[s].rotation.x;
            rotations[3 * s + 1] = sources[s].rotation.y;
            rotations[3 * s + 2] = sources[s].rotation.z;
        }
        check(rt_update_sources(ctx_, static_cast<uint32_t>(sources.size()), types.data(), positions.data(), widths.data(),
                                rotations.data()),
              "updateSources");
    }

    void HipRayTracingBackend::updateDiffuseRadiation(const std::vector<float> &, const std::vector<float> &, const std::vector<helios::vec3> &, const std::vector<float> &, const std::vector<float> &) {
        // the diffuse distribution travels with each launch (RayTracingLaunchParams), as in the Vulkan
        // backend (VulkanComputeBackend.cpp:752-757)
    }

    void HipRayTracingBackend::updateSkyModel(const std::vector<helios::vec4> &sky_radiance_params, const std::vector<float> &camera_sky_radiance, const helios::vec3 &sun_direction, const std::vector<float> &solar_disk_radiance,
                                              float solar_disk_cos_angle, const std::vector<float> &camera_diffuse_flux, const std::vector<uint32_t> &band_emission_flag) {
        const size_t B = sky_radiance_params.size();
        if (camera_sky_radiance.size() != B || solar_disk_radiance.size() != B) {
            helios_runtime_error("ERROR (HipRayTracingBackend::updateSkyModel): per-band vectors differ in length.");
        }
        std::vector<float> prm(4 * B);
        for (size_t b = 0; b < B; b++) {
            prm[4 * b] = sky_radiance_params[b].x;
            prm[4 * b + 1] = sky_radiance_params[b].y;
            prm[4 * b + 2] = sky_radiance_params[b].z;
            prm[4 * b + 3] = sky_radiance_params[b].w;
        }
        const float sun[3] = {sun_direction.x, sun_direction.y, sun_direction.z};
        check(rt_update_sky_model(ctx_, static_cast<uint32_t>(B), prm.data(), camera_sky_radiance.data(), sun, solar_disk_radiance.data(), solar_disk_cos_angle,
                                  camera_diffuse_flux.size() == B ? camera_diffuse_flux.data() : nullptr, band_emission_flag.size() == B ? band_emission_flag.data() : nullptr),
              "updateSkyModel");
    }

    void HipRayTracingBackend::bandTable(const RayTracingLaunchParams &params, const char *who) {
        const size_t Bl = params.num_bands_launch;
        std::vector<uint32_t> flagged;
        for (size_t g = 0; g < params.band_launch_flag.size(); g++) {
            if (params.band_launch_flag[g]) {
                flagged.push_back(static_cast<uint32_t>(g));
            }
        }
        if (params.band_launch_flag.empty()) {
            // no flags: identity map, as VulkanComputeBackend.cpp:1265-1270
            band_map_.resize(Bl);
            for (size_t b = 0; b < Bl; b++) {
                band_map_[b] = static_cast<uint32_t>(b);
            }
            band_active_.assign(Bl, 1);
        } else if (flagged.size() == Bl) {
            band_map_ = flagged;
            band_active_.assign(Bl, 1);
        } else if (flagged.size() < Bl && band_map_.size() == Bl) {
            // a scattering iteration with retired bands (RadiationModel.cpp:4255-4277): each launch band
            // keeps its slot and stops taking energy once its global flag is cleared
            for (size_t b = 0; b < Bl; b++) {
                band_active_[b] = params.band_launch_flag.at(band_map_[b]) ? 1 : 0;
            }
        } else {
            helios_runtime_error(std::string("ERROR (HipRayTracingBackend::") + who + "): band_launch_flag selects " + std::to_string(flagged.size()) + " bands for a launch of " + std::to_string(Bl) + ".");
        }
    }

    rt_launch_params HipRayTracingBackend::cParams(const RayTracingLaunchParams &params) const {
        rt_launch_params p{};
        p.launch_offset = params.launch_offset;
        p.launch_count = params.launch_count;
        p.rays_per_primitive = params.rays_per_primitive;
        p.random_seed = params.random_seed;
        p.num_bands_global = params.num_bands_global;
        p.num_bands_launch = params.num_bands_launch;
        p.band_map = band_map_.data();
        p.band_active = band_active_.data();
        p.launch_face = params.launch_face;
        p.specular_reflection_enabled = params.specular_reflection_enabled;
        p.diffuse_flux = params.diffuse_flux.empty() ? nullptr : params.diffuse_flux.data();
        if (!params.radiation_out_top.empty() && !params.radiation_out_bottom.empty()) {
            p.radiation_out_top = params.radiation_out_top.data();
            p.radiation_out_bottom = params.radiation_out_bottom.data();
        }
        return p;
    }

    void HipRayTracingBackend::launchDirectRays(const RayTracingLaunchParams &params) {
        if (primitive_count_ == 0) {
            return;
        }
        bandTable(params, "launchDirectRays");
        rt_launch_params p = cParams(params);
        check(rt_launch_direct_rays(ctx_, &p), "launchDirectRays");
    }

    void HipRayTracingBackend::launchDiffuseRays(const RayTracingLaunchParams &params) {
        if (primitive_count_ == 0) {
            return;
        }
        bandTable(params, "launchDiffuseRays");
        rt_launch_params p = cParams(params);
        // anisotropic sky: power law from the launch, Prague from updateSkyModel (diffuse_raygen.comp:613-625)
        const size_t Bl = params.num_bands_launch;
        std::vector<float> peak;
        if (params.diffuse_extinction.size() >= Bl) {
            p.diffuse_extinction = params.diffuse_extinction.data();
        }
        if (params.diffuse_dist_norm.size() >= Bl) {
            p.diffuse_dist_norm = params.diffuse_dist_norm.data();
        }
        if (params.diffuse_peak_dir.size() >= Bl) {
            peak.resize(3 * Bl);
            for (size_t b = 0; b < Bl; b++) {
                peak[3 * b] = params.diffuse_peak_dir[b].x;
                peak[3 * b + 1] = params.diffuse_peak_dir[b].y;
                peak[3 * b + 2] = params.diffuse_peak_dir[b].z;
            }
            p.diffuse_peak_dir = peak.data();
        }
        check(rt_launch_diffuse_rays(ctx_, &p), "launchDiffuseRays");
    }

    rt_camera_params HipRayTracingBackend::cameraParams(const RayTracingLaunchParams &params) {
        rt_camera_params c{};
        c.position[0] = params.camera_position.x;
        c.position[1] = params.camera_position.y;
        c.position[2] = params.camera_position.z;
        c.direction[0] = params.camera_direction.x;
        c.direction[1] = params.camera_direction.y;
        c.focal_length = params.camera_focal_length;
        c.lens_diameter = params.camera_lens_diameter;
        c.fov_aspect = params.camera_fov_aspect;
        c.viewplane_length = params.camera_viewplane_length;
        c.resolution[0] = static_cast<uint32_t>(params.camera_resolution.x);
        c.resolution[1] = static_cast<uint32_t>(params.camera_resolution.y);
        c.resolution_full[0] = static_cast<uint32_t>(params.camera_resolution_full.x);
        c.resolution_full[1] = static_cast<uint32_t>(params.camera_resolution_full.y);
        c.pixel_offset[0] = static_cast<uint32_t>(params.camera_pixel_offset.x);
        c.pixel_offset[1] = static_cast<uint32_t>(params.camera_pixel_offset.y);
        c.antialiasing_samples = params.antialiasing_samples;
        c.random_seed = params.random_seed;
        c.specular_reflection_enabled = params.specular_reflection_enabled;
        return c;
    }

    void HipRayTracingBackend::launchCameraRays(const RayTracingLaunchParams &params) {
        if (primitive_count_ == 0) {
            return;
        }
        rt_camera_params c = cameraParams(params);
        check(rt_launch_camera_rays(ctx_, &c), "launchCameraRays");
    }

    void HipRayTracingBackend::launchPixelLabelRays(const RayTracingLaunchParams &params) {
        if (primitive_count_ == 0) {
            return;
        }
        rt_camera_params c = cameraParams(params);
        c.specular_reflection_enabled = 0;
        check(rt_launch_pixel_label_rays(ctx_, &c), "launchPixelLabelRays");
    }

    void HipRayTracingBackend::getRadiationResults(RayTracingResults &results) {
        if (primitive_count_ == 0 || launch_band_count_ == 0) {
            return; // nothing to download (VulkanComputeBackend.cpp:1766-1768)
        }
        const size_t n = primitive_count_ * launch_band_count_;
        results.radiation_in.resize(n);
        results.radiation_out_top.resize(n);
        results.radiation_out_bottom.resize(n);
        results.scatter_buff_top.resize(n);
        results.scatter_buff_bottom.resize(n);
        results.radiation_specular.clear();
        check(rt_get_radiation_results(ctx_, results.radiation_in.data(), results.radiation_out_top.data(), results.radiation_out_bottom.data(), results.scatter_buff_top.data(), results.scatter_buff_bottom.data()),
              "getRadiationResults");
        // camera scatter = the scatter buffers (camera spectral response taken as uniform), which
        // RadiationModel accumulates and uploads as radiation_out before camera launches
        // (VulkanComputeBackend.cpp:1871-1876, RadiationModel.cpp:3992-4012, 4359-4366)
        results.scatter_buff_top_cam = results.scatter_buff_top;
        results.scatter_buff_bottom_cam = results.scatter_buff_bottom;
        results.num_primitives = primitive_count_;
        results.num_bands = launch_band_count_;
    }

    void HipRayTracingBackend::getCameraResults(std::vector<float> &pixel_data, std::vector<uint> &pixel_labels, std::vector<float> &pixel_depths, uint, const helios::int2 &resolution) {
        const size_t n = size_t(resolution.x) * size_t(resolution.y);
        if (n == 0) {
            return;
        }
        pixel_data.assign(n * launch_band_count_, 0.f);
        pixel_labels.assign(n, 0u);
        pixel_depths.assign(n, 0.f);
        if (primitive_count_ == 0) {
            return;
        }
        check(rt_get_camera_results(ctx_, pixel_data.data(), pixel_labels.data(), pixel_depths.data(), static_cast<uint32_t>(resolution.x), static_cast<uint32_t>(resolution.y)), "getCameraResults");
    }

    void HipRayTracingBackend::zeroRadiationBuffers(size_t launch_band_count) {
        launch_band_count_ = launch_band_count;
        if (primitive_count_ == 0 || launch_band_count == 0) {
            return;
        }
        band_map_.clear();
        band_active_.clear();
        check(rt_zero_radiation_buffers(ctx_, static_cast<uint32_t>(launch_band_count)), "zeroRadiationBuffers");
    }

    void HipRayTracingBackend::zeroScatterBuffers() {
        if (primitive_count_ == 0 || launch_band_count_ == 0) {
            return;
        }
        check(rt_zero_scatter_buffers(ctx_), "zeroScatterBuffers");
    }

    void HipRayTracingBackend::zeroCameraPixelBuffers(const helios::int2 &resolution) {
        check(rt_zero_camera_pixel_buffers(ctx_, static_cast<uint32_t>(resolution.x), static_cast<uint32_t>(resolution.y)), "zeroCameraPixelBuffers");
    }

    void HipRayTracingBackend::copyScatterToRadiation() {
        if (primitive_count_ == 0 || launch_band_count_ == 0) {
            return;
        }
        check(rt_copy_scatter_to_radiation(ctx_), "copyScatterToRadiation");
    }

    void HipRayTracingBackend::uploadRadiationOut(const std::vector<float> &radiation_out_top, const std::vector<float> &radiation_out_bottom) {
        if (radiation_out_top.size() != radiation_out_bottom.size()) {
            helios_runtime_error("ERROR (HipRayTracingBackend::uploadRadiationOut): top and bottom buffers differ in size.");
        }
        check(rt_upload_radiation_out(ctx_, radiation_out_top.data(), radiation_out_bottom.data(), radiation_out_top.size()), "uploadRadiationOut");
    }

    void HipRayTracingBackend::zeroCameraScatterBuffers(size_t) {
        // the camera scatter buffers are the scatter buffers here (getRadiationResults), nothing to zero
    }

    void HipRayTracingBackend::uploadSourceFluxes(const std::vector<float> &fluxes) {
        if (fluxes.empty()) {
            return;
        }
        check(rt_upload_source_fluxes(ctx_, fluxes.data(), fluxes.size()), "uploadSourceFluxes");
    }

    void HipRayTracingBackend::uploadSourceFluxesCam(const std::vector<float> &fluxes_cam) {
        if (fluxes_cam.empty()) {
            return; // no camera weights (VulkanComputeBackend.cpp:2221-2223)
        }
        check(rt_upload_source_fluxes_cam(ctx_, fluxes_cam.data(), fluxes_cam.size()), "uploadSourceFluxesCam");
    }

    void HipRayTracingBackend::queryGPUMemory() const {
        size_t free_b = 0, total_b = 0;
        check(rt_query_memory(ctx_, &free_b, &total_b), "queryGPUMemory");
        std::cout << "GPU memory (" << rt_backend_name(ctx_) << "): " << (total_b - free_b) / (1024 * 1024) << " MB used of " << total_b / (1024 * 1024) << " MB" << std::endl;
    }

    std::string HipRayTracingBackend::getBackendName() const {
        return ctx_ ? std::string(rt_backend_name(ctx_)) : std::string("HIP gfx950 (not initialized)");
    }

} // namespace helios
