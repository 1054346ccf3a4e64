// HipRayTracingBackend.h -- helios::RayTracingBackend implemented on the gfx950 C ABI (include/helios_rt.h).
//
// This file builds inside a Helios source tree (it includes plugins/radiation/include/RayTracingBackend.h);
// it is not part of this repository's own build.  Each virtual maps to one or two C calls and every
// failure is rethrown through helios_runtime_error, as the reference backends do.  Radiation launches
// (direct, diffuse) for all five source types, texture-masked primitives and power-law / Prague skies,
// translucent (glass) covers, camera and pixel-label launches (with the Blinn-Phong specular term) and their
// buffer management are supported.
#ifndef HIP_RAYTRACING_BACKEND_H
#define HIP_RAYTRACING_BACKEND_H

#include <cstdint>
#include <vector>

#include "RayTracingBackend.h"
#include "helios_rt.h"

namespace helios {

    class HipRayTracingBackend : public RayTracingBackend {
    public:
        explicit HipRayTracingBackend(int device = 0);
        ~HipRayTracingBackend() override;

        void initialize() override;
        void shutdown() override;

        void updateGeometry(const RayTracingGeometry &geometry) override;
        void buildAccelerationStructure() override;
        void updateMaterials(const RayTracingMaterial &materials) override;
        void updateSources(const std::vector<RayTracingSource> &sources) override;
        void updateDiffuseRadiation(const std::vector<float> &flux, const std::vector<float> &extinction, const std::vector<helios::vec3> &peak_dir, const std::vector<float> &dist_norm,
                                    const std::vector<float> &sky_energy) override;
        void updateSkyModel(const std::vector<helios::vec4> &sky_radiance_params, const std::vector<float> &camera_sky_radiance, const helios::vec3 &sun_direction, const std::vector<float> &solar_disk_radiance,
                            float solar_disk_cos_angle, const std::vector<float> &camera_diffuse_flux, const std::vector<uint32_t> &band_emission_flag) override;

        void launchDirectRays(const RayTracingLaunchParams &params) override;
        void launchDiffuseRays(const RayTracingLaunchParams &params) override;
        void launchCameraRays(const RayTracingLaunchParams &params) override;
        void launchPixelLabelRays(const RayTracingLaunchParams &params) override;

        void getRadiationResults(RayTracingResults &results) override;
        void getCameraResults(std::vector<float> &pixel_data, std::vector<uint> &pixel_labels, std::vector<float> &pixel_depths, uint camera_id, const helios::int2 &resolution) override;

        void zeroRadiationBuffers(size_t launch_band_count) override;
        void zeroScatterBuffers() override;
        void zeroCameraPixelBuffers(const helios::int2 &resolution) override;
        void copyScatterToRadiation() override;
        void uploadRadiationOut(const std::vector<float> &radiation_out_top, const std::vector<float> &radiation_out_bottom) override;
        void zeroCameraScatterBuffers(size_t launch_band_count) override;
        void uploadSourceFluxes(const std::vector<float> &fluxes) override;
        void uploadSourceFluxesCam(const std::vector<float> &fluxes_cam) override;

        void queryGPUMemory() const override;
        std::string getBackendName() const override;

    private:
        void check(int rc, const char *who) const;
        // band_launch_flag [B_global] -> launch -> global map and per-slot active flags
        void bandTable(const RayTracingLaunchParams &params, const char *who);
        rt_launch_params cParams(const RayTracingLaunchParams &params) const;
        static rt_camera_params cameraParams(const RayTracingLaunchParams &params);

        int device_;
        rt_ctx *ctx_ = nullptr;
        size_t primitive_count_ = 0;
        size_t launch_band_count_ = 0;
        std::vector<uint32_t> band_map_;
        std::vector<uint8_t> band_active_;
    };

} // namespace helios

#endif
