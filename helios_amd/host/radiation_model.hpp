// radiation_model.hpp -- C++ host interface of the gfx950 radiation backend with the call surface of
// Helios's RadiationModel (plugins/radiation/include/RadiationModel.h) for the radiation-transfer
// path: bands, ray counts, scattering depth, collimated sources, diffuse flux, updateGeometry,
// runBand and the per-primitive "radiation_flux_<band>" results.
//
// Geometry and per-primitive optical data come in as arrays (what RadiationModel::buildGeometryData
// and buildMaterialData derive from the Context, RadiationModel.cpp:5808-6148, 6313-6532) instead of
// a helios::Context.  runBand executes the whole pass schedule on the device (rt_run_band); errors
// throw std::runtime_error with the reference's messages where the reference throws.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/helios_rt.h"

namespace hiprt {

struct vec3 {
    float x, y, z;
};
struct vec2 {
    float x, y;
};

// Per-band settings (RadiationModel.h RadiationBand; defaults RadiationModel.cpp:40-55)
struct RadiationBand {
    size_t directRayCount = 100;
    size_t diffuseRayCount = 1000;
    float diffuseFlux = -1.f;  // < 0: not set
    uint32_t scatteringDepth = 0;
    bool emissionFlag = true;
    float diffuseExtinction = 0.f;           // power-law sky exponent K (0 = isotropic)
    float diffusePeakDir[3] = {0.f, 0.f, 1.f};
    float diffuseDistNorm = 1.f;
};

// Applies RadiationModel.cpp:6252-6311 to one primitive's properties in one band (clamp rho/tau to
// [0,1]; blackbody override when emitting with depth 0; rho = 1 - eps auto-correction; energy
// conservation checks that throw).
void validate_material(float &rho, float &tau, float eps, bool emission, uint32_t depth, const std::string &band,
                       uint32_t uuid);

class RadiationModel {
public:
    explicit RadiationModel(int device = 0);
    ~RadiationModel();
    RadiationModel(const RadiationModel &) = delete;
    RadiationModel &operator=(const RadiationModel &) = delete;

    std::string getBackendName() const;
    void disableMessages() {}

    // ---- geometry and primitive data (the Context's role) ----
    void setGeometry(const std::vector<float> &transforms, const std::vector<uint32_t> &types,
                     const std::vector<uint32_t> &uuids, const std::vector<uint32_t> &twosided, bool periodic_x = false,
                     bool periodic_y = false);
    // "reflectivity_<band>", "transmissivity_<band>", "emissivity_<band>", "temperature": [P] floats
    void setPrimitiveData(const std::string &label, const std::vector<float> &values);
    bool doesPrimitiveDataExist(const std::string &label) const;
    // also "radiation_flux_<band>" after runBand
    const std::vector<float> &getPrimitiveData(const std::string &label) const;
    size_t getPrimitiveCount() const { return types_.size(); }

    // ---- bands (RadiationModel.h:756-878, 1198) ----
    void addRadiationBand(const std::string &label);
    bool doesBandExist(const std::string &label) const;
    void disableEmission(const std::string &label);
    void enableEmission(const std::string &label);
    void setDirectRayCount(const std::string &label, size_t N);
    void setDiffuseRayCount(const std::string &label, size_t N);
    void setDiffuseRadiationFlux(const std::string &label, float flux);
    // Power-law diffuse sky psi^-K about peak_dir (RadiationModel.cpp:185-217); the normalisation makes
    // the cosine-weighted hemisphere integral of the distribution one.
    void setDiffuseRadiationExtinctionCoeff(const std::string &label, float K, const vec3 &peak_dir);
    float getDiffuseFlux(const std::string &label) const;
    void setScatteringDepth(const std::string &label, uint32_t depth);

    // ---- sources (RadiationModel.h:884-992) ----
    uint32_t addCollimatedRadiationSource();
    uint32_t addCollimatedRadiationSource(const vec3 &direction);
    uint32_t addSphereRadiationSource(const vec3 &position, float radius);
    uint32_t addSunSphereRadiationSource();
    uint32_t addSunSphereRadiationSource(const vec3 &sun_direction);
    uint32_t addRectangleRadiationSource(const vec3 &position, const vec2 &size, const vec3 &rotation_rad);
    uint32_t addDiskRadiationSource(const vec3 &position, float radius, const vec3 &rotation_rad);
    void setSourceFlux(uint32_t source_ID, const std::string &band_label, float flux);
    float getSourceFlux(uint32_t source_ID, const std::string &band_label) const;

    // ---- solve (RadiationModel.h:1532-1561) ----
    void updateGeometry();
    void runBand(const std::string &label);
    void runBand(const std::vector<std::string> &labels);
    std::vector<float> getTotalAbsorbedFlux() const;

    // ---- additions with no reference counterpart ----
    // Seeds of the next runBand: direct pass, then one per diffuse pass (primary + scattering).  By
    // default every runBand draws fresh seeds from a counter (the reference draws them from the clock).
    void setRandomSeeds(uint32_t direct_seed, const std::vector<uint32_t> &diffuse_seeds);
    // Also propagate the primary pass's scatter when emission is off (RT_SCHEDULE_CONSERVING).
    void setEnergyConservingSchedule(bool on) { schedule_ = on ? RT_SCHEDULE_CONSERVING : RT_SCHEDULE_REFERENCE; }

private:
    const RadiationBand &band(const std::string &label, const char *who) const;
    float primitiveValue(const std::string &label, size_t i, float fallback) const;

    rt_ctx *ctx_ = nullptr;
    rt_scene *scene_ = nullptr;
    std::vector<float> transforms_;
    std::vector<uint32_t> types_, uuids_, twosided_;
    bool periodic_x_ = false, periodic_y_ = false;
    bool geometry_dirty_ = true;
    std::map<std::string, RadiationBand> bands_;             // alphabetical = global band order
    uint32_t addSource(uint32_t type, const vec3 &position, const vec2 &width, const vec3 &rotation,
                       float flux_scale, const char *who);

    std::vector<vec3> source_dir_;                           // direction (collimated) or position
    std::vector<uint32_t> source_type_;                      // RT_SOURCE_*
    std::vector<vec2> source_width_;
    std::vector<vec3> source_rot_;
    std::vector<float> source_flux_scale_;                   // setSourceFlux multiplies by it (sun sphere)
    std::vector<std::map<std::string, float>> source_flux_;  // -1 = not set
    std::map<std::string, std::vector<float>> data_;
    uint32_t schedule_ = RT_SCHEDULE_REFERENCE;
    bool fixed_seeds_ = false;
    uint32_t direct_seed_ = 0;
    std::vector<uint32_t> diffuse_seeds_;
    uint32_t seed_counter_ = 1;
};

}  // namespace hiprt
