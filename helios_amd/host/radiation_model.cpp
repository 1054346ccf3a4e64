// radiation_model.cpp -- RadiationModel-style host interface over the C ABI (include/helios_rt.h).
#include "radiation_model.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace hiprt {

namespace {

void fail(const std::string &msg) { throw std::runtime_error(msg); }

void check(int rc, const char *who) {
    if (rc != 0) fail(std::string("ERROR (RadiationModel::") + who + "): " + rt_last_error());
}

}  // namespace

void validate_material(float &rho, float &tau, float eps, bool emission, uint32_t depth, const std::string &band,
                       uint32_t uuid) {
    // RadiationModel.cpp:6252-6311
    rho = std::max(0.f, std::min(1.f, rho));
    tau = std::max(0.f, std::min(1.f, tau));
    const std::string where = "band " + band + ", primitive #" + std::to_string(uuid) + ": eps=" +
                              std::to_string(eps) + ", tau=" + std::to_string(tau) + ", rho=" + std::to_string(rho);
    if (emission) {
        if (depth == 0 && eps != 1.f) {
            rho = 0.f;
            tau = 0.f;
        } else if (eps != 1.f && rho == 0.f && tau == 0.f) {
            rho = 1.f - eps;
        } else if (std::fabs(eps + rho + tau - 1.f) > 1e-5f && eps > 0.f) {
            fail("ERROR (RadiationModel): emissivity, transmissivity and reflectivity must sum to 1 (" + where +
                 "); emission may need to be disabled for this band");
        }
    } else if (rho + tau > 1.f) {
        fail("ERROR (RadiationModel): reflectivity and transmissivity sum to more than 1 (" + where + ")");
    }
}

RadiationModel::RadiationModel(int device) { check(rt_create(device, &ctx_), "RadiationModel"); }

RadiationModel::~RadiationModel() {
    if (scene_) rt_scene_destroy(ctx_, scene_);
    if (ctx_) rt_destroy(ctx_);
}

std::string RadiationModel::getBackendName() const { return rt_backend_name(ctx_); }

void RadiationModel::setGeometry(const std::vector<float> &transforms, const std::vector<uint32_t> &types,
                                 const std::vector<uint32_t> &uuids, const std::vector<uint32_t> &twosided,
                                 bool periodic_x, bool periodic_y) {
    const size_t P = types.size();
    if (transforms.size() != 16 * P || twosided.size() != P || (!uuids.empty() && uuids.size() != P))
        fail("ERROR (RadiationModel::setGeometry): transforms must hold 16 floats and twosided/uuids one value per "
             "primitive");
    transforms_ = transforms;
    types_ = types;
    twosided_ = twosided;
    uuids_ = uuids;
    if (uuids_.empty()) {
        uuids_.resize(P);
        for (size_t i = 0; i < P; i++) uuids_[i] = (uint32_t)i;
    }
    periodic_x_ = periodic_x;
    periodic_y_ = periodic_y;
    geometry_dirty_ = true;
}

void RadiationModel::setPrimitiveData(const std::string &label, const std::vector<float> &values) {
    if (values.size() != types_.size())
        fail("ERROR (RadiationModel::setPrimitiveData): '" + label + "' needs one value per primitive (" +
             std::to_string(types_.size()) + ")");
    data_[label] = values;
}

bool RadiationModel::doesPrimitiveDataExist(const std::string &label) const { return data_.count(label) != 0; }

const std::vector<float> &RadiationModel::getPrimitiveData(const std::string &label) const {
    auto it = data_.find(label);
    if (it == data_.end()) fail("ERROR (RadiationModel::getPrimitiveData): no primitive data '" + label + "'");
    return it->second;
}

float RadiationModel::primitiveValue(const std::string &label, size_t i, float fallback) const {
    auto it = data_.find(label);
    return it == data_.end() ? fallback : it->second[i];
}

void RadiationModel::addRadiationBand(const std::string &label) {
    // RadiationModel.cpp:306-328: a second add of the same label is ignored
    if (bands_.count(label)) return;
    bands_.emplace(label, RadiationBand{});
    for (auto &f : source_flux_) f[label] = -1.f;
}

bool RadiationModel::doesBandExist(const std::string &label) const { return bands_.count(label) != 0; }

const RadiationBand &RadiationModel::band(const std::string &label, const char *who) const {
    auto it = bands_.find(label);
    if (it == bands_.end())
        fail(std::string("ERROR (RadiationModel::") + who + "): band '" + label + "' does not exist (addRadiationBand)");
    return it->second;
}

void RadiationModel::disableEmission(const std::string &label) {
    band(label, "disableEmission");
    bands_[label].emissionFlag = false;
}

void RadiationModel::enableEmission(const std::string &label) {
    band(label, "enableEmission");
    bands_[label].emissionFlag = true;
}

void RadiationModel::setDirectRayCount(const std::string &label, size_t N) {
    band(label, "setDirectRayCount");
    bands_[label].directRayCount = N;
}

void RadiationModel::setDiffuseRayCount(const std::string &label, size_t N) {
    band(label, "setDiffuseRayCount");
    bands_[label].diffuseRayCount = N;
}

void RadiationModel::setDiffuseRadiationFlux(const std::string &label, float flux) {
    band(label, "setDiffuseRadiationFlux");
    bands_[label].diffuseFlux = flux;
}

void RadiationModel::setDiffuseRadiationExtinctionCoeff(const std::string &label, float K, const vec3 &peak_dir) {
    band(label, "setDiffuseRadiationExtinctionCoeff");
    const double m = std::sqrt((double)peak_dir.x * peak_dir.x + (double)peak_dir.y * peak_dir.y + (double)peak_dir.z * peak_dir.z);
    if (!(m > 0.0)) throw std::runtime_error("ERROR (RadiationModel::setDiffuseRadiationExtinctionCoeff): peak direction is zero.");
    const float d[3] = {(float)(peak_dir.x / m), (float)(peak_dir.y / m), (float)(peak_dir.z / m)};
    // midpoint rule over 100 x 100 (theta, phi) cells of the upper hemisphere of
    // fd(psi) cos(theta) sin(theta) dtheta dphi / pi, with psi clamped below at one degree
    const int N = 100;
    const float pi = 3.14159265358979f;
    float norm = 0.f;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            const float th = 0.5f * pi / (float)N * (0.5f + (float)i);
            const float ph = 2.f * pi / (float)N * (0.5f + (float)j);
            const float nx = std::sin(th) * std::cos(ph), ny = std::sin(th) * std::sin(ph), nz = std::cos(th);
            const float c = std::fmin(1.f, std::fmax(-1.f, nx * d[0] + ny * d[1] + nz * d[2]));
            const float psi = std::fmax(std::acos(c), pi / 180.f);
            norm += std::pow(psi, -K) * std::cos(th) * std::sin(th) * pi / (float)(N * N);
        }
    RadiationBand &b = bands_[label];
    b.diffuseExtinction = K;
    for (int a = 0; a < 3; a++) b.diffusePeakDir[a] = d[a];
    b.diffuseDistNorm = 1.f / norm;
}

float RadiationModel::getDiffuseFlux(const std::string &label) const {
    // RadiationModel.cpp:1312-1343 without spectra: a flux that was never set is 0
    const RadiationBand &b = band(label, "getDiffuseFlux");
    return b.diffuseFlux >= 0.f ? b.diffuseFlux : 0.f;
}

void RadiationModel::setScatteringDepth(const std::string &label, uint32_t depth) {
    band(label, "setScatteringDepth");
    bands_[label].scatteringDepth = depth;
}

uint32_t RadiationModel::addCollimatedRadiationSource() { return addCollimatedRadiationSource(vec3{0.f, 0.f, 1.f}); }

uint32_t RadiationModel::addSource(uint32_t type, const vec3 &position, const vec2 &width, const vec3 &rotation,
                                   float flux_scale, const char *who) {
    if (source_dir_.size() + 1 > 256) fail(std::string("ERROR (RadiationModel::") + who + "): at most 256 radiation sources");
    source_dir_.push_back(position);
    source_type_.push_back(type);
    source_width_.push_back(width);
    source_rot_.push_back(rotation);
    source_flux_scale_.push_back(flux_scale);
    std::map<std::string, float> f;
    for (auto &b : bands_) f[b.first] = -1.f;
    source_flux_.push_back(f);
    return (uint32_t)(source_dir_.size() - 1);
}

uint32_t RadiationModel::addCollimatedRadiationSource(const vec3 &direction) {
    // RadiationModel.cpp:976-1010
    if (direction.x == 0.f && direction.y == 0.f && direction.z == 0.f)
        fail("ERROR (RadiationModel::addCollimatedRadiationSource): direction has zero length");
    return addSource(RT_SOURCE_COLLIMATED, direction, vec2{0.f, 0.f}, vec3{0.f, 0.f, 0.f}, 1.f,
                     "addCollimatedRadiationSource");
}

uint32_t RadiationModel::addSphereRadiationSource(const vec3 &position, float radius) {
    // RadiationModel.cpp:1011-1040: width = diameter
    if (!(radius > 0.f)) fail("ERROR (RadiationModel::addSphereRadiationSource): Spherical radiation source radius must be positive.");
    const float w = 2.f * std::fabs(radius);
    return addSource(RT_SOURCE_SPHERE, position, vec2{w, w}, vec3{0.f, 0.f, 0.f}, 1.f, "addSphereRadiationSource");
}

uint32_t RadiationModel::addSunSphereRadiationSource() { return addSunSphereRadiationSource(vec3{0.f, 0.f, 1.f}); }

uint32_t RadiationModel::addSunSphereRadiationSource(const vec3 &sun_direction) {
    // RadiationModel.cpp:1050-1079: the sun as a sphere of diameter 2 x 695.5e6 m at 150e9 m, fluxes set with
    // setSourceFlux scaled by sigma * 5700^4 / 1288.437
    const float m = std::sqrt(sun_direction.x * sun_direction.x + sun_direction.y * sun_direction.y +
                              sun_direction.z * sun_direction.z);
    if (!(m > 0.f)) fail("ERROR (RadiationModel::addSunSphereRadiationSource): direction has zero length");
    const vec3 pos{150e9f * sun_direction.x / m, 150e9f * sun_direction.y / m, 150e9f * sun_direction.z / m};
    const float w = 2.f * 695.5e6f;
    const float sigma = 5.6703744E-8f;
    return addSource(RT_SOURCE_SUN_SPHERE, pos, vec2{w, w}, vec3{0.f, 0.f, 0.f}, sigma * std::pow(5700.f, 4.f) / 1288.437f,
                     "addSunSphereRadiationSource");
}

uint32_t RadiationModel::addRectangleRadiationSource(const vec3 &position, const vec2 &size, const vec3 &rotation_rad) {
    // RadiationModel.cpp:1081-1110
    if (!(size.x > 0.f) || !(size.y > 0.f))
        fail("ERROR (RadiationModel::addRectangleRadiationSource): Radiation source size must be positive.");
    return addSource(RT_SOURCE_RECTANGLE, position, size, rotation_rad, 1.f, "addRectangleRadiationSource");
}

uint32_t RadiationModel::addDiskRadiationSource(const vec3 &position, float radius, const vec3 &rotation_rad) {
    // RadiationModel.cpp:1112-1140: width = (radius, radius)
    if (!(radius > 0.f)) fail("ERROR (RadiationModel::addDiskRadiationSource): Disk radiation source radius must be positive.");
    return addSource(RT_SOURCE_DISK, position, vec2{radius, radius}, rotation_rad, 1.f, "addDiskRadiationSource");
}

void RadiationModel::setSourceFlux(uint32_t source_ID, const std::string &band_label, float flux) {
    band(band_label, "setSourceFlux");
    if (source_ID >= source_dir_.size())
        fail("ERROR (RadiationModel::setSourceFlux): source ID " + std::to_string(source_ID) + " out of bounds");
    // RadiationModel.cpp:1199: a sun-sphere source scales the flux it is given
    source_flux_[source_ID][band_label] = flux * source_flux_scale_[source_ID];
}

float RadiationModel::getSourceFlux(uint32_t source_ID, const std::string &band_label) const {
    // RadiationModel.cpp:1208-1231 without spectra: a flux that was never set is 0
    band(band_label, "getSourceFlux");
    if (source_ID >= source_dir_.size())
        fail("ERROR (RadiationModel::getSourceFlux): source ID " + std::to_string(source_ID) + " out of bounds");
    float f = source_flux_[source_ID].at(band_label);
    return f < 0.f ? 0.f : f;
}

void RadiationModel::updateGeometry() {
    if (scene_) {
        check(rt_scene_destroy(ctx_, scene_), "updateGeometry");
        scene_ = nullptr;
    }
    check(rt_scene_create(ctx_, (uint32_t)types_.size(), transforms_.data(), types_.data(), uuids_.data(),
                          twosided_.data(), periodic_x_ ? 1 : 0, periodic_y_ ? 1 : 0, &scene_),
          "updateGeometry");
    geometry_dirty_ = false;
}

void RadiationModel::setRandomSeeds(uint32_t direct_seed, const std::vector<uint32_t> &diffuse_seeds) {
    fixed_seeds_ = true;
    direct_seed_ = direct_seed;
    diffuse_seeds_ = diffuse_seeds;
}

void RadiationModel::runBand(const std::string &label) { runBand(std::vector<std::string>{label}); }

void RadiationModel::runBand(const std::vector<std::string> &labels) {
    if (labels.empty()) fail("ERROR (RadiationModel::runBand): no bands given");
    for (auto &l : labels) band(l, "runBand");
    if (geometry_dirty_ || !scene_) updateGeometry();
    // launched bands in global (alphabetical) order, RadiationModel.cpp:3470-3560
    std::vector<std::string> launched;
    for (auto &b : bands_)
        if (std::find(labels.begin(), labels.end(), b.first) != labels.end()) launched.push_back(b.first);
    const size_t P = types_.size(), B = launched.size(), S = source_dir_.size();

    std::vector<uint32_t> depth(B), emission(B);
    uint32_t direct_rays = 0, diffuse_rays = 0, maxdepth = 0;
    bool any_emission = false;
    for (size_t b = 0; b < B; b++) {
        const RadiationBand &rb = bands_.at(launched[b]);
        depth[b] = rb.scatteringDepth;
        emission[b] = rb.emissionFlag ? 1u : 0u;
        any_emission |= rb.emissionFlag;
        direct_rays = std::max(direct_rays, (uint32_t)rb.directRayCount);
        diffuse_rays = std::max(diffuse_rays, (uint32_t)rb.diffuseRayCount);
        maxdepth = std::max(maxdepth, rb.scatteringDepth);
    }
    std::vector<float> rho(P * B), tau(P * B), eps(any_emission ? P * B : 0);
    for (size_t b = 0; b < B; b++) {
        const std::string &l = launched[b];
        for (size_t i = 0; i < P; i++) {
            float r = primitiveValue("reflectivity_" + l, i, 0.f);
            float t = primitiveValue("transmissivity_" + l, i, 0.f);
            float e = primitiveValue("emissivity_" + l, i, 1.f);
            validate_material(r, t, e, emission[b] != 0, depth[b], l, uuids_[i]);
            rho[i * B + b] = r;
            tau[i * B + b] = t;
            if (any_emission) eps[i * B + b] = e;
        }
    }
    std::vector<float> temperature;
    if (any_emission && doesPrimitiveDataExist("temperature")) temperature = data_.at("temperature");

    std::vector<float> sdir(S * 3), sflux(S * B), dflux(B), swidth(S * 2), srot(S * 3);
    for (size_t s = 0; s < S; s++) {
        sdir[s * 3] = source_dir_[s].x;
        sdir[s * 3 + 1] = source_dir_[s].y;
        sdir[s * 3 + 2] = source_dir_[s].z;
        swidth[s * 2] = source_width_[s].x;
        swidth[s * 2 + 1] = source_width_[s].y;
        srot[s * 3] = source_rot_[s].x;
        srot[s * 3 + 1] = source_rot_[s].y;
        srot[s * 3 + 2] = source_rot_[s].z;
        for (size_t b = 0; b < B; b++) sflux[s * B + b] = getSourceFlux((uint32_t)s, launched[b]);
    }
    for (size_t b = 0; b < B; b++) dflux[b] = getDiffuseFlux(launched[b]);
    // power-law skies of the launched bands (all bands isotropic: the request carries none)
    std::vector<float> kext(B, 0.f), kpeak(3 * B, 0.f), knorm(B, 0.f);
    bool any_sky = false;
    for (size_t b = 0; b < B; b++) {
        const RadiationBand &rb = bands_.at(launched[b]);
        kext[b] = rb.diffuseExtinction;
        knorm[b] = rb.diffuseDistNorm;
        for (int a = 0; a < 3; a++) kpeak[3 * b + a] = rb.diffusePeakDir[a];
        any_sky = any_sky || rb.diffuseExtinction > 0.f;
    }

    std::vector<uint32_t> dseeds(maxdepth + 1);
    uint32_t dseed;
    if (fixed_seeds_) {
        dseed = direct_seed_;
        for (size_t k = 0; k < dseeds.size(); k++)
            dseeds[k] = k < diffuse_seeds_.size() ? diffuse_seeds_[k]
                                                  : (diffuse_seeds_.empty() ? 1u : diffuse_seeds_.back()) + (uint32_t)k;
    } else {
        dseed = 0x9E3779B9u * seed_counter_++;
        for (auto &v : dseeds) v = 0x9E3779B9u * seed_counter_++;
    }

    rt_band_request q{};
    q.nbands = (uint32_t)B;
    q.nsources = (uint32_t)S;
    q.direct_rays = direct_rays;
    q.diffuse_rays = diffuse_rays;
    q.band_depth = depth.data();
    q.emission = emission.data();
    q.source_dir = S ? sdir.data() : nullptr;
    q.source_flux = S ? sflux.data() : nullptr;
    q.source_type = S ? source_type_.data() : nullptr;
    q.source_width = S ? swidth.data() : nullptr;
    q.source_rotation = S ? srot.data() : nullptr;
    q.diffuse_flux = dflux.data();
    q.rho = rho.data();
    q.tau = tau.data();
    q.eps = any_emission ? eps.data() : nullptr;
    q.temperature = temperature.empty() ? nullptr : temperature.data();
    q.direct_seed = dseed;
    q.diffuse_seed = dseeds.data();
    q.schedule = schedule_;
    if (any_sky) {
        q.diffuse_extinction = kext.data();
        q.diffuse_peak_dir = kpeak.data();
        q.diffuse_dist_norm = knorm.data();
    }
    std::vector<float> R(P * B);
    check(rt_run_band(ctx_, scene_, &q, R.data()), "runBand");
    for (size_t b = 0; b < B; b++) {
        std::vector<float> col(P);
        for (size_t i = 0; i < P; i++) col[i] = R[i * B + b];
        data_["radiation_flux_" + launched[b]] = std::move(col);
    }
}

std::vector<float> RadiationModel::getTotalAbsorbedFlux() const {
    // RadiationModel.cpp:4686-4708: sum of radiation_flux_<band> over every band
    std::vector<float> total(types_.size(), 0.f);
    for (auto &b : bands_) {
        auto it = data_.find("radiation_flux_" + b.first);
        if (it == data_.end()) fail("ERROR (RadiationModel::getTotalAbsorbedFlux): band '" + b.first + "' has not been run");
        for (size_t i = 0; i < total.size(); i++) total[i] += it->second[i];
    }
    return total;
}

}  // namespace hiprt
