// test_radiation_model.cpp -- tests of the C++ RadiationModel-style interface (helios_amd/host), written
// like the reference's plugins/radiation/tests/selfTest.cpp cases they port.
//
//   test_radiation_model            all cases (GPU cases need a gfx950 device)
//   test_radiation_model --cpu      host-only cases
//
// Prints one line per case and exits non-zero if any check failed.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../helios_amd/host/radiation_model.hpp"
#include "../../oracle/rt_oracle.h"

extern "C" uint32_t sg_config1(uint32_t seed, float *T, uint32_t *type, uint32_t *twos);

using hiprt::RadiationModel;
using hiprt::vec2;
using hiprt::vec3;

namespace {

struct Case {
    const char *name;
    bool gpu;
    std::function<void()> fn;
};
std::vector<Case> &cases() {
    static std::vector<Case> c;
    return c;
}
int g_failures = 0;
struct Reg {
    Reg(const char *n, bool gpu, std::function<void()> f) { cases().push_back({n, gpu, std::move(f)}); }
};

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define TEST(name, gpu)                                          \
    static void CAT(test_, __LINE__)();                          \
    static Reg CAT(reg_, __LINE__)(name, gpu, CAT(test_, __LINE__)); \
    static void CAT(test_, __LINE__)()
#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::printf("    CHECK failed at line %d: %s\n", __LINE__, #cond);       \
            g_failures++;                                                            \
        }                                                                            \
    } while (0)
#define CHECK_THROWS(stmt)                                                            \
    do {                                                                              \
        bool thrown_ = false;                                                         \
        try { stmt; } catch (const std::runtime_error &) { thrown_ = true; }          \
        if (!thrown_) {                                                               \
            std::printf("    CHECK_THROWS failed at line %d: %s\n", __LINE__, #stmt); \
            g_failures++;                                                             \
        }                                                                             \
    } while (0)

// Helios Context::addPatch transform (core/src/Context_primitive.cpp:36-56): scale, rotate by
// -elevation about x, by -azimuth about z, translate; row-major 4x4.
std::vector<float> patch_transform(vec3 c, float sx, float sy, float elevation, float azimuth) {
    double ce = std::cos(-elevation), se = std::sin(-elevation), ca = std::cos(-azimuth), sa = std::sin(-azimuth);
    // Rz * Rx
    double R[3][3] = {{ca, -sa * ce, sa * se}, {sa, ca * ce, -ca * se}, {0.0, se, ce}};
    std::vector<float> T(16, 0.f);
    for (int r = 0; r < 3; r++) {
        T[r * 4 + 0] = (float)(R[r][0] * sx);
        T[r * 4 + 1] = (float)(R[r][1] * sy);
        T[r * 4 + 2] = (float)R[r][2];
    }
    T[3] = c.x;
    T[7] = c.y;
    T[11] = c.z;
    T[15] = 1.f;
    return T;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// host-only
// ------------------------------------------------------------------------------------------------
TEST("validate_material rules (RadiationModel.cpp:6252-6311)", false) {
    float r = 1.5f, t = -0.2f;
    hiprt::validate_material(r, t, 1.f, false, 0, "SW", 0);
    CHECK(r == 1.f && t == 0.f);
    r = 0.3f; t = 0.2f;
    hiprt::validate_material(r, t, 0.9f, true, 0, "LW", 1);  // blackbody override
    CHECK(r == 0.f && t == 0.f);
    r = 0.f; t = 0.f;
    hiprt::validate_material(r, t, 0.95f, true, 2, "LW", 2);  // rho = 1 - eps
    CHECK(std::fabs(r - 0.05f) < 1e-7f && t == 0.f);
    r = 0.2f; t = 0.2f;
    CHECK_THROWS(hiprt::validate_material(r, t, 0.95f, true, 2, "LW", 3));
    r = 0.7f; t = 0.4f;
    CHECK_THROWS(hiprt::validate_material(r, t, 1.f, false, 1, "NIR", 4));
    r = 0.45f; t = 0.40f;
    hiprt::validate_material(r, t, 1.f, false, 1, "NIR", 5);
    CHECK(r == 0.45f && t == 0.40f);
}

// ------------------------------------------------------------------------------------------------
// GPU
// ------------------------------------------------------------------------------------------------
TEST("RadiationModel Simple Direct (selfTest.cpp:518-545)", true) {
    RadiationModel radiation;
    radiation.disableMessages();
    radiation.setGeometry(patch_transform({0, 0, 0}, 1, 1, 0, 0), {0}, {0}, {0});
    radiation.setPrimitiveData("reflectivity_SW", {0.f});
    radiation.addRadiationBand("SW");
    radiation.disableEmission("SW");
    uint32_t sun = radiation.addCollimatedRadiationSource(vec3{0, 0, 1});
    radiation.setSourceFlux(sun, "SW", 1000.f);
    radiation.setDirectRayCount("SW", 10000);
    radiation.setScatteringDepth("SW", 0);
    radiation.updateGeometry();
    radiation.runBand("SW");
    float flux = radiation.getPrimitiveData("radiation_flux_SW")[0];
    CHECK(std::fabs(flux - 1000.f) / 1000.f <= 0.01f);
}

TEST("RadiationModel power-law sky: a lone black patch absorbs the sky flux (normalised distribution)", true) {
    // the normalisation makes the cosine-weighted hemisphere integral of psi^-K one, so a horizontal black
    // patch with nothing around absorbs the diffuse flux whatever K and the peak; a wrong normalisation or
    // unit slip moves it far beyond the Monte Carlo error of 2500 rays
    for (float K : {0.5f, 1.2f}) {
        RadiationModel radiation;
        radiation.disableMessages();
        radiation.setGeometry(patch_transform({0, 0, 0}, 1, 1, 0, 0), {0}, {0}, {0});
        radiation.setPrimitiveData("reflectivity_SW", {0.f});
        radiation.addRadiationBand("SW");
        radiation.disableEmission("SW");
        radiation.setDiffuseRadiationFlux("SW", 100.f);
        radiation.setDiffuseRadiationExtinctionCoeff("SW", K, vec3{0.3f, 0.2f, 0.9f});
        radiation.setDiffuseRayCount("SW", 2500);
        radiation.setScatteringDepth("SW", 0);
        radiation.updateGeometry();
        radiation.runBand("SW");
        float flux = radiation.getPrimitiveData("radiation_flux_SW")[0];
        CHECK(std::fabs(flux - 100.f) <= 4.f);
    }
}

TEST("RadiationModel 90 Degree Common-Edge Squares (selfTest.cpp:585-668)", true) {
    const float error_threshold = 0.005f;
    const int Nensemble = 500;
    const uint32_t Ndiffuse_1 = 100000, Ndirect_1 = 5000;
    const float Qs = 1000.f, sigma = 5.6703744E-8f;
    const float shortwave_exact_0 = 0.7f * Qs;
    const float shortwave_exact_1 = 0.3f * 0.2f * Qs;
    const float longwave_exact_0 = 0.f;
    const float longwave_exact_1 = sigma * std::pow(300.f, 4.f) * 0.2f;

    std::vector<float> T = patch_transform({0, 0, 0}, 1, 1, 0, 0);
    std::vector<float> T1 = patch_transform({0.5f, 0, 0.5f}, 1, 1, 0.5f * (float)M_PI, -0.5f * (float)M_PI);
    T.insert(T.end(), T1.begin(), T1.end());
    RadiationModel radiationmodel_1;
    radiationmodel_1.setGeometry(T, {0, 0}, {0, 1}, {0, 0});
    radiationmodel_1.setPrimitiveData("temperature", {300.f, 0.f});
    radiationmodel_1.setPrimitiveData("reflectivity_SW", {0.3f, 0.f});

    radiationmodel_1.addRadiationBand("LW");
    radiationmodel_1.setDirectRayCount("LW", Ndiffuse_1);
    radiationmodel_1.setDiffuseRayCount("LW", Ndiffuse_1);
    radiationmodel_1.setScatteringDepth("LW", 0);

    uint32_t SunSource_1 = radiationmodel_1.addCollimatedRadiationSource(vec3{0, 0, 1});
    radiationmodel_1.addRadiationBand("SW");
    radiationmodel_1.disableEmission("SW");
    radiationmodel_1.setDirectRayCount("SW", Ndirect_1);
    radiationmodel_1.setDiffuseRayCount("SW", Ndirect_1);
    radiationmodel_1.setScatteringDepth("SW", 1);
    radiationmodel_1.setSourceFlux(SunSource_1, "SW", Qs);
    radiationmodel_1.updateGeometry();

    double lw0 = 0, lw1 = 0, sw0 = 0, sw1 = 0;
    for (int r = 0; r < Nensemble; r++) {
        radiationmodel_1.runBand(std::vector<std::string>{"LW", "SW"});
        lw0 += radiationmodel_1.getPrimitiveData("radiation_flux_LW")[0] / (double)Nensemble;
        lw1 += radiationmodel_1.getPrimitiveData("radiation_flux_LW")[1] / (double)Nensemble;
        sw0 += radiationmodel_1.getPrimitiveData("radiation_flux_SW")[0] / (double)Nensemble;
        sw1 += radiationmodel_1.getPrimitiveData("radiation_flux_SW")[1] / (double)Nensemble;
    }
    std::printf("    SW %.4f %.4f (exact %.4f %.4f)  LW %.4f %.4f (exact %.4f %.4f)\n", sw0, sw1, shortwave_exact_0,
                shortwave_exact_1, lw0, lw1, longwave_exact_0, longwave_exact_1);
    CHECK(std::fabs(sw0 - shortwave_exact_0) / shortwave_exact_0 <= error_threshold);
    CHECK(std::fabs(sw1 - shortwave_exact_1) / shortwave_exact_1 <= error_threshold);
    CHECK(lw0 == longwave_exact_0);
    CHECK(std::fabs(lw1 - longwave_exact_1) / longwave_exact_1 <= error_threshold);
    std::vector<float> tot = radiationmodel_1.getTotalAbsorbedFlux();
    CHECK(tot.size() == 2);
    CHECK(tot[1] == radiationmodel_1.getPrimitiveData("radiation_flux_LW")[1] +
                        radiationmodel_1.getPrimitiveData("radiation_flux_SW")[1]);
}

TEST("RadiationModel config-1 canopy against the FP64 oracle (1e-4)", true) {
    const uint32_t P = sg_config1(1, nullptr, nullptr, nullptr);
    std::vector<float> T(P * 16);
    std::vector<uint32_t> type(P), twos(P);
    sg_config1(1, T.data(), type.data(), twos.data());
    std::vector<float> rho(P, 0.10f), tau(P, 0.05f);
    rho[P - 1] = 0.15f;
    tau[P - 1] = 0.f;

    RadiationModel radiation;
    radiation.setGeometry(T, type, {}, twos);
    radiation.setPrimitiveData("reflectivity_PAR", rho);
    radiation.setPrimitiveData("transmissivity_PAR", tau);
    radiation.addRadiationBand("PAR");
    radiation.disableEmission("PAR");
    const float z = 30.f * (float)M_PI / 180.f;
    vec3 sun{0.f, std::sin(z), std::cos(z)};  // zenith 30 deg, azimuth 0 (towards +y)
    uint32_t s = radiation.addCollimatedRadiationSource(sun);
    radiation.setSourceFlux(s, "PAR", 1000.f);
    radiation.setDiffuseRadiationFlux("PAR", 100.f);
    radiation.setDirectRayCount("PAR", 100);
    radiation.setDiffuseRayCount("PAR", 100);
    radiation.setScatteringDepth("PAR", 3);
    const std::vector<uint32_t> dseeds{11, 12, 13, 14};
    radiation.setRandomSeeds(7, dseeds);
    radiation.runBand("PAR");
    const std::vector<float> &R = radiation.getPrimitiveData("radiation_flux_PAR");

    orc_scene *os = orc_scene_create(P, T.data(), type.data(), twos.data(), nullptr, 0, 0);
    const uint32_t depth = 3, emission = 0;
    const float sdir[3] = {sun.x, sun.y, sun.z}, sflux = 1000.f, dflux = 100.f;
    orc_band_params q{};
    q.nbands = 1;
    q.nsources = 1;
    q.direct_rays = 100;
    q.diffuse_rays = 100;
    q.band_depth = &depth;
    q.emission = &emission;
    q.source_dir = sdir;
    q.source_flux = &sflux;
    q.diffuse_flux = &dflux;
    q.rho = rho.data();
    q.tau = tau.data();
    q.direct_seed = 7;
    q.diffuse_seed = dseeds.data();
    std::vector<double> Ro(P);
    CHECK(orc_run_band(os, &q, Ro.data(), nullptr) == 0);
    orc_scene_destroy(os);
    double worst = 0;
    for (uint32_t i = 0; i < P; i++) worst = std::max(worst, std::fabs(R[i] - Ro[i]) / std::max(std::fabs(Ro[i]), 1.0));
    std::printf("    max rel err vs oracle %.3e over %u primitives\n", worst, P);
    CHECK(worst <= 1e-4);
}

TEST("RadiationModel sphere, rectangle and disk sources vs oracle", true) {
    const uint32_t P = 1601;
    std::vector<float> T(P * 16);
    std::vector<uint32_t> type(P), twos(P);
    sg_config1(2, T.data(), type.data(), twos.data());
    std::vector<float> rho(P, 0.10f), tau(P, 0.05f);
    rho[P - 1] = 0.15f;
    tau[P - 1] = 0.f;

    RadiationModel radiation;
    radiation.setGeometry(T, type, {}, twos);
    radiation.setPrimitiveData("reflectivity_PAR", rho);
    radiation.setPrimitiveData("transmissivity_PAR", tau);
    radiation.addRadiationBand("PAR");
    radiation.disableEmission("PAR");
    const vec3 sph{0.5f, -0.3f, 3.f}, rect{0.f, 0.f, 2.5f}, disk{-0.6f, 0.4f, 2.f};
    const vec2 rsize{1.5f, 1.f};
    const vec3 rrot{(float)M_PI, 0.2f, 0.f}, drot{(float)M_PI - 0.2f, 0.f, 0.3f};
    uint32_t a = radiation.addSphereRadiationSource(sph, 0.25f);
    uint32_t b = radiation.addRectangleRadiationSource(rect, rsize, rrot);
    uint32_t c = radiation.addDiskRadiationSource(disk, 0.4f, drot);
    radiation.setSourceFlux(a, "PAR", 2000.f);
    radiation.setSourceFlux(b, "PAR", 800.f);
    radiation.setSourceFlux(c, "PAR", 1200.f);
    radiation.setDiffuseRadiationFlux("PAR", 50.f);
    radiation.setDirectRayCount("PAR", 64);
    radiation.setDiffuseRayCount("PAR", 64);
    radiation.setScatteringDepth("PAR", 2);
    const std::vector<uint32_t> dseeds{21, 22, 23};
    radiation.setRandomSeeds(9, dseeds);
    radiation.runBand("PAR");
    const std::vector<float> &R = radiation.getPrimitiveData("radiation_flux_PAR");

    orc_scene *os = orc_scene_create(P, T.data(), type.data(), twos.data(), nullptr, 0, 0);
    const uint32_t depth = 2, emission = 0;
    const float sdir[9] = {sph.x, sph.y, sph.z, rect.x, rect.y, rect.z, disk.x, disk.y, disk.z};
    const float sflux[3] = {2000.f, 800.f, 1200.f}, dflux = 50.f;
    const uint32_t stype[3] = {RT_SOURCE_SPHERE, RT_SOURCE_RECTANGLE, RT_SOURCE_DISK};
    const float swidth[6] = {0.5f, 0.5f, rsize.x, rsize.y, 0.4f, 0.4f};
    const float srot[9] = {0.f, 0.f, 0.f, rrot.x, rrot.y, rrot.z, drot.x, drot.y, drot.z};
    orc_band_params q{};
    q.nbands = 1;
    q.nsources = 3;
    q.direct_rays = 64;
    q.diffuse_rays = 64;
    q.band_depth = &depth;
    q.emission = &emission;
    q.source_dir = sdir;
    q.source_flux = sflux;
    q.diffuse_flux = &dflux;
    q.rho = rho.data();
    q.tau = tau.data();
    q.direct_seed = 9;
    q.diffuse_seed = dseeds.data();
    q.source_type = stype;
    q.source_width = swidth;
    q.source_rotation = srot;
    std::vector<double> Ro(P);
    CHECK(orc_run_band(os, &q, Ro.data(), nullptr) == 0);
    orc_scene_destroy(os);
    double worst = 0, lit = 0;
    for (uint32_t i = 0; i < P; i++) {
        worst = std::max(worst, std::fabs(R[i] - Ro[i]) / std::max(std::fabs(Ro[i]), 1.0));
        lit = std::max(lit, Ro[i]);
    }
    std::printf("    max rel err vs oracle %.3e over %u primitives (max flux %.1f)\n", worst, P, lit);
    CHECK(lit > 100.0);
    CHECK(worst <= 1e-4);
}

TEST("RadiationModel errors", true) {
    RadiationModel radiation;
    radiation.setGeometry(patch_transform({0, 0, 0}, 1, 1, 0, 0), {0}, {0}, {0});
    radiation.addRadiationBand("SW");
    CHECK_THROWS(radiation.runBand("NIR"));
    CHECK_THROWS(radiation.setSourceFlux(3, "SW", 1.f));
    CHECK_THROWS(radiation.setScatteringDepth("NIR", 1));
    CHECK_THROWS(radiation.setPrimitiveData("reflectivity_SW", {0.1f, 0.2f}));
    CHECK_THROWS(radiation.addCollimatedRadiationSource(vec3{0, 0, 0}));
    CHECK_THROWS(radiation.addSphereRadiationSource(vec3{0, 0, 1}, -1.f));
    CHECK_THROWS(radiation.addRectangleRadiationSource(vec3{0, 0, 1}, vec2{0.f, 1.f}, vec3{0, 0, 0}));
    CHECK_THROWS(radiation.addDiskRadiationSource(vec3{0, 0, 1}, 0.f, vec3{0, 0, 0}));
    CHECK_THROWS(radiation.getPrimitiveData("radiation_flux_SW"));
    CHECK(radiation.getBackendName().find("gfx950") != std::string::npos);
}

int main(int argc, char **argv) {
    bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
    int run = 0;
    for (auto &c : cases()) {
        if (c.gpu && cpu_only) continue;
        int before = g_failures;
        std::printf("[case] %s\n", c.name);
        std::fflush(stdout);
        try {
            c.fn();
        } catch (const std::exception &e) {
            std::printf("    unexpected exception: %s\n", e.what());
            g_failures++;
        }
        std::printf("[%s] %s\n", g_failures == before ? "pass" : "FAIL", c.name);
        std::fflush(stdout);
        run++;
    }
    std::printf("%d cases, %d failed checks\n", run, g_failures);
    return g_failures ? 1 : 0;
}
