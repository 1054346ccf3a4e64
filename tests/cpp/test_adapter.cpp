// test_adapter.cpp -- runs helios::HipRayTracingBackend (helios_amd/host/HipRayTracingBackend.cpp) on a CPU
// through the reference's RayTracingBackend interface (plugins/radiation/include/RayTracingBackend.h), with
// the C ABI below it replaced by a recording test double (adapter_fake_abi.cpp).  It drives the calls
// RadiationModel::runBand makes (R/src/RadiationModel.cpp:3804-4356) and checks what reaches the C ABI:
// the geometry conversion (UUIDs, types, twosided flags, periodic flags), the texture-mask expansion of the
// compact uv_data, materials and sources, the band table of launches with retired bands, the launch
// parameters, the results copied back, and the errors for unsupported features.
//
// Built and run by tests/test_host_cpp.py on the CPU when a Helios source tree is present; the only
// reference code it links is RayTracingTypes.cpp (RayTracingGeometry::validate), compiled where it lies.
// helios::helios_runtime_error lives in the reference's global.cpp, which needs zlib/libpng/libjpeg/
// tinyexr (absent here): the test double below throws std::runtime_error(message), as that function does.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "HipRayTracingBackend.h"
#include "adapter_fake_abi.h"

void helios::helios_runtime_error(const std::string &error_message) { throw std::runtime_error(error_message); }

static int g_fail = 0;
#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);              \
            g_fail++;                                                             \
        }                                                                         \
    } while (0)

template <typename F>
static bool throws_with(F f, const char *needle) {
    try {
        f();
    } catch (const std::runtime_error &e) {
        return std::strstr(e.what(), needle) != nullptr;
    }
    return false;
}

using namespace helios;

// four primitives as RadiationModel::buildGeometryData lays them out (R/src/RadiationModel.cpp:5890-6040):
// UUIDs 10..13 at positions 0..3: patch, triangle, tile, patch; twosided 1, 0, 2, 1; masks on 1 and 3
static RayTracingGeometry make_geometry() {
    RayTracingGeometry g;
    const uint P = 4;
    g.primitive_count = P;
    const uint types[P] = {0, 1, 3, 0};
    for (uint p = 0; p < P; p++) {
        for (int k = 0; k < 16; k++) g.transform_matrices.push_back(k % 5 == 0 ? 1.f : 0.01f * (float)(16 * p + k));
        g.primitive_types.push_back(types[p]);
        g.primitive_UUIDs.push_back(10 + p);
        g.primitive_IDs.push_back(0);
        g.object_IDs.push_back(p);
        g.object_subdivisions.push_back(make_int2(1, 1));
        g.solid_fractions.push_back(1.f);
    }
    g.twosided_flags = {1, 0, 2, 1};
    g.primitive_positions.assign(14, UINT_MAX);
    for (uint p = 0; p < P; p++) g.primitive_positions[10 + p] = p;
    for (int v = 0; v < 4; v++) g.patches.vertices.push_back(make_vec3((float)v, 0.f, 0.f));
    for (int v = 0; v < 4; v++) g.patches.vertices.push_back(make_vec3((float)v, 3.f, 0.f));
    g.patches.UUIDs = {10, 13};
    g.patch_count = 2;
    for (int v = 0; v < 3; v++) g.triangles.vertices.push_back(make_vec3((float)v, 1.f, 0.f));
    g.triangles.UUIDs = {11};
    g.triangle_count = 1;
    for (int v = 0; v < 4; v++) g.tiles.vertices.push_back(make_vec3((float)v, 2.f, 0.f));
    g.tiles.UUIDs = {12};
    g.tile_count = 1;
    g.periodic_flag = make_vec2(1.f, 0.f);
    // masks: 2x2 (id 0) on the triangle, 3x1 (id 1) on the last patch; compact uv_data holds the four UVs
    // of each primitive with uv_IDs >= 0, in primitive order
    g.mask_sizes = {make_int2(2, 2), make_int2(3, 1)};
    g.mask_data = {true, false, false, true, false, true, true};
    g.mask_IDs = {-1, 0, -1, 1};
    g.uv_IDs = {-1, 0, -1, 4};
    for (int v = 0; v < 8; v++) g.uv_data.push_back(make_vec2(0.1f * (float)v, 1.f - 0.1f * (float)v));
    return g;
}

static RayTracingLaunchParams launch(uint Bg, uint Bl, std::vector<bool> flags, uint face = 1) {
    RayTracingLaunchParams p;
    p.launch_offset = 0;
    p.launch_count = 4;
    p.rays_per_primitive = 100;
    p.random_seed = 77;
    p.num_bands_global = Bg;
    p.num_bands_launch = Bl;
    p.band_launch_flag = std::move(flags);
    p.launch_face = face;
    return p;
}

int main() {
    FakeAbi &F = fake();
    HipRayTracingBackend be(3);
    RayTracingBackend *rb = &be;

    // nothing reaches the C ABI before geometry exists (VulkanComputeBackend.cpp:592-598, 1766-1768)
    rb->initialize();
    CHECK(F.device == 3 && F.calls.size() == 1 && F.calls[0] == "rt_create");
    RayTracingMaterial m0;
    rb->updateMaterials(m0);
    rb->launchDirectRays(launch(1, 1, {true}));
    CHECK(F.calls.size() == 1 && F.direct.empty());

    // geometry
    RayTracingGeometry g = make_geometry();
    rb->updateGeometry(g);
    CHECK(F.P == 4 && F.px == 1 && F.py == 0);
    CHECK((F.types == std::vector<uint32_t>{0, 1, 3, 0}));
    CHECK((F.uuids == std::vector<uint32_t>{10, 11, 12, 13}));
    CHECK((F.twosided == std::vector<uint32_t>{1, 0, 2, 1}));
    CHECK(F.transforms == g.transform_matrices);
    // masks: bytes, sizes, ids and the compact UVs expanded to [P][4][2]
    CHECK(F.nmasks == 2 && (F.sizes == std::vector<uint32_t>{2, 2, 3, 1}));
    CHECK((F.texels == std::vector<uint8_t>{1, 0, 0, 1, 0, 1, 1}));
    CHECK((F.mask_ids == std::vector<int32_t>{-1, 0, -1, 1}) && (F.uv_ids == std::vector<int32_t>{-1, 0, -1, 4}));
    for (int p = 0; p < 4; p++)
        for (int v = 0; v < 4; v++) {
            const bool has = p == 1 || p == 3;
            const int src = p == 1 ? v : 4 + v;
            CHECK(F.uv[8 * p + 2 * v] == (has ? 0.1f * (float)src : 0.f));
            CHECK(F.uv[8 * p + 2 * v + 1] == (has ? 1.f - 0.1f * (float)src : 0.f));
        }
    // a geometry without masks sends none
    RayTracingGeometry g2 = make_geometry();
    g2.mask_IDs.assign(4, -1);
    F.nmasks = 99;
    rb->updateGeometry(g2);
    CHECK(F.nmasks == 99 && F.calls.back() == "rt_update_geometry");
    rb->updateGeometry(g);
    rb->buildAccelerationStructure();
    CHECK(F.calls.back() == "rt_build_acceleration_structure");

    // materials [S][P][B]: forwarded; a glass cover is rejected by name
    RayTracingMaterial mat;
    mat.num_sources = 1;
    mat.num_primitives = 4;
    mat.num_bands = 3;
    for (int k = 0; k < 12; k++) {
        mat.reflectivity.push_back(0.01f * (float)k);
        mat.transmissivity.push_back(0.02f * (float)k);
    }
    mat.is_glass.assign(12, 0);
    rb->updateMaterials(mat);
    CHECK(F.mat_S == 1 && F.mat_P == 4 && F.mat_B == 3 && F.rho == mat.reflectivity && F.tau == mat.transmissivity);
    CHECK(F.spec_exp.empty());  // no specular fields, no specular call
    RayTracingMaterial shiny = mat;
    shiny.specular_exponent = {50.f, -1.f, 10.f, 0.f};
    shiny.specular_scale = {0.5f, 0.f, 1.f, 2.f};
    rb->updateMaterials(shiny);
    CHECK(F.spec_exp == shiny.specular_exponent && F.spec_scale == shiny.specular_scale);
    shiny.specular_scale.resize(3);
    CHECK(throws_with([&] { rb->updateMaterials(shiny); }, "specular"));
    // translucent covers: the source-0 slice reaches rt_update_glass_materials; no glass clears the covers
    CHECK(F.is_glass.empty() && F.calls.back() == "rt_update_glass_materials");
    RayTracingMaterial glass = mat;
    glass.is_glass[5] = 1;
    glass.glass_n.assign(12, 1.5f);
    glass.glass_KL.assign(12, 0.1f);
    rb->updateMaterials(glass);
    CHECK(F.is_glass.size() == 12 && F.is_glass[5] == 1 && F.is_glass[4] == 0 && F.glass_n[5] == 1.5f && F.glass_KL[11] == 0.1f);
    glass.glass_n.resize(4);
    CHECK(throws_with([&] { rb->updateMaterials(glass); }, "glass_n"));

    // sources: flattened in order
    std::vector<RayTracingSource> src(2);
    src[0].type = 0;
    src[0].position = make_vec3(0.f, 0.f, 1.f);
    src[1].type = 3;
    src[1].position = make_vec3(1.f, 2.f, 3.f);
    src[1].width = make_vec2(0.5f, 0.25f);
    src[1].rotation = make_vec3(0.1f, 0.2f, 0.3f);
    rb->updateSources(src);
    CHECK((F.src_types == std::vector<uint32_t>{0, 3}));
    CHECK((F.src_pos == std::vector<float>{0.f, 0.f, 1.f, 1.f, 2.f, 3.f}));
    CHECK((F.src_width == std::vector<float>{0.f, 0.f, 0.5f, 0.25f}));
    CHECK((F.src_rot == std::vector<float>{0.f, 0.f, 0.f, 0.1f, 0.2f, 0.3f}));

    // band table: 3 global bands, a launch of bands 0 and 2
    rb->zeroRadiationBuffers(2);
    CHECK(F.launch_bands == 2);
    rb->uploadSourceFluxes({5.f, 6.f, 7.f, 8.f});
    CHECK((F.fluxes == std::vector<float>{5.f, 6.f, 7.f, 8.f}));
    rb->launchDirectRays(launch(3, 2, {true, false, true}));
    CHECK(F.direct.size() == 1);
    {
        const FakeLaunch &l = F.direct.back();
        CHECK((l.band_map == std::vector<uint32_t>{0, 2}) && (l.band_active == std::vector<uint8_t>{1, 1}));
        CHECK(l.p.launch_count == 4 && l.p.rays_per_primitive == 100 && l.p.random_seed == 77);
        CHECK(l.p.num_bands_global == 3 && l.p.num_bands_launch == 2 && !l.has_out);
    }
    // primary diffuse pass of the top face with sky flux and a power-law sky
    RayTracingLaunchParams d = launch(3, 2, {true, false, true}, 1);
    d.diffuse_flux = {50.f, 60.f};
    d.diffuse_extinction = {0.5f, 0.f};
    d.diffuse_dist_norm = {1.2f, 1.f};
    d.diffuse_peak_dir = {make_vec3(0.f, 0.f, 1.f), make_vec3(0.f, 1.f, 0.f)};
    d.radiation_out_top.assign(8, 1.f);
    d.radiation_out_bottom.assign(8, 2.f);
    rb->launchDiffuseRays(d);
    {
        const FakeLaunch &l = F.diffuse.back();
        CHECK(l.p.launch_face == 1 && l.has_out);
        CHECK((l.diffuse_flux == std::vector<float>{50.f, 60.f}));
        CHECK((l.extinction == std::vector<float>{0.5f, 0.f}) && (l.dist_norm == std::vector<float>{1.2f, 1.f}));
        CHECK((l.peak_dir == std::vector<float>{0.f, 0.f, 1.f, 0.f, 1.f, 0.f}));
        CHECK((l.band_active == std::vector<uint8_t>{1, 1}));
    }
    // a scattering iteration after band 2 ran out of depth (RadiationModel.cpp:4255-4277): the launch keeps
    // two bands, band 2 keeps its slot and takes no energy
    rb->copyScatterToRadiation();
    rb->zeroScatterBuffers();
    rb->launchDiffuseRays(launch(3, 2, {true, false, false}, 0));
    {
        const FakeLaunch &l = F.diffuse.back();
        CHECK((l.band_map == std::vector<uint32_t>{0, 2}) && (l.band_active == std::vector<uint8_t>{1, 0}));
        CHECK(l.p.launch_face == 0 && l.diffuse_flux.empty() && l.peak_dir.empty());
    }
    // flags that select more bands than the launch holds
    CHECK(throws_with([&] { rb->launchDiffuseRays(launch(3, 2, {true, true, true})); }, "selects 3 bands"));
    // a new runBand resets the table; no flags = identity map
    rb->zeroRadiationBuffers(2);
    rb->launchDirectRays(launch(2, 2, {}));
    CHECK((F.direct.back().band_map == std::vector<uint32_t>{0, 1}));
    // the retired-band form needs a table from an earlier launch of the same runBand
    rb->zeroRadiationBuffers(2);
    CHECK(throws_with([&] { rb->launchDirectRays(launch(3, 2, {true, false, false})); }, "selects 1 bands"));
    rb->zeroRadiationBuffers(2);
    // specular reflection: the mode travels with direct and camera launches, the camera weights and the
    // per-primitive exponent / scale reach their entry points
    RayTracingLaunchParams sp = launch(3, 2, {true, false, true});
    sp.specular_reflection_enabled = 2;
    rb->launchDirectRays(sp);
    CHECK(F.direct.back().p.specular_reflection_enabled == 2);
    rb->launchCameraRays(sp);
    CHECK(F.camera.back().specular_reflection_enabled == 2);
    rb->uploadSourceFluxesCam({0.5f, 0.25f});
    CHECK((F.fluxes_cam == std::vector<float>{0.5f, 0.25f}));
    rb->uploadSourceFluxesCam({});
    CHECK(F.calls.back() == "rt_upload_source_fluxes_cam" && F.fluxes_cam.size() == 2);

    // results: [P][B_launch] copies; the camera scatter buffers equal the scatter buffers
    rb->uploadRadiationOut(std::vector<float>(8, 3.f), std::vector<float>(8, 4.f));
    CHECK(F.out_top.size() == 8 && F.out_bottom[7] == 4.f);
    RayTracingResults r;
    rb->getRadiationResults(r);
    CHECK(r.num_primitives == 4 && r.num_bands == 2 && r.radiation_in.size() == 8);
    CHECK(r.radiation_in[7] == 7.f && r.radiation_out_top[3] == 1003.f && r.radiation_out_bottom[0] == 2000.f);
    CHECK(r.scatter_buff_top[5] == 3005.f && r.scatter_buff_bottom[6] == 4006.f);
    CHECK(r.scatter_buff_top_cam == r.scatter_buff_top && r.scatter_buff_bottom_cam == r.scatter_buff_bottom);
    CHECK(throws_with([&] { rb->uploadRadiationOut(std::vector<float>(8), std::vector<float>(7)); }, "differ"));

    // camera: the launch parameters map field by field
    RayTracingLaunchParams c = launch(3, 2, {true, false, true});
    c.camera_position = make_vec3(1.f, 2.f, 3.f);
    c.camera_direction = make_vec2(0.25f, 1.5f);
    c.camera_focal_length = 2.f;
    c.camera_lens_diameter = 0.05f;
    c.camera_fov_aspect = 1.5f;
    c.camera_viewplane_length = 0.7f;
    c.camera_resolution = make_int2(16, 8);
    c.camera_resolution_full = make_int2(32, 8);
    c.camera_pixel_offset = make_int2(16, 0);
    c.antialiasing_samples = 4;
    rb->zeroCameraPixelBuffers(make_int2(32, 8));
    rb->launchCameraRays(c);
    rb->launchPixelLabelRays(c);
    {
        const rt_camera_params &k = F.camera.back();
        CHECK(k.position[0] == 1.f && k.position[2] == 3.f && k.direction[0] == 0.25f && k.direction[1] == 1.5f);
        CHECK(k.focal_length == 2.f && k.lens_diameter == 0.05f && k.fov_aspect == 1.5f && k.viewplane_length == 0.7f);
        CHECK(k.resolution[0] == 16 && k.resolution[1] == 8 && k.resolution_full[0] == 32 && k.pixel_offset[0] == 16);
        CHECK(k.antialiasing_samples == 4 && k.random_seed == 77 && k.specular_reflection_enabled == 0);
        CHECK(F.label.size() == 1 && F.label.back().resolution[0] == 16);
    }
    std::vector<float> px, dep;
    std::vector<uint> lab;
    rb->getCameraResults(px, lab, dep, 0, make_int2(4, 2));
    CHECK(px.size() == 16 && lab.size() == 8 && dep.size() == 8 && lab[7] == 8u && dep[2] == 1.f && px[3] == 1.f);

    // a C ABI failure is rethrown with its message
    F.fail_next = 1;
    F.fail_msg = "device lost";
    CHECK(throws_with([&] { rb->buildAccelerationStructure(); }, "device lost"));

    rb->shutdown();
    CHECK(F.calls.back() == "rt_destroy");
    if (g_fail == 0) std::printf("adapter: all checks passed (%zu C ABI calls)\n", F.calls.size());
    return g_fail == 0 ? 0 : 1;
}
