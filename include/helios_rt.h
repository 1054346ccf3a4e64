/*
 * helios_rt.h -- C ABI of the MI355X (gfx950) radiation-transfer backend.
 *
 * Plain C: opaque handles, POD arrays, sizes, int status (0 = success, -1 = failure with the
 * message in rt_last_error()).  No C++ or torch types cross this boundary.  Each entry point names
 * the reference interface it replaces (Helios tree: R = plugins/radiation, CD = plugins/collisiondetection).
 *
 * Conventions shared with the reference:
 *   - primitive arrays are in the order the caller passes them; that position is the "primitive
 *     index" that seeds the per-ray RNG (R/shaders/direct_raygen.comp:360-362);
 *   - transforms are 4x4 row-major; canonical patch (-0.5..0.5)^2, canonical triangle
 *     (0,0,0),(0,1,0),(1,1,0) (R/src/backends/VulkanComputeBackend.cpp:328-329);
 *   - types: 0 patch, 1 triangle, 3 tile (traced as patch), 4 voxel; scene creation and
 *     rt_update_geometry reject any other type (disks, 2, included) with an error naming it, so no primitive
 *     silently drops out of a scene (the Vulkan backend traces patch/tile/triangle only,
 *     R/shaders/common/bvh_traversal.glsl:311-325; the OptiX 8 backend rejects disks and voxels,
 *     R/src/optix8/OptiX8DeviceCode.cu:492-504);
 *   - a voxel is the axis-aligned box of its eight transformed unit-cube corners: castRays hits it as
 *     CollisionDetection does (CD/src/CollisionDetection_RayTracing.cpp:1625-1711); the radiation passes,
 *     cameras and pixel labels neither hit it nor launch rays from it (the Vulkan traversal never hits
 *     one, bvh_traversal.glsl:118-140, and its ray generation has no sample point for one,
 *     R/shaders/direct_raygen.comp:417-426), so its radiation_flux is 0;
 *   - at most 2^29 primitives per scene; one direct pass covers at most 2^32 - 2^25 (sun, source,
 *     primitive) work items, and larger requests fail with an error saying so;
 *   - twosided: raw Helios flag 0..3 (R/src/RadiationModel.cpp:5941);
 *   - radiation buffers are [primitive][band] (R/include/BufferIndexing.h:65-262).
 */
#ifndef HELIOS_RT_H
#define HELIOS_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rt_ctx rt_ctx;
typedef struct rt_scene rt_scene;

#define RT_MISS_UUID 0xFFFFFFFFu

/* Message of the last failed call on this thread. */
const char *rt_last_error(void);

/* ---------------------------------------------------------------------------------------------
 * Context.  Replaces RayTracingBackend::create("auto") + initialize()
 * (R/src/backends/BackendFactory.cpp:46-105, R/include/RayTracingBackend.h:54) and shutdown()
 * (RayTracingBackend.h:61).  Fails (returns -1) when no gfx950 device is usable: there is no CPU
 * fallback.
 * ------------------------------------------------------------------------------------------- */
int rt_create(int device, rt_ctx **out);
int rt_destroy(rt_ctx *ctx);
/* RayTracingBackend::getBackendName (RayTracingBackend.h:272) */
const char *rt_backend_name(const rt_ctx *ctx);
/* RayTracingBackend::queryGPUMemory (RayTracingBackend.h:265) */
int rt_query_memory(rt_ctx *ctx, size_t *free_bytes, size_t *total_bytes);

/* ---------------------------------------------------------------------------------------------
 * Resident scenes.  Replaces RayTracingBackend::updateGeometry + buildAccelerationStructure
 * (RayTracingBackend.h:73, :81; VulkanComputeBackend.cpp:213-590) and CollisionDetection::buildBVH
 * (CD/src/CollisionDetection.cpp:341-513): world vertices, normals and a Morton LBVH are built on
 * the GPU and stay in HBM until rt_scene_destroy.
 * ------------------------------------------------------------------------------------------- */
int rt_scene_create(rt_ctx *ctx, uint32_t nprims, const float *transforms, const uint32_t *types,
                    const uint32_t *uuids, const uint32_t *twosided, int periodic_x, int periodic_y,
                    rt_scene **out);
int rt_scene_destroy(rt_ctx *ctx, rt_scene *scene);

typedef struct {
    uint32_t nprims;       /* primitives passed in                        */
    uint32_t nleaves;      /* traced primitives                           */
    uint32_t depth;        /* deepest leaf of the BVH                     */
    uint32_t nbbox;        /* periodic boundary faces                     */
    float domain[4];       /* xmin, xmax, ymin, ymax of the periodic faces */
    uint64_t device_bytes; /* resident bytes                              */
    uint32_t builder;      /* 1 = PLOC tree, 0 = Morton LBVH (Karras)     */
    uint32_t depth4;       /* levels of the 4-wide tree, 0 if not used    */
} rt_scene_info;
int rt_scene_get_info(const rt_scene *scene, rt_scene_info *info);

/* Texture masks (RayTracingGeometry mask_data / mask_sizes / mask_IDs / uv_data / uv_IDs,
 * R/include/RayTracingTypes.h:72-77; R/shaders/common/texture_mask.glsl:20-66).  nmasks masks of
 * sizes[m] = (width, height) texels, one byte per texel (0 = transparent), concatenated row-major from
 * the top row in `texels`; mask_ids[P] (-1 = none); uv[P][4][2] per-primitive vertex UVs (NULL = zeros)
 * and uv_ids[P] (NULL or < 0: a patch uses its parametric UVs).  A transparent texel at a hit lets the
 * ray pass; direct and diffuse ray origins are redrawn off transparent texels (direct_raygen.comp:431-473,
 * diffuse_raygen.comp:411-455).  Replaces the scene's masks; nmasks = 0 removes them.  castRays ignores
 * masks, as CollisionDetection does. */
int rt_scene_set_masks(rt_ctx *ctx, rt_scene *scene, uint32_t nmasks, const uint8_t *texels, const uint32_t *sizes,
                       const int32_t *mask_ids, const float *uv, const int32_t *uv_ids);

/* ---------------------------------------------------------------------------------------------
 * CollisionDetection::castRaysSoA (CD/include/CollisionDetection.h:395,
 * CD/src/CollisionDetection_RayTracing.cpp:551-680): closest hit per ray with the CD predicates.
 * origins/dirs [n][3] (directions are not normalised, as on the reference CPU path); max_dist <= 0
 * means unbounded; a miss writes RT_MISS_UUID and leaves out_t / out_normal untouched.
 * ------------------------------------------------------------------------------------------- */
int rt_cast_rays_soa(rt_ctx *ctx, const rt_scene *scene, const float *origins, const float *dirs, size_t n,
                     float max_dist, float *out_t, float *out_normal, uint32_t *out_uuid);
/* CollisionDetection::castRaysSoA_packets (CD/include/CollisionDetection.h:395-422): rays grouped in
 * pulses of packet_size contiguous sub-rays; results identical to rt_cast_rays_soa, and the same kernel,
 * as the reference's own GPU-resident path (CD/src/CollisionDetection_RayTracing.cpp:714-720). */
int rt_cast_rays_soa_packets(rt_ctx *ctx, const rt_scene *scene, const float *origins, const float *dirs, size_t n,
                             size_t packet_size, float max_dist, float *out_t, float *out_normal, uint32_t *out_uuid);
/* LiDARcloud::castRaysUnified (plugins/lidar/src/LiDAR.cpp:558-598), the LiDAR synthetic scan's ray cast:
 * castRaysSoA with max distance miss_t, then per ray hit_id = UUID and hit_fnorm = dot(dir, normal) on a
 * hit; hit_t = miss_t, hit_id = -1, hit_fnorm = 1e6 on a miss.  The conversion runs on the device. */
int rt_lidar_cast_rays(rt_ctx *ctx, const rt_scene *scene, const float *origins, const float *dirs, size_t n,
                       size_t packet_size, float miss_t, float *hit_t, float *hit_fnorm, int32_t *hit_id);
/* rt_cast_rays_soa with every array already in device memory; runs on the context stream and waits for it. */
int rt_cast_rays_soa_device(rt_ctx *ctx, const rt_scene *scene, const float *d_origins, const float *d_dirs,
                            size_t n, float max_dist, float *d_out_t, float *d_out_normal, uint32_t *d_out_uuid);
/* Allocation-free LiDAR cast.  No counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  rt_lidar_cast_rays
 * with every array in device memory (origins/dirs [n][3]; hit_t, hit_fnorm, hit_id [n]) and no
 * allocation of any kind: the castRays normals and UUIDs that the conversion reads, the work counter and
 * the error flag live in a fixed scratch array in the library's device data (16 bytes per ray plus 256,
 * i.e. 16 MiB + 256 B per device, held until the process ends).  At most RT_NOALLOC_MAX_RAYS rays per
 * call (larger n fails with an error; cast a longer scan in slices).  The scratch is shared by every
 * call on one device and keeps the last call's contents; the library serialises calls per device in
 * this process (concurrent calls wait), and a caller must not read the scratch.  Runs on the context
 * stream and waits for it.  Results equal rt_lidar_cast_rays bit for bit. */
#define RT_NOALLOC_MAX_RAYS (1u << 20)
int rt_lidar_cast_rays_noalloc(rt_ctx *ctx, const rt_scene *scene, const float *d_origins, const float *d_dirs,
                               size_t n, float miss_t, float *d_hit_t, float *d_hit_fnorm, int32_t *d_hit_id);

/* ---------------------------------------------------------------------------------------------
 * Whole runBand on the device.  Replaces the pass schedule of RadiationModel::runBand
 * (R/src/RadiationModel.cpp:3470-4672: direct pass, primary diffuse pass on both faces, scattering
 * iterations, final radiation_flux = radiation_in + scatter_top + scatter_bottom) with no host
 * round trips between passes.  Seeds are explicit (the reference draws them from the clock).
 * ------------------------------------------------------------------------------------------- */
#define RT_SCHEDULE_REFERENCE 0u  /* pass schedule exactly as RadiationModel.cpp:4236-4356          */
#define RT_SCHEDULE_CONSERVING 1u /* also propagates the primary pass's scatter when emission is off */

#define RT_SOURCE_COLLIMATED 0u
#define RT_SOURCE_SPHERE 1u
#define RT_SOURCE_SUN_SPHERE 2u
#define RT_SOURCE_RECTANGLE 3u
#define RT_SOURCE_DISK 4u

typedef struct {
    uint32_t nbands;             /* B, launched bands (alphabetical order)                         */
    uint32_t nsources;           /* S, radiation sources                                           */
    uint32_t direct_rays;        /* max directRayCount over the bands (d = ceil(sqrt))             */
    uint32_t diffuse_rays;       /* max diffuseRayCount over the bands (n = ceil(sqrt))            */
    const uint32_t *band_depth;  /* [B] scatteringDepth                                            */
    const uint32_t *emission;    /* [B] emission flag                                              */
    const float *source_dir;     /* [S][3] collimated: direction towards the source; other types:  */
                                 /* source position (RayTracingSource::position)                   */
    const float *source_flux;    /* [S][B]                                                         */
    const float *diffuse_flux;   /* [B]                                                            */
    const float *rho;            /* [P][B] reflectivity (validated, source-0 slice)                */
    const float *tau;            /* [P][B] transmissivity                                          */
    const float *eps;            /* [P][B] emissivity, NULL = 1                                    */
    const float *temperature;    /* [P] kelvin, NULL = 300                                         */
    uint32_t direct_seed;
    const uint32_t *diffuse_seed;/* [1 + max depth]: primary diffuse pass, then each scattering pass */
    uint32_t schedule;           /* RT_SCHEDULE_*                                                   */
    /* Source shapes (RayTracingTypes.h:189-196; direct_raygen.comp:491-561).  source_type NULL = every
     * source collimated.  Sphere / sun sphere: width[0] = diameter; rectangle: width = (w, h); disk:
     * width[0] = radius; rotation = ZYX Euler angles (radians) of a rectangle or disk. */
    const uint32_t *source_type; /* [S] RT_SOURCE_*, or NULL                                        */
    const float *source_width;   /* [S][2], required when a source is not collimated               */
    const float *source_rotation;/* [S][3], required when a source is a rectangle or disk          */
    /* Sky radiance distribution of the diffuse flux (R/shaders/common/sky_distribution.glsl:40-116,
     * diffuse_raygen.comp:613-625), per band, all NULL = isotropic: power law psi^-K * norm about the
     * peak direction when K > 0, else the Prague pattern when sky_params[b][3] (its normalisation) > 0. */
    const float *diffuse_extinction; /* [B] power-law exponent K                               */
    const float *diffuse_peak_dir;   /* [B][3] unit direction of the peak (the sun)            */
    const float *diffuse_dist_norm;  /* [B] power-law normalisation                            */
    const float *sky_params;         /* [B][4] Prague (circumsolar strength, width in degrees, */
                                     /* horizon brightness, normalisation)                     */
    /* Translucent covers (RayTracingMaterial is_glass / glass_n / glass_KL, R/include/RayTracingTypes.h:
     * 173-178; R/shaders/common/glass_cover.glsl:15-53, direct_raygen.comp:587-664,
     * diffuse_raygen.comp:484-560), [P][B] each, is_glass NULL = no covers.  A primitive that is glass in
     * every band of the request is a cover: rays pass through it, multiplying each band's flux by the
     * sheet's Fresnel + Bouguer transmittance at the crossing angle and pushing the absorbed and reflected
     * shares into the cover (direct rays: of the source flux; diffuse rays: of the sky flux); a sky ray
     * is attenuated by the covers it crossed, a surface hit behind a cover is not.  A solve with covers
     * takes at most RT_GLASS_MAX_CHANNELS channels (suns x bands). */
    const uint8_t *is_glass;
    const float *glass_n;            /* refractive index (>= 1)                                */
    const float *glass_KL;           /* absorption product K*L, NULL = 0                       */
} rt_band_request;
#define RT_GLASS_MAX_CHANNELS 8u

/* R_out: [P][B] radiation_flux_<band> (host memory). */
int rt_run_band(rt_ctx *ctx, const rt_scene *scene, const rt_band_request *req, float *R_out);

/* Batched solves of one scene: K suns (own directions, fluxes, sky flux and direct seed) share
 * the bands, materials and diffuse seeds of `base`, so every diffuse/scattering pass is traced once
 * and gathered for all K*B channels.  Each sun's result equals rt_run_band with that sun alone.
 * base->source_dir / source_flux / diffuse_flux / direct_seed are ignored. */
typedef struct {
    uint32_t nsuns;              /* K                                      */
    const float *source_dir;     /* [K][S][3]                              */
    const float *source_flux;    /* [K][S][B]                              */
    const float *diffuse_flux;   /* [K][B]                                 */
    const uint32_t *direct_seed; /* [K]                                    */
    const float *diffuse_peak_dir; /* [K][B][3] sky peak of each sun, NULL = base's */
} rt_sun_batch;
/* R_out: [K][P][B]; if out_on_device != 0, R_out is a device pointer. */
int rt_run_band_batch(rt_ctx *ctx, const rt_scene *scene, const rt_band_request *base, const rt_sun_batch *suns,
                      float *R_out, int out_on_device);

/* FP64 energy bookkeeping of the scattering iterations (no reference counterpart: the reference checks
 * conservation only in its closed-form tests).  Off by default; when on, every rt_run_band /
 * rt_run_band_batch on the context also records, per channel (channel = sun*B + band), area-weighted
 * FP64 totals in W (flux density x m^2):
 *   row 0 source   direct incident flux of the sources;
 *   row 1 sky      sky flux arriving at primitives in the primary diffuse pass;
 *   row 2 emitted  sigma eps T^4 from every radiating face (the top, and the bottom of a two-sided one);
 *   row 3 absorbed sum of radiation_flux (includes row 5);
 *   row 4 escaped  flux handed on by a pass that reaches the sky, summed over the passes (each primitive's
 *                  radiated flux times its own rays' Monte Carlo estimate of its view factor to the sky);
 *   row 5 leftover scatter still undelivered after the last pass (the reference deposits it as absorbed);
 *   row 6 residual (source + sky + emitted) - (absorbed + escaped).
 * The residual is zero up to Monte Carlo error when the schedule hands every scattered flux on exactly once
 * (RT_SCHEDULE_CONSERVING); under RT_SCHEDULE_REFERENCE with direct rays, diffuse/emission off and
 * scattering it measures the flux the reference gathers twice and drops (RadiationModel.cpp:4286-4289).
 * Costs one extra [P][C] x 8-byte buffer and a reduction per pass. */
#define RT_ENERGY_ROWS 7
#define RT_ENERGY_PASS_ROWS 4
int rt_set_energy_accounting(rt_ctx *ctx, int on);
/* out: [RT_ENERGY_ROWS][nchannels] of the last solve (fails unless it ran with the bookkeeping on). */
int rt_energy_residual_fp64(rt_ctx *ctx, double *out, uint32_t nchannels);
/* Per diffuse/scattering pass g in run order (the primary pass first): out[g][row][nchannels] with rows
 * 0 radiated (flux handed on by the launching faces), 1 arrived (gathered from surfaces), 2 sky (arrived
 * from the sky), 3 escaped; radiated = arrived + escaped up to Monte Carlo error.  Writes at most
 * max_passes passes (out may be NULL) and the number of passes to *npasses. */
int rt_energy_passes_fp64(rt_ctx *ctx, double *out, uint32_t nchannels, uint32_t max_passes, uint32_t *npasses);

/* ---------------------------------------------------------------------------------------------
 * Call-by-call backend: helios::RayTracingBackend (R/include/RayTracingBackend.h:43-312) as C
 * functions over one resident scene owned by the context, with the reference backend's five
 * [P][B_launch] buffers (radiation_in, scatter_top, scatter_bottom, radiation_out_top,
 * radiation_out_bottom) kept in HBM between calls.  A RadiationModel that keeps its own pass loop
 * (R/src/RadiationModel.cpp:3804-4356) calls these where it calls VulkanComputeBackend.  Launches
 * return after their kernels finish (traversal errors are reported by the launch call).
 * ------------------------------------------------------------------------------------------- */
/* updateGeometry (RayTracingBackend.h:73; VulkanComputeBackend.cpp:213): stages the primitive
 * arrays, same meaning as rt_scene_create; uuids NULL = 0..P-1. */
int rt_update_geometry(rt_ctx *ctx, uint32_t nprims, const float *transforms, const uint32_t *types,
                       const uint32_t *uuids, const uint32_t *twosided, int periodic_x, int periodic_y);
/* The texture-mask fields of updateGeometry's RayTracingGeometry (as rt_scene_set_masks, with the
 * reference's compact uv_data expanded to [P][4][2]); staged, applied by the next build.  Call after
 * rt_update_geometry, which clears staged masks. */
int rt_update_texture_masks(rt_ctx *ctx, uint32_t nprims, uint32_t nmasks, const uint8_t *texels,
                            const uint32_t *sizes, const int32_t *mask_ids, const float *uv, const int32_t *uv_ids);
/* buildAccelerationStructure (RayTracingBackend.h:81): GPU LBVH of the staged geometry. */
int rt_build_acceleration_structure(rt_ctx *ctx);
/* updateMaterials (RayTracingBackend.h:90; R/include/RayTracingTypes.h:157-184): rho, tau are
 * [S][P][B_global]; the source-0 slice serves every source, as in the Vulkan backend's shaders
 * (R/shaders/direct_raygen.comp:684-687, diffuse_raygen.comp:575-579). */
int rt_update_materials(rt_ctx *ctx, uint32_t nsources, uint32_t nprims, uint32_t nbands_global, const float *rho,
                        const float *tau);
/* updateMaterials' translucent-cover fields (RayTracingTypes.h:173-178; VulkanComputeBackend.cpp:656-690):
 * is_glass, glass_n (refractive index >= 1), glass_KL (absorption K*L >= 0, NULL = 0), each
 * [P][B_global] (the source-0 slice); call after rt_update_materials (which clears them), is_glass NULL =
 * no covers.  A launch
 * passes rays through the primitives that are glass in every launch band (at most 8 launch bands then). */
int rt_update_glass_materials(rt_ctx *ctx, uint32_t nprims, uint32_t nbands_global, const uint8_t *is_glass,
                              const float *glass_n, const float *glass_KL);
/* updateSources (RayTracingBackend.h:99; RayTracingTypes.h:189-196): types[S] (RT_SOURCE_*),
 * positions[S][3] (direction towards a collimated source, else the source position), widths[S][2] and
 * rotations[S][3] (RayTracingSource::width / ::rotation; NULL when every source is collimated). */
int rt_update_sources(rt_ctx *ctx, uint32_t nsources, const uint32_t *types, const float *positions,
                      const float *widths, const float *rotations);
/* zeroRadiationBuffers (RayTracingBackend.h:194; VulkanComputeBackend.cpp:1935): sets B_launch
 * (1..256) and zeroes all five radiation buffers. */
int rt_zero_radiation_buffers(rt_ctx *ctx, uint32_t launch_band_count);
/* zeroScatterBuffers (RayTracingBackend.h:199) */
int rt_zero_scatter_buffers(rt_ctx *ctx);
/* copyScatterToRadiation (RayTracingBackend.h:218): radiation_out_top/bottom := scatter_top/bottom */
int rt_copy_scatter_to_radiation(rt_ctx *ctx);
/* uploadRadiationOut (RayTracingBackend.h:228): count = P*B_launch values per face */
int rt_upload_radiation_out(rt_ctx *ctx, const float *top, const float *bottom, size_t count);
/* uploadSourceFluxes (RayTracingBackend.h:247): [S][B_launch] */
int rt_upload_source_fluxes(rt_ctx *ctx, const float *fluxes, size_t count);

/* updateMaterials' specular fields (RayTracingTypes.h:170-171; VulkanComputeBackend.cpp:631-655):
 * specular_exponent [P] (not positive = no specular term) and specular_scale [P] (NULL = 0), read by
 * camera launches with specular_reflection_enabled. */
int rt_update_specular_materials(rt_ctx *ctx, uint32_t nprims, const float *exponent, const float *scale);
/* uploadSourceFluxesCam (RayTracingBackend.h:250-258; VulkanComputeBackend.cpp:2220-2237): the camera's
 * spectral weight of each source and launch band, [S][B_launch]; needed by camera launches with
 * specular on. */
int rt_upload_source_fluxes_cam(rt_ctx *ctx, const float *fluxes_cam, size_t count);
/* RayTracingResults::radiation_specular (RayTracingTypes.h:265) of this backend's single camera slot:
 * [S][P][B_launch] host memory, count = S*P*B_launch; zeros when no direct launch with specular on ran
 * since rt_zero_radiation_buffers. */
int rt_get_radiation_specular(rt_ctx *ctx, float *out, size_t count);

/* RayTracingLaunchParams (RayTracingTypes.h:201-249), radiation fields.  The reference passes a
 * [B_global] band_launch_flag and derives the launch -> global band map from it; here the map is
 * explicit and a launch band that stops scattering keeps its slot with band_active = 0 (the
 * caller-side adapter builds both from band_launch_flag). */
typedef struct {
    uint32_t launch_offset;        /* first primitive launched                                */
    uint32_t launch_count;         /* primitives launched                                     */
    uint32_t rays_per_primitive;   /* rays per primitive face (sqrt rounded up per side)      */
    uint32_t random_seed;
    uint32_t num_bands_global;
    uint32_t num_bands_launch;
    const uint32_t *band_map;      /* [B_launch] global band of each launch band, NULL = identity */
    const uint8_t *band_active;    /* [B_launch] 0 = band takes no energy in this launch, NULL = all */
    uint32_t launch_face;          /* diffuse launches: 1 = top face, 0 = bottom face         */
    const float *diffuse_flux;     /* [B_launch] sky flux of a diffuse launch, NULL = 0       */
    const float *radiation_out_top;    /* [P][B_launch] uploaded before a diffuse launch, or NULL */
    const float *radiation_out_bottom; /* [P][B_launch], NULL iff radiation_out_top is NULL      */
    /* anisotropic sky of a diffuse launch (RayTracingLaunchParams diffuse_extinction / _peak_dir /
     * _dist_norm), [B_launch] each, NULL = 0; the Prague parameters come from rt_update_sky_model, whose
     * sun direction is the peak when diffuse_peak_dir is NULL.  Bands without sky flux ignore them. */
    const float *diffuse_extinction;
    const float *diffuse_peak_dir;     /* [B_launch][3] */
    const float *diffuse_dist_norm;
    /* direct launches (RayTracingLaunchParams::specular_reflection_enabled): non-zero adds each source's
     * flux x strength over the unoccluded rays to radiation_specular [S][P][B_launch]
     * (R/shaders/direct_raygen.comp:717-726), which camera launches with specular on read */
    uint32_t specular_reflection_enabled;
} rt_launch_params;
/* launchDirectRays (RayTracingBackend.h:139; R/shaders/direct_raygen.comp): every source, then
 * radiation_in / scatter_top / scatter_bottom += the deposits. */
int rt_launch_direct_rays(rt_ctx *ctx, const rt_launch_params *params);
/* launchDiffuseRays (RayTracingBackend.h:146; R/shaders/diffuse_raygen.comp): one face. */
int rt_launch_diffuse_rays(rt_ctx *ctx, const rt_launch_params *params);
/* getRadiationResults (RayTracingBackend.h:171): each output [P][B_launch] host memory, NULL skips. */
int rt_get_radiation_results(rt_ctx *ctx, float *radiation_in, float *radiation_out_top, float *radiation_out_bottom,
                             float *scatter_top, float *scatter_bottom);
/* Backend state: out[4] = primitives of the built scene, B_global, B_launch, sources. */
int rt_get_backend_info(rt_ctx *ctx, uint32_t *out);

/* Cameras: RayTracingLaunchParams' camera fields (RayTracingTypes.h:221-235). */
typedef struct {
    float position[3];
    float direction[2];              /* elevation, azimuth (radians)                              */
    float focal_length;              /* working distance                                          */
    float lens_diameter;             /* 0 = pinhole                                               */
    float fov_aspect;                /* HFOV / VFOV                                               */
    float viewplane_length;          /* from the effective HFOV and zoom                          */
    uint32_t resolution[2];          /* this tile                                                 */
    uint32_t resolution_full[2];     /* whole image                                               */
    uint32_t pixel_offset[2];        /* tile origin in the whole image                            */
    uint32_t antialiasing_samples;   /* 0 is taken as 1                                           */
    uint32_t random_seed;
    uint32_t specular_reflection_enabled; /* 0 off; 1 Blinn-Phong term with scale 1; 2 with the
                                             primitives' specular_scale (camera_raygen.comp:362-409) */
} rt_camera_params;
/* updateSkyModel (RayTracingBackend.h:129): per launch band, sky_radiance_params [B][4] (Prague
 * circumsolar strength, circumsolar width in degrees, horizon brightness, normalisation; a
 * normalisation <= 0 means an isotropic sky), camera_sky_radiance [B], solar_disk_radiance [B],
 * camera_diffuse_flux [B] and band_emission_flag [B] (both NULL = 0); sun_direction[3] unit. */
int rt_update_sky_model(rt_ctx *ctx, uint32_t nbands, const float *sky_radiance_params,
                        const float *camera_sky_radiance, const float *sun_direction, const float *solar_disk_radiance,
                        float solar_disk_cos_angle, const float *camera_diffuse_flux,
                        const uint32_t *band_emission_flag);
/* launchCameraRays (RayTracingBackend.h:153; R/shaders/camera_raygen.comp): adds one tile's pixels
 * to camera_radiation ([B_launch][full pixels]), which the tile at offset (0, 0) zeroes first.  Hit
 * pixels read radiation_out_top/bottom (the last passes' emitted and scattered radiation), misses
 * the sky model; sources in view add their radiance.  Each pixel sums its samples in order. */
int rt_launch_camera_rays(rt_ctx *ctx, const rt_camera_params *params);
/* zeroCameraPixelBuffers (RayTracingBackend.h:210): sizes and zeroes labels and depths. */
int rt_zero_camera_pixel_buffers(rt_ctx *ctx, uint32_t width, uint32_t height);
/* launchPixelLabelRays (RayTracingBackend.h:160; R/shaders/pixel_label_raygen.comp): the
 * pixel-centre ray's hit UUID + 1 (0 = sky) and its depth along the view axis (-1 = sky). */
int rt_launch_pixel_label_rays(rt_ctx *ctx, const rt_camera_params *params);
/* getCameraResults (RayTracingBackend.h:182): pixel_data [pixels][B_launch], labels [pixels],
 * depths [pixels] for a width x height image; buffers that hold another image read as zeros;
 * NULL skips an output. */
int rt_get_camera_results(rt_ctx *ctx, float *pixel_data, uint32_t *pixel_labels, float *pixel_depths, uint32_t width,
                          uint32_t height);

/* ---------------------------------------------------------------------------------------------
 * Request server (no reference counterpart: the reference solves one runBand per call).  Models
 * (a resident scene + bands, ray counts, depths, optical properties kept in HBM, diffuse seeds) are
 * registered once; each request is one sun for one model.  Worker slots (own stream + workspace)
 * take the oldest request together with every queued request of the same model (FIFO, up to
 * max_batch_suns), shrink that batch until its workspace fits the memory budget, and solve it as
 * one batched run; slots execute concurrently.  A request's result equals rt_run_band of that sun
 * alone with the model's settings.  A request whose single-sun footprint exceeds the budget fails.
 * The scene must outlive the server.
 * ------------------------------------------------------------------------------------------- */
typedef struct rt_server rt_server;
typedef struct {
    uint32_t slots;          /* concurrent batches (streams), 0 = 4, at most 16          */
    uint32_t max_batch_suns; /* requests per batch, 0 = 64 (also capped at 256 / bands) */
    uint64_t memory_budget;  /* device bytes for solve workspaces, 0 = half the free HBM */
} rt_server_cfg;
typedef struct {
    const float *source_dir;   /* [S][3] */
    const float *source_flux;  /* [S][B] */
    const float *diffuse_flux; /* [B]    */
    uint32_t direct_seed;
} rt_solve_sun;
typedef struct {
    uint64_t submitted, completed, failed;
    uint64_t batches;              /* batched solves run                                    */
    uint64_t batched_requests;     /* requests they carried                                 */
    uint32_t max_batch;            /* largest batch run                                     */
    uint32_t queued;               /* requests waiting now                                  */
    double p50_ms, p99_ms, mean_ms; /* submit -> result-on-host latency of completed requests */
    uint64_t memory_budget;
    uint64_t peak_workspace_bytes; /* high-water mark of slot workspaces + outputs          */
    uint64_t resident_model_bytes; /* optical properties of the registered models           */
} rt_server_stats;
int rt_server_create(rt_ctx *ctx, const rt_server_cfg *cfg, rt_server **out);
/* Finishes every queued request, then frees the server. */
int rt_server_destroy(rt_server *srv);
/* base: bands, ray counts, depths, emission, rho/tau/eps/temperature, translucent covers, diffuse
 * seeds, schedule; its source_dir / source_flux / diffuse_flux / direct_seed are ignored (they come per
 * request).  A model with covers is batched at most 8 channels (suns x bands) at a time. */
int rt_server_add_model(rt_server *srv, const rt_scene *scene, const rt_band_request *base, uint32_t *model_id);
/* R_out: [P][B] host memory, written before the request completes; keep it alive until rt_server_wait. */
int rt_server_submit(rt_server *srv, uint32_t model_id, const rt_solve_sun *sun, float *R_out, uint64_t *ticket);
/* Blocks until the request is done; returns its status (and message) and releases the ticket. */
int rt_server_wait(rt_server *srv, uint64_t ticket, double *latency_ms);
int rt_server_get_stats(rt_server *srv, rt_server_stats *out);
/* Device bytes (workspace + output) a server slot needs for a batch of nsuns suns of this request
 * over nprims primitives -- the figure admission compares with the budget.  Host-only (no device call);
 * fails, as the solve would, when nsuns x sources x nprims exceeds one launch's work-item limit. */
int rt_solve_footprint(uint32_t nprims, const rt_band_request *req, uint32_t nsuns, uint64_t *bytes);
int rt_server_reset_stats(rt_server *srv);

/* ---------------------------------------------------------------------------------------------
 * Launch tuning (no reference counterpart): waves per workgroup, persistent workgroups per CU and
 * work-queue chunking (primitives taken per queue fetch) of the direct, diffuse and gather kernels,
 * and the occlusion-grid density of the direct pass, set by hand or picked by rt_autotune, which
 * times candidates on a sample of the caller's own scene and request.  The traversal stack depth is
 * picked per launch from the BVH depth.
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    uint32_t waves_per_block;   /* 1, 2 or 4 (the gather kernels always use 4)  */
    uint32_t blocks_per_cu;     /* persistent workgroups per CU, 1..64          */
    uint32_t chunk;             /* work items per queue fetch, 0 or 1 = one     */
    uint32_t grid_density;      /* direct kernel: occlusion-grid cells per primitive, 1..64 (0 = 2) */
} rt_launch_cfg;
#define RT_KERNEL_DIRECT 0
#define RT_KERNEL_DIFFUSE 1
#define RT_KERNEL_GATHER 2
int rt_set_launch_cfg(rt_ctx *ctx, int kernel, const rt_launch_cfg *cfg);
int rt_get_launch_cfg(rt_ctx *ctx, int kernel, rt_launch_cfg *cfg);
typedef struct {
    rt_launch_cfg cfg[3];   /* chosen shape per kernel class (also installed on the context)      */
    double ms[3];           /* sample time with the chosen shape                                   */
    double default_ms[3];   /* sample time with the default shape (4 waves, 4 blocks/CU, chunk 1) */
    uint32_t candidates;    /* shapes timed                                                        */
    uint32_t sample_prims;  /* primitives in the timing sample                                     */
} rt_tune_result;
/* Times candidate shapes for the direct, diffuse and gather kernels on a sample of `scene` with the
 * ray counts, first source direction (an oblique default when it has none or it is zero) and band
 * count of `req`, for batches of `nsuns` suns. */
int rt_autotune(rt_ctx *ctx, const rt_scene *scene, const rt_band_request *req, uint32_t nsuns, rt_tune_result *out);

/* ---------------------------------------------------------------------------------------------
 * Timing of the last solve's kernels (HIP events on the context stream), milliseconds.  After a
 * castRays or camera launch, direct_ms / direct_rays hold that kernel's time and rays.
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    double direct_ms;   /* direct trace kernels            */
    double diffuse_ms;  /* diffuse/scatter trace kernels   */
    double gather_ms;   /* gather kernels                  */
    double other_ms;    /* occlusion-grid builds           */
    uint64_t direct_rays;
    uint64_t diffuse_rays;
    uint32_t direct_launches;
    uint32_t diffuse_launches;
    uint32_t direct_grid;   /* 1: the direct pass tested occlusion through per-direction grids, 0: BVH */
    uint32_t grid_entries;  /* list entries per direction per traced primitive (x100) of those grids */
} rt_timing;
int rt_enable_timing(rt_ctx *ctx, int on);
/* Intersection predicates of every later trace on the context (radiation passes, cameras, castRays, debug traces).
 * RT_PREDICATES_REFERENCE (default): the reference's FP32 tests, operation for operation
 * (R/shaders/common/intersections.glsl:26-112; castRays: CD/src/CollisionDetection_RayTracing.cpp:1793-1884),
 * the parity mode.  RT_PREDICATES_WATERTIGHT: the watertight ray/triangle test of Woop, Benthin and Wald (2013),
 * patches as two triangles, closed edges, so no ray passes between primitives that share an edge's vertices
 * (the reference's open patch interval lets a ray through a shared patch edge); the distance window is the
 * reference's.  Hit decisions differ only for rays that graze an edge. */
#define RT_PREDICATES_REFERENCE 0
#define RT_PREDICATES_WATERTIGHT 1
int rt_set_predicates(rt_ctx *ctx, int mode);
/* Occlusion test of collimated direct rays: 1 (default) = per-direction occlusion grids when the scene is
 * not periodic and they fit, 0 = always traverse the BVH.  Results are identical either way; grids cost
 * memory (about 100 B per primitive per direction) and are faster. */
int rt_set_direct_grids(rt_ctx *ctx, int on);
/* Traversal work counters (zeroed by rt_set_counting): out[12] = direct (BVH node fetches, primitive
 * tests, rays), diffuse/scatter (same three), castRays (same three), camera (same three).  They
 * define the counted algorithmic bytes per ray used for the roofline. */
int rt_set_counting(rt_ctx *ctx, int on);
int rt_get_counters(rt_ctx *ctx, uint64_t *out);
int rt_get_timing(rt_ctx *ctx, rt_timing *out);

/* Debug/parity access: one diffuse pass's hit words [P][2][n*n] (word = prim | bottom<<31,
 * 0xFFFFFFFF miss, 0xFFFFFFFE face not launched) and one direct pass's strength sums [P][S][2]. */
int rt_debug_diffuse_hits(rt_ctx *ctx, const rt_scene *scene, uint32_t n, uint32_t seed, uint32_t *out);
int rt_debug_direct_sums(rt_ctx *ctx, const rt_scene *scene, uint32_t d, uint32_t seed, uint32_t nsrc,
                         const float *src_dir, float *out);
/* The same for sources of any type (src_pos as rt_band_request.source_dir; types/widths/rotations as
 * rt_band_request.source_type/_width/_rotation). */
int rt_debug_direct_sums_src(rt_ctx *ctx, const rt_scene *scene, uint32_t d, uint32_t seed, uint32_t nsrc,
                             const float *src_pos, const uint32_t *types, const float *widths,
                             const float *rotations, float *out);
/* Generated diffuse rays of one pass, [P][2][n*n][6] = origin.xyz, direction.xyz (no tracing). */
int rt_debug_diffuse_rays(rt_ctx *ctx, const rt_scene *scene, uint32_t n, uint32_t seed, float *out);
/* Closest hit of arbitrary rays [n][6] with the radiation predicates, skipping skip[r]; brute != 0
 * tests every primitive instead of traversing the BVH.  hit = primitive index or 0xFFFFFFFF. */
int rt_debug_trace(rt_ctx *ctx, const rt_scene *scene, uint32_t n, const float *rays, const uint32_t *skip, float tmin,
                   int brute, uint32_t *hit, float *t);

#ifdef __cplusplus
}
#endif
#endif
