/*
 * rt_oracle.h -- CPU restatement of the Helios radiation-transfer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (helios_amd/, include/) links, loads or calls
 * this library; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
 * checker.  It restates, in plain C, the per-ray FP32 arithmetic of the reference shaders and the
 * pass schedule of RadiationModel::runBand, with FP64 per-primitive accumulation.
 *
 * Every function cites the reference file:line it follows (paths relative to the Helios tree;
 * R = plugins/radiation, CD = plugins/collisiondetection).
 *
 * Pinning (how each part is checked against something other than itself; see tests/):
 *   orc_tea            published TEA test vectors (32 rounds, key 0) -- tests/test_oracle_rng.py
 *   orc_lcg_next       Numerical Recipes ranqd1 published sequence from seed 0
 *   orc_sincos_2pi     libm sin/cos in double, max error <= 2 ulp over a dense sweep
 *   predicates         CD selfTest golden cases (CD/tests/selfTest.cpp:3253-3297, 4522-4630,
 *                      5111-5237), closed-form hit distances
 *   orc_trace / castRays  SAH-BVH traversal == brute force on small random scenes
 *   orc_run_band       Simple Direct (R/tests/selfTest.cpp:518-545), black parallel rectangles
 *                      view factor (:670-729), Beer's-law canopy (:1372-1468), furnace-style
 *                      second-law enclosure (:1061-1132), closed-form pass-schedule cases
 */
#ifndef HELIOS_RT_ORACLE_H
#define HELIOS_RT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG (R/shaders/common/random.glsl:10-45) ---- */
uint32_t orc_tea(uint32_t v0, uint32_t v1, uint32_t rounds, const uint32_t key[4], uint32_t *v1_out);
uint32_t orc_tea16(uint32_t val0, uint32_t val1);
uint32_t orc_lcg_next(uint32_t state);
float orc_rnd(uint32_t *state);
void orc_sincos_2pi(float r, float *s, float *c);

/* ---- scene ---- */
typedef struct orc_scene orc_scene;

/* transforms: [P][16] row-major; types: 0 patch, 1 triangle, 3 tile (treated as patch);
 * twosided: raw Helios flag 0..3; uuids: [P] (returned by castRays). */
orc_scene *orc_scene_create(uint32_t P, const float *transforms, const uint32_t *types,
                            const uint32_t *twosided, const uint32_t *uuids, int periodic_x, int periodic_y);
void orc_scene_destroy(orc_scene *s);

/* world vertices [P][4][3] (triangles use 3), normals [P][3], diffuse frame [P][4] (ce,se,ca,sa),
 * domain bounds [4] (xmin,xmax,ymin,ymax of the periodic faces), bbox count */
void orc_scene_geometry(const orc_scene *s, float *verts, float *normals, float *frame, float *domain, uint32_t *nbbox);

/* Closest hit with the radiation predicates (intersections.glsl).  Returns prim index or 0xFFFFFFFF.
 * brute != 0 tests every primitive (pin for the BVH path). */
uint32_t orc_trace(const orc_scene *s, const float o[3], const float d[3], float tmin, uint32_t skip,
                   int brute, float *t_out);

/* One full radiation trace step with the periodic wrap loop (direct_raygen.comp:603-668 without glass);
 * returns final hit prim or 0xFFFFFFFF. */
uint32_t orc_trace_periodic(const orc_scene *s, const float o[3], const float d[3], float tmin, uint32_t skip);

/* Ray generation (bit-exact FP32 restatement). */
void orc_direct_ray(const orc_scene *s, uint32_t prim, uint32_t lx, uint32_t ly, uint32_t d, uint32_t seed,
                    float origin[3]);
void orc_diffuse_ray(const orc_scene *s, uint32_t prim, uint32_t lx, uint32_t ly, uint32_t n, uint32_t face,
                     uint32_t seed, float origin[3], float dir[3]);

/* Diffuse hit words for one pass: out[P][2][n*n], word = prim | (bottom_face_hit << 31), miss 0xFFFFFFFF,
 * face slot not launched = 0xFFFFFFFE. */
void orc_diffuse_hits(const orc_scene *s, uint32_t n, uint32_t seed, uint32_t *out);

/* Direct visibility sums: out[P][S][2] = (sum strength lit-from-top, sum strength lit-from-bottom) */
void orc_direct_sums(const orc_scene *s, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_dir, double *out);

/* Direct sums for sources of any type (direct_raygen.comp:482-628): src_pos [S][3] (collimated: direction),
 * types [S] (NULL = all collimated; 1 sphere, 2 sun sphere, 3 rectangle, 4 disk), widths [S][2],
 * rotations [S][3] (ZYX Euler angles of rectangles and disks). */
void orc_direct_sums_src(const orc_scene *s, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_pos,
                         const uint32_t *types, const float *widths, const float *rotations, double *out);
void orc_direct_sums_src_range(const orc_scene *s, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_pos,
                               const uint32_t *types, const float *widths, const float *rotations, uint32_t p0,
                               uint32_t p1, double *out);
void orc_rotation_matrix(const float rot[3], float m[9]);

/* Range versions (primitives [p0, p1) only; output indexed from p0) -- used to time bounded samples. */
void orc_diffuse_hits_range(const orc_scene *s, uint32_t n, uint32_t seed, uint32_t p0, uint32_t p1, uint32_t *out);
void orc_direct_sums_range(const orc_scene *s, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_dir,
                           uint32_t p0, uint32_t p1, double *out);

/* Texture masks (one byte per texel, 0 = transparent; mask_ids [P], -1 = none; uv [P][8] or NULL;
 * uv_ids [P] or NULL).  Returns -1 on a mask id out of range. */
int orc_scene_set_masks(orc_scene *s, uint32_t nmasks, const uint8_t *texels, const uint32_t *sizes,
                        const int32_t *mask_ids, const float *uv, const int32_t *uv_ids);

/* CollisionDetection::castRaysSoA semantics. */
void orc_cast_rays(const orc_scene *s, uint64_t n, const float *origins, const float *dirs, float max_dist,
                   float *out_t, float *out_normal, uint32_t *out_uuid, int brute);

/* ---- cameras (R/shaders/camera_raygen.comp, pixel_label_raygen.comp) ---- */
typedef struct {
    float position[3];
    float direction[2];        /* elevation, azimuth (radians)                          */
    float focal_length, lens_diameter, fov_aspect, viewplane_length;
    uint32_t resolution[2];    /* tile                                                  */
    uint32_t resolution_full[2];
    uint32_t pixel_offset[2];
    uint32_t antialiasing_samples;
    uint32_t random_seed;
} orc_camera;
typedef struct {
    uint32_t nbands, nsources;
    const float *out_top, *out_bottom;  /* [P][B] radiation_out of the last passes         */
    const float *source_flux;           /* [S][B]                                          */
    const float *source_pos;            /* [S][3]                                          */
    const uint32_t *source_type;        /* [S]                                             */
    const float *source_width;          /* [S][2]                                          */
    const float *source_rotation;       /* [S][3] Euler angles                             */
    const float *sky_params;            /* [B][4]                                          */
    const float *sky_radiance;          /* [B]                                             */
    const float *solar_radiance;        /* [B]                                             */
    const float *camera_diffuse_flux;   /* [B]                                             */
    const uint32_t *emission;           /* [B]                                             */
    float sun[3];
    float solar_cos;
    /* specular reflection (spec NULL: off) */
    const float *spec;                  /* [S][P][B] radiation of each source delivered to each primitive */
    const float *spec_weight;           /* [S][B] camera weights of the sources             */
    const float *spec_exp;              /* [P] specular exponents                           */
    const float *spec_scale;            /* [P] specular scales (mode 2)                     */
    uint32_t spec_mode;                 /* 1 = scale 1, 2 = spec_scale                      */
} orc_camera_light;
/* Adds the tile's samples to pixels [B][full_x*full_y] (FP64 accumulation of FP32 samples). */
void orc_camera_pixels(const orc_scene *s, const orc_camera *c, const orc_camera_light *L, double *pixels);
/* Pixel labels (UUID + 1, 0 = sky) and depths (-1 = sky) of the tile, [full_x*full_y] each. */
void orc_pixel_labels(const orc_scene *s, const orc_camera *c, uint32_t *labels, float *depths);

/* ---- runBand ---- */
typedef struct {
    uint32_t nbands;           /* B (launched bands, alphabetical order)               */
    uint32_t nsources;         /* S (collimated)                                        */
    uint32_t direct_rays;      /* max directRayCount over bands                         */
    uint32_t diffuse_rays;     /* max diffuseRayCount over bands                        */
    const uint32_t *band_depth;/* [B] scatteringDepth                                   */
    const uint32_t *emission;  /* [B] emission flag                                     */
    const float *source_dir;   /* [S][3] direction towards each collimated source       */
    const float *source_flux;  /* [S][B]                                                */
    const float *diffuse_flux; /* [B]                                                   */
    const float *rho;          /* [P][B]                                                */
    const float *tau;          /* [P][B]                                                */
    const float *eps;          /* [P][B] (used for emission bands; NULL = 1)            */
    const float *temperature;  /* [P] (NULL = 300 K)                                    */
    uint32_t direct_seed;
    const uint32_t *diffuse_seed; /* [1 + max depth]: primary diffuse pass, then each scattering pass */
    uint32_t schedule;         /* 0 = reference pass schedule, 1 = energy-conserving    */
    const uint32_t *source_type;   /* [S] or NULL (all collimated)                       */
    const float *source_width;     /* [S][2]                                             */
    const float *source_rotation;  /* [S][3]                                             */
    const float *diffuse_extinction; /* [B] power-law K, NULL = isotropic                   */
    const float *diffuse_peak_dir;   /* [B][3]                                             */
    const float *diffuse_dist_norm;  /* [B]                                                */
    const float *sky_params;         /* [B][4] Prague                                      */
    /* translucent covers (RayTracingMaterial is_glass / glass_n / glass_KL, RayTracingTypes.h:173-178),
     * [P][B] each, is_glass NULL = none: a primitive that is glass in every band is a cover */
    const uint8_t *is_glass;
    const float *glass_n;
    const float *glass_KL;
} orc_band_params;

/* Angular transmittance, reflectance and absorptance of a glass sheet (Fresnel, both polarisations,
 * internal reflections in closed form, Bouguer absorption; R/shaders/common/glass_cover.glsl:15-53):
 * out[3] = tau, rho, alpha. */
void orc_glass_tra(float cos_theta, float n, float KL, float out[3]);

/* R[P][B] = radiation_flux_<band>; diag (may be NULL) receives [B] FP64 energy totals:
 * diag[0*B+b] absorbed sum(A*R), diag[1*B+b] primary diffuse incident, diag[2*B+b] direct incident */
int orc_run_band(const orc_scene *s, const orc_band_params *p, double *R, double *diag);

/* Intersection predicates of every later call: 0 = the reference's (default), 1 = watertight (Woop, Benthin
 * and Wald 2013, as rt_set_predicates).  Process-wide. */
void orc_set_predicates(int mode);

/* primitive areas [P] (patch |a x b|, triangle |e1 x e2|/2) in FP64 */
void orc_areas(const orc_scene *s, double *out);

#ifdef __cplusplus
}
#endif
#endif
