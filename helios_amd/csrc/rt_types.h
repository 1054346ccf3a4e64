// rt_types.h -- device-resident scene and job descriptors shared by the kernels and the host runtime.
#pragma once

#include <stdint.h>

#include "rt_device.h"

// Traversal stack: one LDS column per lane.  The LBVH build rejects trees deeper than this.
#ifndef RT_STACK
#define RT_STACK 64
#endif

#define RT_LEAF_BIT 0x80000000u
// traversal flags of a launch: honour texture masks, watertight intersection predicates
#define RT_FL_MASK 1
#define RT_FL_WT 2
// host-side launch flag (not a kernel template argument): every job of a direct launch has an occlusion grid
#define RT_FL_GRIDS 4
#define RT_NOT_LAUNCHED 0xFFFFFFFEu
#define RT_HIT_BOTTOM 0x80000000u

// Periodic boundary face in predicate form (R/shaders/common/periodic_boundary.glsl:28-59):
// anchor v0, edge a = v1-v0, edge b = v3-v0, plane normal n = normalize(cross(v1-v0, v2-v0)).
struct BBoxFace {
    float v0[3], a[3], b[3], n[3], aa, bb;
};

// Everything a kernel needs to trace one scene.  Lives in device memory; kernels read it through a
// pointer held in the job descriptor.
struct DevScene {
    const float4 *rec;     // [P][4] primitive records in leaf (Morton) order
    const float4 *node;    // [P-1][4] BVH2 nodes: both child boxes + child links
    const float4 *node4;   // [nnode4][8] BVH4 nodes (same tree collapsed 4-wide), breadth-first
    const float4 *nrm;     // [P] world normal .xyz, .w = bits(twosided flag)
    const float4 *frame;   // [P] (cos elev, sin elev, cos az, sin az) of the normal
    const float *T;        // [P][16] row-major transforms
    const uint32_t *type;  // [P] 0 patch, 1 triangle, 3 tile, other = not traced
    const uint32_t *uuid;  // [P]
    const uint32_t *order; // [P] traced primitives in leaf (Morton) order, then the others
    uint32_t P;            // primitives (original order)
    uint32_t nleaf;        // traced primitives (leaves)
    uint32_t depth;        // max leaf depth
    uint32_t nnode4;       // BVH4 nodes
    uint32_t depth4;       // BVH4 levels
    uint32_t nbbox;        // periodic faces (0, 2 or 4)
    int px, py;
    float dom[4];          // xmin, xmax, ymin, ymax of the periodic faces
    BBoxFace bbox[4];
    float smax;            // largest |coordinate| of any vertex (scale of FP32 rounding in the scene)
    const uint32_t *mdesc; // [P] mask descriptor index (float4 units from rec) per primitive, RT_MISS = none;
                           // null when no primitive has a texture mask
    const float4 *qnode4;  // [nnode4][4] the BVH4 nodes with 8-bit child boxes (64 B per node), null = none
};

// Occlusion grid of one collimated-source direction: every traced primitive's padded projection onto the
// plane perpendicular to the direction, binned into a uniform grid (CSR: start[cell] .. start[cell + 1]
// index ents[]).  A shadow ray can only hit primitives binned in the cell of its origin's projection.
// A list entry is (leaf slot | RT_GRID_REC3 when the test needs the record's fourth float4 -- patches and
// masked primitives --, bits of the depth key): one 8-byte load per candidate.  0xFFFFFFFF = no primitive.
#define RT_GRID_REC3 0x40000000u
#define RT_GRID_SLOT 0x3FFFFFFFu
struct SunGrid {
    float u[3], v[3];      // orthonormal basis of the projection plane
    float w[3];            // unit direction towards the source (depth axis of the keys)
    float g0u, g0v;        // grid origin
    float inv_cs;          // 1 / cell size
    uint32_t nx, ny;
    uint32_t sorted;       // 1: each cell's list is in descending key order (a scan may stop at the first
                           // key below the ray origin's depth)
    uint32_t sbase;        // start[] values count from sbase (modulo 2^32): cell c lists
                           // ents[start[c] - sbase, start[c + 1] - sbase)
    const uint32_t *start; // [nx*ny + 1]
    const uint2 *ents;     // this direction's list: (slot | flag, key); the key is the primitive's largest
                           // depth along w (max over its vertices of v.w), padded
};

// one direct trace job: one scene, one collimated source
struct DirectJob {
    const DevScene *sc;
    float sdir[3];         // collimated: raw direction towards the source (normalised in-kernel);
                           // other types: source position
    uint32_t d;            // rays per side (d*d rays per primitive)
    uint32_t seed;
    float *out;            // [P][2] strength sums lit-from-top / lit-from-bottom
    uint32_t item_begin;   // first work item of this job
    uint32_t P;
    const uint32_t *order; // item -> primitive map (null: item k is primitive k)
    const SunGrid *grid;   // occlusion grid of this direction (null: traverse the BVH)
    uint32_t stype;        // source type: 0 collimated, 1 sphere, 2 sun sphere, 3 rectangle, 4 disk
    uint32_t skip_draws;   // RNG draws of the request's earlier non-collimated sources (2 each)
    float width[2];        // sphere diameter | rectangle width, height | disk radius
    float rot[9];          // rotation matrix (rows) of a rectangle or disk source
};

// One camera launch (R/shaders/camera_raygen.comp, pixel_label_raygen.comp): camera pose and lens, the
// tile of pixels, and what a hit or a miss reads.  rot = (sin, cos) of theta = elevation - pi/2 and of
// phi = pi/2 - azimuth, evaluated in double on the host and rounded.
struct CameraLaunch {
    const DevScene *sc;
    float pos[3];
    float st, ct, sp, cp;
    float focal, lens, aspect, viewplane;
    uint32_t res_x, res_y, full_x, full_y, off_x, off_y;
    uint32_t spp, seed;
    uint32_t B, S;
    const float *out_top, *out_bot;    // [P][B] radiation_out of the last passes
    const float *flux;                 // [S][B] source fluxes
    const float *src_pos;              // [S][3]
    const uint32_t *src_type;          // [S]
    const float *src_width;            // [S][2]
    const float *src_rot;              // [S][9] rotation matrix rows
    const float *sky_params;           // [B][4] Prague (circ strength, circ width, horizon, norm)
    const float *sky_rad;              // [B] camera sky radiance
    const float *solar_rad;            // [B] solar disk radiance
    const float *cam_dflux;            // [B] sky flux seen on a miss when the band emits
    const uint32_t *emis;              // [B] band emission flags
    float sun[3];
    float solar_cos;
    float *pixels;                     // [B][full_x * full_y]
    uint32_t *labels;                  // [full_x * full_y]
    float *depths;                     // [full_x * full_y]
    // specular reflection (spec NULL: off)
    const float *spec;                 // [S][spec_P][B] radiation_specular of the direct launch
    const float *fcam;                 // [S][B] camera weights of the sources
    const float *spec_exp;             // [P] specular exponents
    const float *spec_scale;           // [P] specular scales (mode 2)
    uint32_t spec_P, spec_mode;        // primitives; 1 = scale 1, 2 = spec_scale
};

// Translucent covers of one solve (device copy): primitives glass in every band pass rays through, and the
// kernels push their absorbed / reflected shares into `push` (FP64 atomics, added to the FP32 buffers after
// the pass, so the sums do not depend on the atomics' order beyond FP64 rounding).
#define RT_GLASS_MAX_CH 8
struct GlassCtx {
    const uint8_t *cover;  // [P] 1 = cover for this solve's bands
    const float *gn;       // [P][B] refractive index
    const float *gkl;      // [P][B] absorption product K*L
    double *push;          // [P][C][3] radiation_in, scatter_top, scatter_bottom
    uint32_t B, C;
};

// one direct job of a solve with covers: DirectJob's source and rays, plus its channels' source fluxes and
// the per-band strength sums it writes
struct GlassDirectJob {
    DirectJob dj;
    const float *flux;     // [B] source flux of this (sun, source) per band
    float *gsum;           // [P][B][2] strength x cover transmittance, lit from top / bottom
    uint32_t chan0;        // first channel of this job's sun (sun * B)
};

// one diffuse trace pass: writes hit words [P][2][n*n]
struct DiffuseJob {
    const DevScene *sc;
    uint32_t n;
    uint32_t seed;
    uint32_t *hits;
    uint32_t item_begin;
    uint32_t P;
    uint32_t faces;        // bit 0: top launch (face 1), bit 1: bottom launch (face 0); 0 means both
    uint32_t pad;
    const uint32_t *order; // item -> primitive map (null: item k is primitive k)
};

// one gather over a hit table: C channels, band of channel c is c % B
struct GatherJob {
    const DevScene *sc;
    const uint32_t *hits;  // [P][2][n*n]
    uint32_t n;
    uint32_t C, B;
    const float *out_top;  // [P][C]
    const float *out_bot;  // [P][C]
    const float *dflux;    // [C] sky flux (0 in scattering passes)
    const uint8_t *active; // [C]
    const float *rho;      // [P][B]
    const float *tau;      // [P][B]
    float *in;             // [P][C]
    float *st;             // [P][C]
    float *sb;             // [P][C]
    uint32_t item_begin;
    uint32_t P;
    const uint32_t *order; // item -> primitive map (null: item k is primitive k)
    const float *sky;      // [C][9] anisotropic sky of each channel (sky_weight parameters), null = isotropic
    uint32_t seed;         // seed of the diffuse pass whose hit words these are (miss directions regenerated)
    // energy bookkeeping (null = off): per primitive and channel the incident flux density of this pass and
    // its sky part ([P][C][2]; 0 for a channel the pass does not deposit), and per primitive the rays of each
    // face that reached the sky ([P][2]: top launch, bottom launch)
    float *acct;
    uint32_t *miss;
};
