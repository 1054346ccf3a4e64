// rt_solve.hip -- device-resident runBand pass schedule (piece d) and its elementwise kernels.
//
// The schedule follows R/src/RadiationModel.cpp:3470-4672 step by step, but every buffer stays in
// HBM: the scatter -> radiation_out hand-offs that the reference does on the host between passes
// (VulkanComputeBackend.cpp:1765-1877, 2055-2160) are elementwise kernels here, and a whole solve is
// one stream of launches with a single synchronisation at the end.
//
// Channels: a solve works on C = K*B channels (K suns sharing scene, bands, materials and diffuse
// seeds; B bands).  Buffers are [P][C]; channel c belongs to sun c / B and band c % B.  Every
// per-sun decision of the reference schedule (rundirect, diffuse enabled, copy at s = 0) is a
// per-channel mask, so each sun's channels evolve exactly as in a solve of that sun alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rt_solve.h"

namespace rt {

// radiation_in / scatter_top / scatter_bottom += source flux x strength sums x optical properties; dsum
// [K][S][P][2] (strength per lit face), or with covers gsum [K][S][P][B][2] (strength x the covers'
// transmittance per band) when gsum is given
__global__ void k_deposit_direct(uint32_t P, uint32_t C, uint32_t B, uint32_t S, const float *__restrict__ dsum,
                                 const float *__restrict__ flux, const float *__restrict__ rho,
                                 const float *__restrict__ tau, const uint8_t *__restrict__ mask, float *in, float *st,
                                 float *sb, const float *__restrict__ gsum) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)P * C) return;
    uint32_t i = (uint32_t)(q / C), c = (uint32_t)(q % C);
    if (!mask[c]) return;
    uint32_t k = c / B, b = c % B;
    float r = rho[(size_t)i * B + b], t = tau[(size_t)i * B + b];
    float a_in = 0.f, a_st = 0.f, a_sb = 0.f;
    for (uint32_t s = 0; s < S; s++) {
        const float *ds = gsum ? gsum + ((((size_t)k * S + s) * P + i) * B + b) * 2
                               : dsum + (((size_t)k * S + s) * P + i) * 2;
        float stp = ds[0], sbt = ds[1];
        float F = flux[((size_t)k * S + s) * B + b];
        a_in += F * (stp + sbt) * (1.0f - r - t);
        a_st += F * (stp * r + sbt * t);
        a_sb += F * (stp * t + sbt * r);
    }
    in[q] += a_in;
    st[q] += a_st;
    sb[q] += a_sb;
}

// scattered direct energy -> radiation_out, one-sided primitives combine both faces
// (RadiationModel.cpp:3999-4036)
__global__ void k_after_direct(uint32_t P, uint32_t C, const float4 *__restrict__ nrm, const uint8_t *__restrict__ mask,
                               float *st, float *sb, float *ot, float *ob) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)P * C) return;
    uint32_t i = (uint32_t)(q / C), c = (uint32_t)(q % C);
    if (!mask[c]) return;
    float t = st[q], b = sb[q];
    if (__float_as_uint(nrm[i].w) == 0u) { float tot = t + b; t = tot; b = tot; }
    ot[q] = t;
    ob[q] = b;
    st[q] = 0.f;
    sb[q] = 0.f;
}

// emitted flux sigma*eps*T^4 added to radiation_out (RadiationModel.cpp:4041-4137)
__global__ void k_emission(uint32_t P, uint32_t C, uint32_t B, const float4 *__restrict__ nrm,
                           const float *__restrict__ eps, const float *__restrict__ temp,
                           const uint32_t *__restrict__ depth, const uint8_t *__restrict__ mask, float *ot, float *ob) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)P * C) return;
    uint32_t i = (uint32_t)(q / C), c = (uint32_t)(q % C);
    if (!mask[c]) return;
    uint32_t b = c % B;
    float e = eps ? eps[(size_t)i * B + b] : 1.0f;
    if (depth[b] == 0 && e != 1.f) e = 1.f;
    float T = temp ? temp[i] : 300.f;
    if (T < 0.f) T = 300.f;
    float out = (float)(5.6703744E-8 * (double)e * ((double)T * (double)T * (double)T * (double)T));
    ot[q] += out;
    if (__float_as_uint(nrm[i].w) != 0u) ob[q] += out;
}

// per-iteration scatter hand-off: optional deposit of a retiring band's leftover, copy
// scatter -> radiation_out, zero scatter (RadiationModel.cpp:4255-4289)
__global__ void k_copy_zero(uint32_t P, uint32_t C, const uint8_t *__restrict__ retire, const uint8_t *__restrict__ copy,
                            const uint8_t *__restrict__ zero, float *in, float *st, float *sb, float *ot, float *ob) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)P * C) return;
    uint32_t c = (uint32_t)(q % C);
    float t = st[q], b = sb[q];
    if (retire[c]) { in[q] += t + b; t = 0.f; b = 0.f; st[q] = 0.f; sb[q] = 0.f; }
    if (copy[c]) { ot[q] = t; ob[q] = b; }
    if (zero[c]) { st[q] = 0.f; sb[q] = 0.f; }
}

// R = radiation_in + scatter_top + scatter_bottom (RadiationModel.cpp:4631-4654), written [K][P][B]
__global__ void k_final(uint32_t P, uint32_t C, uint32_t B, const float *__restrict__ in, const float *__restrict__ st,
                        const float *__restrict__ sb, float *__restrict__ R) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)P * C) return;
    uint32_t i = (uint32_t)(q / C), c = (uint32_t)(q % C);
    uint32_t k = c / B, b = c % B;
    R[((size_t)k * P + i) * B + b] = in[q] + st[q] + sb[q];
}

// Emitted flux density of primitive i in band b (RadiationModel.cpp:4041-4137): sigma eps T^4, FP64.
__device__ __forceinline__ double emitted(uint32_t i, uint32_t B, uint32_t b, const float *eps, const float *temp,
                                          const uint32_t *depth) {
    float e = eps ? eps[(size_t)i * B + b] : 1.0f;
    if (depth[b] == 0 && e != 1.f) e = 1.f;
    float T = temp ? temp[i] : 300.f;
    if (T < 0.f) T = 300.f;
    return 5.6703744E-8 * (double)e * ((double)T * (double)T * (double)T * (double)T);
}

// FP64 per-channel totals weighted by primitive area, one block per channel (rows of out, [row][C]):
// 0 source (direct incident), 3 absorbed (sum of radiation_flux), 5 leftover scatter (deposited as absorbed by
// the final combine, part of row 3), 2 emitted (sigma eps T^4 from each radiating face: the top, and the bottom
// of a two-sided primitive) when emit_mask is given.  Rows 1 and 4 (sky, escaped) come from the passes.
__global__ void k_energy(uint32_t P, uint32_t C, uint32_t B, uint32_t S, const double *__restrict__ area,
                         const float *__restrict__ R, const float *__restrict__ dsum, const float *__restrict__ flux,
                         const float *__restrict__ st, const float *__restrict__ sb, const float4 *__restrict__ nrm,
                         const uint8_t *__restrict__ emit_mask, const float *__restrict__ eps,
                         const float *__restrict__ temp, const uint32_t *__restrict__ depth, double *out) {
    uint32_t c = blockIdx.x;
    uint32_t k = c / B, b = c % B;
    const bool em = emit_mask && emit_mask[c];
    double a_abs = 0.0, a_src = 0.0, a_left = 0.0, a_emit = 0.0;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
        double A = area[i];
        a_abs += A * (double)R[((size_t)k * P + i) * B + b];
        a_left += A * ((double)st[(size_t)i * C + c] + (double)sb[(size_t)i * C + c]);
        if (dsum) {
            for (uint32_t s = 0; s < S; s++) {
                const float *ds = dsum + (((size_t)k * S + s) * P + i) * 2;
                a_src += A * (double)flux[((size_t)k * S + s) * B + b] * ((double)ds[0] + (double)ds[1]);
            }
        }
        if (em && A > 0.0) {
            const double faces = __float_as_uint(nrm[i].w) != 0u ? 2.0 : 1.0;
            a_emit += A * faces * emitted(i, B, b, eps, temp, depth);
        }
    }
    __shared__ double r[4][256];
    r[0][threadIdx.x] = a_src;
    r[1][threadIdx.x] = a_emit;
    r[2][threadIdx.x] = a_abs;
    r[3][threadIdx.x] = a_left;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 4; q++) r[q][threadIdx.x] += r[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0 * C + c] = r[0][0];
        out[2 * C + c] = r[1][0];
        out[3 * C + c] = r[2][0];
        out[5 * C + c] = r[3][0];
    }
}

// Energy bookkeeping of one diffuse/scattering pass, FP64 partial sums of one block over its range of
// primitives, thread per channel (partial[block][row][C]):
//   0 radiated: A (out_top + out_bottom) of the faces that launch rays (the top of every traced primitive,
//     the bottom of a two-sided one) -- what the gather hands on from each primitive;
//   1 arrived:  A x (incident - sky) -- flux gathered from surfaces;
//   2 sky:      A x sky incident;
//   3 escaped:  A (out_top m_top + out_bottom m_bottom) / n^2, with m the face's rays that reached the sky:
//     the radiated flux times the primitive's own Monte Carlo estimate of its view factor to the sky.
// A channel the pass does not deposit (a retired band) moves no energy and adds nothing.  Radiated =
// arrived + escaped up to Monte Carlo error when the pass hands on exactly radiation_out.
__global__ void k_energy_pass(uint32_t P, uint32_t C, uint32_t nn, const double *__restrict__ area,
                              const float4 *__restrict__ nrm, const uint32_t *__restrict__ type,
                              const uint8_t *__restrict__ active, const float *__restrict__ acct,
                              const uint32_t *__restrict__ miss, const float *__restrict__ ot,
                              const float *__restrict__ ob, double *partial) {
    const uint32_t per = (P + gridDim.x - 1) / gridDim.x;
    const uint32_t i0 = min(P, blockIdx.x * per), i1 = min(P, i0 + per);
    const double inv = 1.0 / (double)nn;
    for (uint32_t c = threadIdx.x; c < C; c += blockDim.x) {
        double rad = 0.0, arr = 0.0, sky = 0.0, esc = 0.0;
        if (active[c]) {
            for (uint32_t i = i0; i < i1; i++) {
                const uint32_t ty = type[i];
                if (!(ty == 0u || ty == 1u || ty == 3u)) continue;
                const double A = area[i];
                const size_t q = (size_t)i * C + c;
                const double t = ot[q];
                const double bt = __float_as_uint(nrm[i].w) != 0u ? (double)ob[q] : 0.0;
                const double in_tot = acct[2 * q], in_sky = acct[2 * q + 1];
                rad += A * (t + bt);
                arr += A * (in_tot - in_sky);
                sky += A * in_sky;
                esc += A * (t * (double)miss[2 * (size_t)i] + bt * (double)miss[2 * (size_t)i + 1]) * inv;
            }
        }
        double *pp = partial + (size_t)blockIdx.x * 4 * C;
        pp[0 * C + c] = rad;
        pp[1 * C + c] = arr;
        pp[2 * C + c] = sky;
        pp[3 * C + c] = esc;
    }
}

// Sums the partial rows of k_energy_pass over its blocks in block order (deterministic): out[row][C].
__global__ void k_energy_sum(uint32_t nblocks, uint32_t C, const double *__restrict__ partial, double *out) {
    for (uint32_t q = threadIdx.x; q < 4 * C; q += blockDim.x) {
        double a = 0.0;
        for (uint32_t g = 0; g < nblocks; g++) a += partial[(size_t)g * 4 * C + q];
        out[q] = a;
    }
}

// primitive areas in FP64 from the transforms (|T e_x x T e_y| for patches, half that for triangles)
__global__ void k_area(uint32_t P, const float *__restrict__ T, const uint32_t *__restrict__ type, double *area) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float *t = T + (size_t)i * 16;
    uint32_t ty = type[i];
    double ax, ay, az, bx, by, bz;
    if (ty == 0 || ty == 3) { ax = t[0]; ay = t[4]; az = t[8]; bx = t[1]; by = t[5]; bz = t[9]; }
    else { ax = t[1]; ay = t[5]; az = t[9]; bx = (double)t[0] + t[1]; by = (double)t[4] + t[5]; bz = (double)t[8] + t[9]; }
    double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    double a = sqrt(cx * cx + cy * cy + cz * cz);
    area[i] = (ty == 0 || ty == 3) ? a : (ty == 1 ? 0.5 * a : 0.0);
}

#define CK(x)                                                                     \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #x); \
            return -1;                                                            \
        }                                                                         \
    } while (0)
#define LK()                                                                      \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) {                                                   \
            set_error(std::string("kernel launch failed: ") + hipGetErrorString(e_)); \
            return -1;                                                            \
        }                                                                         \
    } while (0)

static inline uint32_t nb(size_t n) { return (uint32_t)((n + 255) / 256); }

int launch_deposit_direct(uint32_t P, uint32_t C, uint32_t B, uint32_t S, const float *dsum, const float *flux,
                          const float *rho, const float *tau, const uint8_t *mask, float *in, float *st, float *sb,
                          hipStream_t s, const float *gsum) {
    size_t PC = (size_t)P * C;
    if (PC == 0) return 0;
    k_deposit_direct<<<nb(PC), 256, 0, s>>>(P, C, B, S, dsum, flux, rho, tau, mask, in, st, sb, gsum);
    LK();
    return 0;
}

int source_shapes(uint32_t S, const uint32_t *types, const float *widths, const float *rotations,
                  std::vector<SourceShape> &out, const char *who) {
    out.assign(S, SourceShape{});
    if (!types) return 0;
    for (uint32_t s = 0; s < S; s++) {
        SourceShape &sh = out[s];
        sh.type = types[s];
        if (sh.type > 4u) {
            set_error(std::string(who) + ": source " + std::to_string(s) + " has unknown type " +
                      std::to_string(sh.type) + " (0 collimated, 1 sphere, 2 sun sphere, 3 rectangle, 4 disk)");
            return -1;
        }
        if (sh.type == 0u) continue;
        if (!widths) {
            set_error(std::string(who) + ": source widths are required for non-collimated sources");
            return -1;
        }
        sh.width[0] = widths[2 * s];
        sh.width[1] = widths[2 * s + 1];
        const bool bad = !(sh.width[0] > 0.f) || !std::isfinite(sh.width[0]) ||
                         (sh.type == 3u && (!(sh.width[1] > 0.f) || !std::isfinite(sh.width[1])));
        if (bad) {
            set_error(std::string(who) + ": source " + std::to_string(s) + " needs a positive finite width");
            return -1;
        }
        if (sh.type == 3u || sh.type == 4u) {
            if (!rotations) {
                set_error(std::string(who) + ": source rotations are required for rectangle and disk sources");
                return -1;
            }
            const double rx = rotations[3 * s], ry = rotations[3 * s + 1], rz = rotations[3 * s + 2];
            const double cx = std::cos(rx), sx = std::sin(rx), cy = std::cos(ry), sy = std::sin(ry);
            const double cz = std::cos(rz), sz = std::sin(rz);
            const double m[9] = {cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx,
                                 sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx,
                                 -sy,     cy * sx,                cy * cx};
            for (int k = 0; k < 9; k++) sh.rot[k] = (float)m[k];
        }
    }
    return 0;
}

void set_job_source(DirectJob &j, const std::vector<SourceShape> &shapes, uint32_t s) {
    const SourceShape &sh = shapes[s];
    j.stype = sh.type;
    j.width[0] = sh.width[0];
    j.width[1] = sh.width[1];
    for (int k = 0; k < 9; k++) j.rot[k] = sh.rot[k];
    uint32_t skip = 0;
    for (uint32_t q = 0; q < s; q++) skip += shapes[q].type != 0u ? 2u : 0u;
    j.skip_draws = skip;
}

int Workspace::reserve(size_t bytes) {
    if (bytes <= cap) return 0;
    if (base) (void)hipFree(base);
    base = nullptr;
    cap = 0;
    size_t want = capacity_for(bytes);
    CK(hipMalloc(&base, want));
    cap = want;
    return 0;
}

void Workspace::release() {
    if (base) (void)hipFree(base);
    base = nullptr;
    cap = 0;
}

namespace {
struct Carve {
    char *p;
    size_t off = 0;
    template <typename T>
    T *take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T *r = reinterpret_cast<T *>(p + off);
        off += n * sizeof(T);
        return r;
    }
};
}  // namespace

static void tick(SolveCtx &sx, int cls, hipEvent_t a, hipEvent_t b) { sx.ev.push_back({cls, a, b}); }

static int ev_pair(SolveCtx &sx, hipEvent_t &a, hipEvent_t &b) {
    if (!sx.timing) { a = b = nullptr; return 0; }
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    return 0;
}

// Per-channel sky_weight parameters ([C][9]: K, norm, peak xyz, Prague 4) of a request whose diffuse sky
// is anisotropic in some band; empty when every band's sky is isotropic.
static int sky_table(const rt_band_request &q, const SunSet &suns, uint32_t K, std::vector<float> &out) {
    const uint32_t B = q.nbands;
    out.clear();
    bool any = false;
    for (uint32_t b = 0; b < B; b++) {
        const bool pl = q.diffuse_extinction && q.diffuse_extinction[b] > 0.f;
        const bool pr = q.sky_params && q.sky_params[4 * b + 3] > 0.f;
        if ((pl || pr) && !q.diffuse_peak_dir && !suns.peak) {
            set_error("run_band: an anisotropic sky (diffuse_extinction or sky_params) needs diffuse_peak_dir");
            return -1;
        }
        if (pl && !q.diffuse_dist_norm) {
            set_error("run_band: a power-law sky (diffuse_extinction > 0) needs diffuse_dist_norm");
            return -1;
        }
        any = any || pl || pr;
    }
    if (!any) return 0;
    out.assign((size_t)K * B * 9, 0.f);
    for (uint32_t k = 0; k < K; k++)
        for (uint32_t b = 0; b < B; b++) {
            float *p = out.data() + ((size_t)k * B + b) * 9;
            p[0] = q.diffuse_extinction ? q.diffuse_extinction[b] : 0.f;
            p[1] = q.diffuse_dist_norm ? q.diffuse_dist_norm[b] : 0.f;
            const float *pk = suns.peak ? suns.peak + ((size_t)k * B + b) * 3 : (q.diffuse_peak_dir ? q.diffuse_peak_dir + 3 * b : nullptr);
            if (pk) { p[2] = pk[0]; p[3] = pk[1]; p[4] = pk[2]; }
            if (q.sky_params)
                for (int a = 0; a < 4; a++) p[5 + a] = q.sky_params[4 * b + a];
        }
    return 0;
}

static constexpr uint32_t ENERGY_BLOCKS = 1024;  // blocks of k_energy_pass

size_t solve_workspace_bytes(uint32_t P, const rt_band_request &q, uint32_t K, bool resident_materials, bool acct) {
    const uint32_t B = q.nbands, Sn = q.nsources, C = K * B;
    uint32_t maxdepth = 0;
    for (uint32_t b = 0; b < B; b++) maxdepth = std::max(maxdepth, q.band_depth[b]);
    const uint32_t n = (uint32_t)std::ceil(std::sqrt((double)q.diffuse_rays));
    const size_t nn = (size_t)n * n, PC = (size_t)P * C;
    // upper bound: the hit table is counted whenever diffuse rays are requested (whether a diffuse
    // pass runs also depends on the per-sun sky fluxes)
    bool need_hits = n > 0;
    const uint32_t nm = 6 + 4 * std::max(maxdepth, 1u);
    size_t bytes = 5 * PC * 4 + (need_hits ? (size_t)P * 2 * nn * 4 : 0) + (size_t)K * Sn * P * 2 * 4 + (size_t)P * 8 +
                   (size_t)K * Sn * B * 4 + (size_t)C * 4 * 2 + (size_t)nm * C + (size_t)B * 4 + (size_t)C * 8 * 3 +
                   (size_t)K * Sn * sizeof(DirectJob) + (sizeof(DiffuseJob) + sizeof(GatherJob)) * (1 + maxdepth) +
                   (size_t)C * 9 * 4 + 64 * 64;
    if (!resident_materials) bytes += (size_t)P * B * 4 * 3 + (size_t)P * 4;
    // translucent covers: flags, glass parameters, pushed shares [P][C][3] FP64, per-band direct sums, jobs
    if (q.is_glass)
        bytes += (size_t)P + (size_t)P * B * 8 + PC * 24 + (size_t)K * Sn * P * B * 8 +
                 (size_t)K * Sn * sizeof(GlassDirectJob) + sizeof(GlassCtx) + 4 * 256;
    // energy bookkeeping: [P][C][2] pass densities, [P][2] sky-bound rays, partials, per-pass rows, totals
    if (acct)
        bytes += PC * 8 + (size_t)P * 8 + (size_t)ENERGY_BLOCKS * 4 * C * 8 + (size_t)(1 + maxdepth) * 4 * C * 8 +
                 (size_t)C * RT_ENERGY_ROWS * 8;
    return bytes + 40 * 256;  // carve alignment
}

size_t Workspace::capacity_for(size_t bytes) { return bytes + bytes / 8 + (1u << 20); }

int run_solve(SolveCtx &sx, const Scene &S, const rt_band_request &q, const SunSet &suns, float *R_out,
              int out_on_device, const DevMaterials *dm) {
    hipStream_t st = sx.stream;
    const uint32_t P = S.h.P, B = q.nbands, K = suns.K, Sn = q.nsources, C = K * B;
    if (B == 0 || K == 0) { set_error("run_band: no bands or no suns"); return -1; }
    if (C > 256) { set_error("run_band: more than 256 channels (suns x bands) in one solve"); return -1; }
    for (auto &e : sx.ev) { if (e.a) (void)hipEventDestroy(e.a); if (e.b) (void)hipEventDestroy(e.b); }
    sx.ev.clear();
    sx.timing_info = rt_timing{};
    sx.last_direct_grid = false;
    uint32_t maxdepth = 0;
    int scattering = 0, emission = 0;
    for (uint32_t b = 0; b < B; b++) {
        maxdepth = std::max(maxdepth, q.band_depth[b]);
        if (q.band_depth[b] > 0) scattering = 1;
        if (q.emission[b]) emission = 1;
    }
    std::vector<uint8_t> rundirect(K, 0), diffuse(K, 0);
    for (uint32_t k = 0; k < K; k++) {
        for (uint32_t s = 0; s < Sn; s++)
            for (uint32_t b = 0; b < B; b++)
                if (suns.flux[((size_t)k * Sn + s) * B + b] > 0.f) rundirect[k] = 1;
        for (uint32_t b = 0; b < B; b++)
            if (suns.dflux[(size_t)k * B + b] > 0.f) diffuse[k] = 1;
    }
    const uint32_t d = (uint32_t)std::ceil(std::sqrt((double)q.direct_rays));
    const uint32_t n = (uint32_t)std::ceil(std::sqrt((double)q.diffuse_rays));
    const uint32_t nn = n * n;
    bool any_direct = false, any_primary = false;
    for (uint32_t k = 0; k < K; k++) {
        if (rundirect[k] && Sn > 0 && d > 0) any_direct = true;
        if (emission || diffuse[k]) any_primary = true;
    }
    bool need_hits = (any_primary || scattering) && n > 0;
    // one direct launch covers every (sun, source, primitive) work item with a 32-bit work counter
    if (any_direct && check_work_items((uint64_t)K * Sn * P, "run_band: direct pass (suns x sources x primitives)"))
        return -1;

    // ---- masks [nm][C] ----
    enum { M_RD, M_AFTER, M_PRIM, M_EMIT, M_POST, M_NONE, M_BASE };
    const uint32_t nm = M_BASE + 4 * std::max(maxdepth, 1u);
    std::vector<uint8_t> mk((size_t)nm * C, 0);
    auto M = [&](uint32_t m, uint32_t c) -> uint8_t & { return mk[(size_t)m * C + c]; };
    for (uint32_t c = 0; c < C; c++) {
        uint32_t k = c / B, b = c % B;
        bool rd = rundirect[k] && Sn > 0 && d > 0;
        bool prim = emission || diffuse[k];
        M(M_RD, c) = rd;
        M(M_AFTER, c) = scattering && rundirect[k];
        M(M_PRIM, c) = prim && n > 0;
        M(M_EMIT, c) = prim && q.emission[b];
        M(M_POST, c) = scattering && prim && !rundirect[k];
        bool loop = scattering && (emission || diffuse[k] || rundirect[k]);
        bool was_active = loop;
        for (uint32_t s = 0; s < maxdepth; s++) {
            bool active = loop && (s + 1 <= q.band_depth[b]);
            bool retire = (q.schedule == 1) && was_active && !active;
            bool copy = loop && ((s > 0) || (q.schedule == 1 ? (prim && rundirect[k]) : (emission && rundirect[k])));
            M(M_BASE + 4 * s + 0, c) = retire;
            M(M_BASE + 4 * s + 1, c) = copy;
            M(M_BASE + 4 * s + 2, c) = loop;
            M(M_BASE + 4 * s + 3, c) = active && n > 0;
            was_active = active;
        }
    }
    std::vector<float> dfl((size_t)C, 0.f);
    for (uint32_t c = 0; c < C; c++) dfl[c] = suns.dflux[(size_t)(c / B) * B + c % B];
    // anisotropic sky of each channel (power law and/or Prague), when any band has one
    std::vector<float> hsky;
    if (sky_table(q, suns, K, hsky)) return -1;

    // translucent covers: primitives glass in every band of the request (direct_raygen.comp:629-637)
    std::vector<uint8_t> hcover;
    bool glass = false;
    if (q.is_glass && P) {
        if (!q.glass_n) { set_error("run_band: is_glass needs glass_n"); return -1; }
        hcover.assign(P, 0);
        for (uint32_t i = 0; i < P; i++) {
            bool all = true;
            for (uint32_t b = 0; b < B; b++) {
                const size_t k = (size_t)i * B + b;
                if (!q.is_glass[k]) { all = false; continue; }
                if (!(q.glass_n[k] >= 1.0f) || (q.glass_KL && !(q.glass_KL[k] >= 0.0f))) {
                    set_error("run_band: primitive " + std::to_string(i) + " band " + std::to_string(b) +
                              ": glass needs a refractive index >= 1 and an absorption K*L >= 0");
                    return -1;
                }
            }
            hcover[i] = all ? 1 : 0;
            glass = glass || all;
        }
        if (glass && C > RT_GLASS_MAX_CH) {
            set_error("run_band: a solve with translucent covers takes at most " + std::to_string(RT_GLASS_MAX_CH) +
                      " channels (suns x bands), this one has " + std::to_string(C));
            return -1;
        }
    }
    if (glass) need_hits = false;  // covers trace and gather in one kernel

    // ---- workspace ----
    size_t PC = (size_t)P * C;
    size_t bytes = solve_workspace_bytes(P, q, K, dm != nullptr, sx.energy_acct);
    if (sx.ws.reserve(bytes)) return -1;
    Carve cv{(char *)sx.ws.base};
    float *in = cv.take<float>(PC), *stb = cv.take<float>(PC), *sbb = cv.take<float>(PC);
    float *ot = cv.take<float>(PC), *ob = cv.take<float>(PC);
    uint32_t *hits = need_hits ? cv.take<uint32_t>((size_t)P * 2 * nn) : nullptr;
    float *dsum = cv.take<float>((size_t)K * Sn * P * 2 + 1);
    // optical properties: resident device copies when the caller has them, else staged per solve
    const float *rho = dm ? dm->rho : cv.take<float>((size_t)P * B);
    const float *tau = dm ? dm->tau : cv.take<float>((size_t)P * B);
    const float *eps = dm ? dm->eps : (q.eps ? cv.take<float>((size_t)P * B) : nullptr);
    const float *temp = dm ? dm->temp : (q.temperature ? cv.take<float>(P) : nullptr);
    float *flux = cv.take<float>((size_t)K * Sn * B + 1);
    float *dflux = cv.take<float>(C), *zflux = cv.take<float>(C);
    uint8_t *masks = cv.take<uint8_t>(mk.size());
    uint32_t *depth = cv.take<uint32_t>(B);
    double *area = cv.take<double>(P);
    const bool acct = sx.energy_acct && !glass;  // the bookkeeping follows the hit-table passes only
    double *energy = cv.take<double>((size_t)C * RT_ENERGY_ROWS);
    uint8_t *gcover = glass ? cv.take<uint8_t>(P) : nullptr;
    float *ggn = glass ? cv.take<float>((size_t)P * B) : nullptr;
    float *ggkl = glass ? cv.take<float>((size_t)P * B) : nullptr;
    double *gpush = glass ? cv.take<double>(PC * 3) : nullptr;
    float *ggsum = glass ? cv.take<float>((size_t)K * Sn * P * B * 2 + 1) : nullptr;
    GlassDirectJob *gdjobs = glass ? cv.take<GlassDirectJob>((size_t)K * Sn + 1) : nullptr;
    GlassCtx *gctx = glass ? cv.take<GlassCtx>(1) : nullptr;
    float *eacct = acct ? cv.take<float>(PC * 2) : nullptr;
    uint32_t *emiss = acct ? cv.take<uint32_t>((size_t)P * 2) : nullptr;
    double *epart = acct ? cv.take<double>((size_t)ENERGY_BLOCKS * 4 * C) : nullptr;
    double *epass = acct ? cv.take<double>((size_t)(1 + maxdepth) * 4 * C) : nullptr;
    DirectJob *djobs = cv.take<DirectJob>((size_t)K * Sn + 1);
    DiffuseJob *fjobs = cv.take<DiffuseJob>(1 + maxdepth);
    GatherJob *gjobs = cv.take<GatherJob>(1 + maxdepth);
    float *dsky = hsky.empty() ? nullptr : cv.take<float>(hsky.size());
    uint32_t *counter = cv.take<uint32_t>(16);
    int *err = reinterpret_cast<int *>(counter + 8);

    // ---- uploads (pageable host -> device, stream ordered) ----
    CK(hipMemsetAsync(in, 0, PC * 4, st));
    CK(hipMemsetAsync(stb, 0, PC * 4, st));
    CK(hipMemsetAsync(sbb, 0, PC * 4, st));
    CK(hipMemsetAsync(ot, 0, PC * 4, st));
    CK(hipMemsetAsync(ob, 0, PC * 4, st));
    CK(hipMemsetAsync(err, 0, sizeof(int), st));
    if (!dm) {
        CK(hipMemcpyAsync((void *)rho, q.rho, (size_t)P * B * 4, hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync((void *)tau, q.tau, (size_t)P * B * 4, hipMemcpyHostToDevice, st));
        if (eps) CK(hipMemcpyAsync((void *)eps, q.eps, (size_t)P * B * 4, hipMemcpyHostToDevice, st));
        if (temp) CK(hipMemcpyAsync((void *)temp, q.temperature, (size_t)P * 4, hipMemcpyHostToDevice, st));
    }
    if (Sn) CK(hipMemcpyAsync(flux, suns.flux, (size_t)K * Sn * B * 4, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(dflux, dfl.data(), (size_t)C * 4, hipMemcpyHostToDevice, st));
    CK(hipMemsetAsync(zflux, 0, (size_t)C * 4, st));
    if (dsky) CK(hipMemcpyAsync(dsky, hsky.data(), hsky.size() * 4, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(masks, mk.data(), mk.size(), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(depth, q.band_depth, (size_t)B * 4, hipMemcpyHostToDevice, st));
    GlassCtx hg_ctx{};
    if (glass) {
        CK(hipMemcpyAsync(gcover, hcover.data(), P, hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(ggn, q.glass_n, (size_t)P * B * 4, hipMemcpyHostToDevice, st));
        if (q.glass_KL) CK(hipMemcpyAsync(ggkl, q.glass_KL, (size_t)P * B * 4, hipMemcpyHostToDevice, st));
        else CK(hipMemsetAsync(ggkl, 0, (size_t)P * B * 4, st));
        CK(hipMemsetAsync(gpush, 0, PC * 3 * 8, st));
        CK(hipMemsetAsync(ggsum, 0, ((size_t)K * Sn * P * B * 2 + 1) * 4, st));
        hg_ctx.cover = gcover; hg_ctx.gn = ggn; hg_ctx.gkl = ggkl; hg_ctx.push = gpush; hg_ctx.B = B; hg_ctx.C = C;
        CK(hipMemcpyAsync(gctx, &hg_ctx, sizeof(GlassCtx), hipMemcpyHostToDevice, st));
    }

    std::vector<DirectJob> hd;
    uint32_t items = 0, ncol = 0;
    std::vector<SourceShape> shapes;
    if (source_shapes(Sn, q.source_type, q.source_width, q.source_rotation, shapes, "run_band")) return -1;
    if (any_direct) {
        // collimated sources first (k_direct), then the other types (k_direct_src)
        for (int pass = 0; pass < 2; pass++)
            for (uint32_t k = 0; k < K; k++) {
                if (!rundirect[k]) continue;
                for (uint32_t s = 0; s < Sn; s++) {
                    if ((shapes[s].type == 0u) != (pass == 0)) continue;
                    DirectJob j{};
                    j.sc = S.d;
                    const float *sd = suns.dir + ((size_t)k * Sn + s) * 3;
                    j.sdir[0] = sd[0]; j.sdir[1] = sd[1]; j.sdir[2] = sd[2];
                    j.d = d;
                    j.seed = suns.seed[k];
                    j.out = dsum + ((size_t)k * Sn + s) * P * 2;
                    j.item_begin = items;
                    j.P = P;
                    j.order = S.h.order;  // primitives in leaf order: concurrent waves trace nearby rays
                    set_job_source(j, shapes, s);
                    items += P;
                    hd.push_back(j);
                    if (pass == 0) ncol++;
                }
            }
        // occlusion grids, one per collimated (sun, source) direction, when the scene is not periodic
        sx.last_direct_grid = false;
        if (sx.use_grid && ncol && !glass) {
            std::vector<float> jd((size_t)ncol * 3);
            for (size_t j = 0; j < ncol; j++)
                for (int a = 0; a < 3; a++) jd[j * 3 + a] = hd[j].sdir[a];
            hipEvent_t ga, gb;
            if (ev_pair(sx, ga, gb)) return -1;
            if (ga) CK(hipEventRecord(ga, st));
            sx.grids.density = (float)sx.cfg[0].grid_density;
            int g = build_sun_grids(sx.grids, S, jd.data(), ncol, st);
            if (g < 0) return -1;
            if (gb) { CK(hipEventRecord(gb, st)); tick(sx, 3, ga, gb); }
            if (g == 0) {
                for (size_t j = 0; j < ncol; j++) hd[j].grid = sx.grids.dev + j;
                sx.last_direct_grid = true;
                const double per = (double)sx.grids.entries / ((double)ncol * std::max<uint32_t>(1u, S.h.nleaf));
                sx.last_grid_entries = (uint32_t)std::min(4.0e9, std::round(per * 100.0));
            }
        }
        CK(hipMemcpyAsync(djobs, hd.data(), hd.size() * sizeof(DirectJob), hipMemcpyHostToDevice, st));
    }
    std::vector<GlassDirectJob> hgd;
    if (glass && any_direct) {
        for (const DirectJob &j : hd) {
            GlassDirectJob g{};
            g.dj = j;
            // job (sun k, source s): its outputs were laid out at dsum + (k * Sn + s) * P * 2
            const size_t ks = (size_t)(j.out - dsum) / ((size_t)P * 2);
            g.flux = flux + ks * B;
            g.gsum = ggsum + ks * P * B * 2;
            g.chan0 = (uint32_t)(ks / Sn) * B;
            hgd.push_back(g);
        }
        CK(hipMemcpyAsync(gdjobs, hgd.data(), hgd.size() * sizeof(GlassDirectJob), hipMemcpyHostToDevice, st));
    }
    std::vector<DiffuseJob> hf(1 + maxdepth);
    for (uint32_t g = 0; g < hf.size(); g++) {
        DiffuseJob &fj = hf[g];
        fj.sc = S.d; fj.n = n; fj.seed = q.diffuse_seed[g]; fj.hits = hits; fj.item_begin = 0; fj.P = P;
        fj.order = S.h.order;
    }
    CK(hipMemcpyAsync(fjobs, hf.data(), hf.size() * sizeof(DiffuseJob), hipMemcpyHostToDevice, st));
    std::vector<GatherJob> hg(1 + maxdepth);
    for (uint32_t g = 0; g < hg.size(); g++) {
        GatherJob &J = hg[g];
        J.sc = S.d; J.hits = hits; J.n = n; J.C = C; J.B = B;
        J.out_top = ot; J.out_bot = ob;
        J.dflux = g == 0 ? dflux : zflux;
        J.active = g == 0 ? masks + (size_t)M_PRIM * C : masks + (size_t)(M_BASE + 4 * (g - 1) + 3) * C;
        J.rho = rho; J.tau = tau; J.in = in; J.st = stb; J.sb = sbb;
        J.item_begin = 0; J.P = P;
        J.order = S.h.order;
        J.sky = g == 0 ? dsky : nullptr;  // the sky is seen by the primary diffuse pass only
        J.seed = q.diffuse_seed[g];
        J.acct = eacct;
        J.miss = emiss;
    }
    CK(hipMemcpyAsync(gjobs, hg.data(), hg.size() * sizeof(GatherJob), hipMemcpyHostToDevice, st));
    // the host vectors above must outlive the async copies
    CK(hipStreamSynchronize(st));

    k_area<<<nb(P), 256, 0, st>>>(P, S.h.T, S.h.type, area);
    LK();

    // ---- DIRECT (RadiationModel.cpp:3847-3982) ----
    if (any_direct) {
        hipEvent_t a, b;
        if (ev_pair(sx, a, b)) return -1;
        if (a) CK(hipEventRecord(a, st));
        if (glass) {
            if (launch_direct_glass(gdjobs, (uint32_t)hgd.size(), items, S.h.depth, counter, gctx, err, sx.ncu, st,
                                    scene_flags(S, sx)))
                return -1;
        } else if (launch_direct_jobs(hd.data(), djobs, (uint32_t)hd.size(), items, S.h.depth, counter, err,
                                      sx.stat_ptr(0), sx.cfg[0], sx.ncu, st, scene_flags(S, sx))) {
            return -1;
        }
        if (b) { CK(hipEventRecord(b, st)); tick(sx, 0, a, b); }
        sx.timing_info.direct_rays += (uint64_t)items * d * d;
        sx.timing_info.direct_launches++;
        k_deposit_direct<<<nb(PC), 256, 0, st>>>(P, C, B, Sn, dsum, flux, rho, tau, masks + (size_t)M_RD * C, in,
                                                   stb, sbb, glass ? ggsum : nullptr);
        LK();
        if (glass && launch_add_push(PC, gpush, in, stb, sbb, st)) return -1;
    }
    if (scattering) {
        k_after_direct<<<nb(PC), 256, 0, st>>>(P, C, S.h.nrm, masks + (size_t)M_AFTER * C, stb, sbb, ot, ob);
        LK();
    }
    uint32_t npass = 0;  // passes with energy rows
    auto diffuse_pass = [&](uint32_t g) -> int {
        hipEvent_t a, b;
        if (ev_pair(sx, a, b)) return -1;
        if (a) CK(hipEventRecord(a, st));
        if (glass) {
            if (launch_diffuse_glass(fjobs + g, gjobs + g, P, S.h.depth, counter, gctx, err, sx.ncu, st,
                                     scene_flags(S, sx)))
                return -1;
            if (b) { CK(hipEventRecord(b, st)); tick(sx, 1, a, b); }
            sx.timing_info.diffuse_rays += (uint64_t)P * 2 * nn;
            sx.timing_info.diffuse_launches++;
            return launch_add_push(PC, gpush, in, stb, sbb, st);
        }
        if (launch_diffuse(fjobs + g, P, S.h.depth, counter, err, sx.stat_ptr(1), sx.cfg[1], sx.ncu, st,
                           scene_flags(S, sx)))
            return -1;
        if (b) { CK(hipEventRecord(b, st)); tick(sx, 1, a, b); }
        sx.timing_info.diffuse_rays += (uint64_t)P * 2 * nn;
        sx.timing_info.diffuse_launches++;
        hipEvent_t c2, d2;
        if (ev_pair(sx, c2, d2)) return -1;
        if (c2) CK(hipEventRecord(c2, st));
        if (launch_gather(gjobs + g, 1, P, C, counter, sx.cfg[2], sx.ncu, st, acct)) return -1;
        if (d2) { CK(hipEventRecord(d2, st)); tick(sx, 2, c2, d2); }
        if (acct) {
            const uint8_t *act = hg[g].active;
            k_energy_pass<<<ENERGY_BLOCKS, 256, 0, st>>>(P, C, nn, area, S.h.nrm, S.h.type, act, eacct, emiss, ot, ob,
                                                         epart);
            LK();
            k_energy_sum<<<1, 256, 0, st>>>(ENERGY_BLOCKS, C, epart, epass + (size_t)npass * 4 * C);
            LK();
            npass++;
        }
        return 0;
    };
    // ---- EMISSION + PRIMARY DIFFUSE (RadiationModel.cpp:4040-4222) ----
    if (any_primary && n > 0) {
        if (emission) {
            k_emission<<<nb(PC), 256, 0, st>>>(P, C, B, S.h.nrm, eps, temp, depth, masks + (size_t)M_EMIT * C, ot, ob);
            LK();
        }
        if (diffuse_pass(0)) return -1;
    }
    if (scattering) {
        // !rundirect: radiation_out := scatter after the primary pass (RadiationModel.cpp:4225-4234)
        const uint8_t *none = masks + (size_t)M_NONE * C;  // all-zero row
        k_copy_zero<<<nb(PC), 256, 0, st>>>(P, C, none, masks + (size_t)M_POST * C, masks + (size_t)M_POST * C, in,
                                             stb, sbb, ot, ob);
        LK();
        // ---- SCATTERING ITERATIONS (RadiationModel.cpp:4236-4356) ----
        for (uint32_t s = 0; s < maxdepth; s++) {
            const uint8_t *row = masks + (size_t)(M_BASE + 4 * s) * C;
            k_copy_zero<<<nb(PC), 256, 0, st>>>(P, C, row, row + C, row + 2 * C, in, stb, sbb, ot, ob);
            LK();
            bool any_active = false;
            for (uint32_t c = 0; c < C; c++) any_active |= mk[(size_t)(M_BASE + 4 * s + 3) * C + c] != 0;
            if (any_active && n > 0)
                if (diffuse_pass(1 + s)) return -1;
        }
    }
    // ---- final (RadiationModel.cpp:4631-4654) ----
    float *R = R_out;
    float *Rd = nullptr;
    if (!out_on_device) {
        Rd = ot;  // reuse a [P][C] buffer: radiation_out is dead here
        R = Rd;
    }
    k_final<<<nb(PC), 256, 0, st>>>(P, C, B, in, stb, sbb, R);
    LK();
    sx.energy_valid = false;
    sx.energy_glass = sx.energy_acct && glass;
    if (acct) {
        CK(hipMemsetAsync(energy, 0, (size_t)C * RT_ENERGY_ROWS * 8, st));
        const bool emits = any_primary && n > 0 && emission;
        k_energy<<<C, 256, 0, st>>>(P, C, B, Sn, area, R, any_direct ? dsum : nullptr, flux, stb, sbb, S.h.nrm,
                                    emits ? masks + (size_t)M_EMIT * C : nullptr, eps, temp, depth, energy);
        LK();
    }
    if (!out_on_device) CK(hipMemcpyAsync(R_out, Rd, PC * 4, hipMemcpyDeviceToHost, st));
    if (acct) {
        sx.energy.assign((size_t)C * RT_ENERGY_ROWS, 0.0);
        sx.energy_passes.assign((size_t)npass * 4 * C, 0.0);
        CK(hipMemcpyAsync(sx.energy.data(), energy, (size_t)C * RT_ENERGY_ROWS * 8, hipMemcpyDeviceToHost, st));
        if (npass)
            CK(hipMemcpyAsync(sx.energy_passes.data(), epass, (size_t)npass * 4 * C * 8, hipMemcpyDeviceToHost, st));
    }
    CK(hipStreamSynchronize(st));
    if (acct) {
        // sky (row 1) and escaped (row 4) from the passes; residual (row 6) = in - out
        for (uint32_t c = 0; c < C; c++) {
            double sky = 0.0, esc = 0.0;
            for (uint32_t g = 0; g < npass; g++) {
                sky += sx.energy_passes[((size_t)g * 4 + 2) * C + c];
                esc += sx.energy_passes[((size_t)g * 4 + 3) * C + c];
            }
            double *e = sx.energy.data();
            e[1 * C + c] = sky;
            e[4 * C + c] = esc;
            e[6 * C + c] = (e[0 * C + c] + sky + e[2 * C + c]) - (e[3 * C + c] + esc);
        }
        sx.energy_C = C;
        sx.energy_npass = npass;
        sx.energy_valid = true;
    }
    int herr = 0;
    CK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) { set_error("traversal error code " + std::to_string(herr) + " (1 = stack overflow, 2 = malformed tree)"); return -1; }
    for (auto &e : sx.ev) {
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e.a, e.b));
        if (e.cls == 0) sx.timing_info.direct_ms += ms;
        else if (e.cls == 1) sx.timing_info.diffuse_ms += ms;
        else if (e.cls == 2) sx.timing_info.gather_ms += ms;
        else sx.timing_info.other_ms += ms;
    }
    return 0;
}

int debug_diffuse_hits(SolveCtx &sx, const Scene &S, uint32_t n, uint32_t seed, uint32_t *out) {
    hipStream_t st = sx.stream;
    const uint32_t P = S.h.P, nn = n * n;
    size_t bytes = (size_t)P * 2 * nn * 4 + sizeof(DiffuseJob) + 4096;
    if (sx.ws.reserve(bytes)) return -1;
    Carve cv{(char *)sx.ws.base};
    uint32_t *hits = cv.take<uint32_t>((size_t)P * 2 * nn);
    DiffuseJob *fjob = cv.take<DiffuseJob>(1);
    uint32_t *counter = cv.take<uint32_t>(16);
    int *err = reinterpret_cast<int *>(counter + 8);
    DiffuseJob fj{};
    fj.sc = S.d; fj.n = n; fj.seed = seed; fj.hits = hits; fj.item_begin = 0; fj.P = P;
    CK(hipMemsetAsync(err, 0, 4, st));
    CK(hipMemcpyAsync(fjob, &fj, sizeof(fj), hipMemcpyHostToDevice, st));
    if (launch_diffuse(fjob, P, S.h.depth, counter, err, sx.stat_ptr(1), sx.cfg[1], sx.ncu, st, scene_flags(S, sx)))
        return -1;
    CK(hipMemcpyAsync(out, hits, (size_t)P * 2 * nn * 4, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    int herr = 0;
    CK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) { set_error("traversal error code " + std::to_string(herr) + " (1 = stack overflow, 2 = malformed tree)"); return -1; }
    return 0;
}

int debug_direct_sums(SolveCtx &sx, const Scene &S, uint32_t d, uint32_t seed, uint32_t nsrc, const float *src_dir,
                      const uint32_t *types, const float *widths, const float *rotations, float *out) {
    hipStream_t st = sx.stream;
    const uint32_t P = S.h.P;
    std::vector<SourceShape> shapes;
    if (source_shapes(nsrc, types, widths, rotations, shapes, "rt_debug_direct_sums")) return -1;
    if (check_work_items((uint64_t)nsrc * P, "rt_debug_direct_sums (sources x primitives)")) return -1;
    size_t bytes = (size_t)nsrc * P * 2 * 4 + nsrc * sizeof(DirectJob) + 4096;
    if (sx.ws.reserve(bytes)) return -1;
    Carve cv{(char *)sx.ws.base};
    float *sums = cv.take<float>((size_t)nsrc * P * 2);
    DirectJob *jobs = cv.take<DirectJob>(nsrc);
    uint32_t *counter = cv.take<uint32_t>(16);
    int *err = reinterpret_cast<int *>(counter + 8);
    // collimated sources first (their items precede the others')
    std::vector<DirectJob> hj;
    std::vector<float> coldir;
    for (int pass = 0; pass < 2; pass++)
        for (uint32_t s = 0; s < nsrc; s++) {
            if ((shapes[s].type == 0u) != (pass == 0)) continue;
            DirectJob j{};
            j.sc = S.d;
            j.sdir[0] = src_dir[s * 3]; j.sdir[1] = src_dir[s * 3 + 1]; j.sdir[2] = src_dir[s * 3 + 2];
            j.d = d; j.seed = seed; j.out = sums + (size_t)s * P * 2; j.item_begin = (uint32_t)hj.size() * P; j.P = P;
            set_job_source(j, shapes, s);
            if (pass == 0) coldir.insert(coldir.end(), j.sdir, j.sdir + 3);
            hj.push_back(j);
        }
    const uint32_t ncol = (uint32_t)(coldir.size() / 3);
    if (sx.use_grid && ncol) {
        sx.grids.density = (float)sx.cfg[0].grid_density;
        int g = build_sun_grids(sx.grids, S, coldir.data(), ncol, st);
        if (g < 0) return -1;
        if (g == 0)
            for (uint32_t s = 0; s < ncol; s++) hj[s].grid = sx.grids.dev + s;
    }
    CK(hipMemsetAsync(err, 0, 4, st));
    CK(hipMemcpyAsync(jobs, hj.data(), nsrc * sizeof(DirectJob), hipMemcpyHostToDevice, st));
    if (launch_direct_jobs(hj.data(), jobs, nsrc, nsrc * P, S.h.depth, counter, err, sx.stat_ptr(0), sx.cfg[0], sx.ncu,
                           st, scene_flags(S, sx)))
        return -1;
    std::vector<float> tmp((size_t)nsrc * P * 2);
    CK(hipMemcpyAsync(tmp.data(), sums, tmp.size() * 4, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    int herr = 0;
    CK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) { set_error("traversal error code " + std::to_string(herr) + " (1 = stack overflow, 2 = malformed tree)"); return -1; }
    // [S][P][2] -> [P][S][2]
    for (uint32_t s = 0; s < nsrc; s++)
        for (uint32_t i = 0; i < P; i++) {
            out[((size_t)i * nsrc + s) * 2] = tmp[((size_t)s * P + i) * 2];
            out[((size_t)i * nsrc + s) * 2 + 1] = tmp[((size_t)s * P + i) * 2 + 1];
        }
    return 0;
}

}  // namespace rt
