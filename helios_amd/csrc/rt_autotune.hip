// rt_autotune.hip -- launch-shape search for the persistent ray kernels.
//
// Candidates (waves per workgroup x workgroups per CU x work-queue chunk) are timed with HIP events on a
// prefix sample of the caller's scene: a direct pass towards the request's first source, a diffuse pass,
// and a gather over the resulting hit table with the batch's channel count.  The fastest shape per kernel
// class is installed on the context.  The sample keeps every launch within tens of milliseconds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "rt_ctx.h"
#include "rt_solve.h"

using namespace rt;

namespace {

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Runs `launch` twice with `cfg` and returns the second run's time in ms (-1 on failure).
template <typename F>
double time_cfg(hipStream_t st, Timer &tm, const LaunchCfg &cfg, F launch) {
    if (launch(cfg)) return -1.0;
    if (hipEventRecord(tm.a, st) != hipSuccess) return -1.0;
    if (launch(cfg)) return -1.0;
    if (hipEventRecord(tm.b, st) != hipSuccess) return -1.0;
    if (hipEventSynchronize(tm.b) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, tm.a, tm.b) != hipSuccess) return -1.0;
    return ms;
}

template <typename F>
int search(hipStream_t st, Timer &tm, const std::vector<LaunchCfg> &cands, F launch, const LaunchCfg &def,
           LaunchCfg &best, double &best_ms, double &default_ms, uint32_t &count) {
    default_ms = time_cfg(st, tm, def, launch);
    if (default_ms < 0) return -1;
    best = def;
    best_ms = default_ms;
    count++;
    for (const LaunchCfg &c : cands) {
        double ms = time_cfg(st, tm, c, launch);
        if (ms < 0) return -1;
        count++;
        if (ms < 0.97 * best_ms) {  // keep the default unless a shape is clearly faster
            best = c;
            best_ms = ms;
        }
    }
    return 0;
}

}  // namespace

extern "C" int rt_autotune(rt_ctx *ctx, const rt_scene *scene, const rt_band_request *req, uint32_t nsuns,
                           rt_tune_result *out) {
    NN(ctx, "ctx");
    NN(scene, "scene");
    NN(req, "request");
    NN(out, "out");
    if (req->nbands == 0 || nsuns == 0) { set_error("rt_autotune: no bands or no suns"); return -1; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    CKR(hipSetDevice(ctx->device));
    const Scene &S = scene->s;
    const uint32_t P = S.h.P;
    if (P == 0) { set_error("rt_autotune: empty scene"); return -1; }
    const uint32_t sample = std::min<uint32_t>(P, std::max<uint32_t>(65536u, P / 8));
    const uint32_t C = std::min<uint32_t>(256u, nsuns * req->nbands);
    const uint32_t d = std::max(1u, (uint32_t)std::ceil(std::sqrt((double)req->direct_rays)));
    const uint32_t n = std::max(1u, (uint32_t)std::ceil(std::sqrt((double)req->diffuse_rays)));
    const size_t nn = (size_t)n * n;
    hipStream_t st = ctx->sx.stream;

    // scratch: direct sums, hit table (sample rows, indexed by primitive), gather buffers
    size_t bytes = (size_t)P * 2 * 4 + (size_t)sample * 2 * nn * 4 + 2 * (size_t)P * C * 4 + 3 * (size_t)sample * C * 4 +
                   2 * (size_t)sample * C * 4 + (size_t)C * 8 + 4096 * 4;
    void *mem = nullptr;
    CKR(hipMalloc(&mem, bytes));
    char *p = (char *)mem;
    auto take = [&](size_t b) {
        char *r = p;
        p += (b + 255) & ~(size_t)255;
        return r;
    };
    float *dsum = (float *)take((size_t)P * 2 * 4);
    uint32_t *hits = (uint32_t *)take((size_t)sample * 2 * nn * 4);
    float *otop = (float *)take((size_t)P * C * 4), *obot = (float *)take((size_t)P * C * 4);
    float *in = (float *)take((size_t)sample * C * 4), *stb = (float *)take((size_t)sample * C * 4),
          *sbb = (float *)take((size_t)sample * C * 4);
    float *rho = (float *)take((size_t)sample * C * 4), *tau = (float *)take((size_t)sample * C * 4);
    float *dfl = (float *)take((size_t)C * 4);
    uint8_t *act = (uint8_t *)take(C);
    DirectJob *dj = (DirectJob *)take(sizeof(DirectJob));
    DiffuseJob *fj = (DiffuseJob *)take(sizeof(DiffuseJob));
    GatherJob *gj = (GatherJob *)take(sizeof(GatherJob));
    uint32_t *counter = (uint32_t *)take(64);
    int *err = (int *)take(64);

    int rc = 0;
    Timer tm;
    if (hipEventCreate(&tm.a) != hipSuccess || hipEventCreate(&tm.b) != hipSuccess) rc = -1;
    if (rc == 0) {
        DirectJob hdj{};
        hdj.sc = S.d;
        // the request's first source direction, or an oblique default when it has none (a batch's base
        // request carries no sun of its own)
        const float up[3] = {0.3f, 0.2f, 0.93f};
        const bool col0 = !req->source_type || (req->nsources && req->source_type[0] == 0u);
        const float *sd = (req->nsources && req->source_dir && col0) ? req->source_dir : up;
        if (!(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2] > 1e-12f)) sd = up;
        hdj.sdir[0] = sd[0]; hdj.sdir[1] = sd[1]; hdj.sdir[2] = sd[2];
        hdj.d = d; hdj.seed = 12345u; hdj.out = dsum; hdj.item_begin = 0; hdj.P = P;
        DiffuseJob hfj{};
        hfj.sc = S.d; hfj.n = n; hfj.seed = 777u; hfj.hits = hits; hfj.item_begin = 0; hfj.P = P;
        GatherJob hgj{};
        hgj.sc = S.d; hgj.hits = hits; hgj.n = n; hgj.C = C; hgj.B = C;
        hgj.out_top = otop; hgj.out_bot = obot; hgj.dflux = dfl; hgj.active = act; hgj.rho = rho; hgj.tau = tau;
        hgj.in = in; hgj.st = stb; hgj.sb = sbb; hgj.item_begin = 0; hgj.P = P;
        std::vector<float> ones((size_t)C, 1.f);
        std::vector<uint8_t> acts((size_t)C, 1);
        if (hipMemsetAsync(otop, 0, (size_t)P * C * 4, st) != hipSuccess ||
            hipMemsetAsync(obot, 0, (size_t)P * C * 4, st) != hipSuccess ||
            hipMemsetAsync(rho, 0, (size_t)sample * C * 4, st) != hipSuccess ||
            hipMemsetAsync(tau, 0, (size_t)sample * C * 4, st) != hipSuccess ||
            hipMemsetAsync(in, 0, (size_t)sample * C * 4, st) != hipSuccess ||
            hipMemsetAsync(stb, 0, (size_t)sample * C * 4, st) != hipSuccess ||
            hipMemsetAsync(sbb, 0, (size_t)sample * C * 4, st) != hipSuccess ||
            hipMemsetAsync(err, 0, 4, st) != hipSuccess ||
            hipMemcpyAsync(dfl, ones.data(), (size_t)C * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(act, acts.data(), C, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(dj, &hdj, sizeof(hdj), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(fj, &hfj, sizeof(hfj), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(gj, &hgj, sizeof(hgj), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("rt_autotune: staging failed");
            rc = -1;
        }
    }
    const uint32_t depth = S.h.depth;
    const int ncu = ctx->sx.ncu;
    std::vector<LaunchCfg> trace_c, gather_c;
    for (uint32_t w : {1u, 2u, 4u})
        for (uint32_t b : {2u, 4u, 8u, 16u})
            for (uint32_t c : {1u, 4u}) trace_c.push_back(LaunchCfg{w, b, c});
    for (uint32_t b : {2u, 4u, 8u, 16u})
        for (uint32_t c : {1u, 4u, 16u}) gather_c.push_back(LaunchCfg{4u, b, c});
    LaunchCfg best[3];
    double best_ms[3] = {0, 0, 0}, def_ms[3] = {0, 0, 0};
    uint32_t count = 0;
    // occlusion-grid density of the direct pass (non-periodic scenes).  A density costs, summed over a few
    // sun directions (the request's first collimated source, if any, plus zenith 20, 50 and 75 degrees:
    // the lists grow as the sun gets lower), the grid build plus a direct pass over all P primitives with
    // the default shape (a full pass, not the sample: the list lengths depend on the whole scene and a short
    // launch is dominated by its tail).  The shapes are then searched at the chosen density.
    GridSet gs;
    uint32_t density = 2;
    bool use_grid = false;
    if (rc == 0 && ctx->sx.use_grid) {
        DirectJob hdj{};
        if (hipMemcpy(&hdj, dj, sizeof(hdj), hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
        std::vector<std::array<float, 3>> dirs;
        const bool own = req->nsources && req->source_dir && (!req->source_type || req->source_type[0] == 0u) &&
                         (req->source_dir[0] * req->source_dir[0] + req->source_dir[1] * req->source_dir[1] +
                              req->source_dir[2] * req->source_dir[2] > 1e-12f);
        if (own) dirs.push_back({req->source_dir[0], req->source_dir[1], req->source_dir[2]});
        const double zen[3] = {20.0, 50.0, 75.0}, az[3] = {30.0, 150.0, 260.0};
        for (int k = 0; k < 3; k++) {
            const double z = zen[k] * M_PI / 180.0, a = az[k] * M_PI / 180.0;
            dirs.push_back({(float)(std::sin(z) * std::sin(a)), (float)(std::sin(z) * std::cos(a)), (float)std::cos(z)});
        }
        double best_d = -1.0;
        for (uint32_t dens : {2u, 1u, 4u, 8u}) {  // the default first: others must beat it by 3%
            if (rc) break;
            gs.density = (float)dens;
            double cost = 0.0;
            bool applies = true;
            for (const auto &dv : dirs) {
                if (hipStreamSynchronize(st) != hipSuccess) { rc = -1; break; }
                const auto t0 = std::chrono::steady_clock::now();
                int g = build_sun_grids(gs, S, dv.data(), 1, st);
                if (g < 0) { rc = -1; break; }
                if (g > 0) { applies = false; break; }  // grids do not apply (or not at this density)
                const double build_ms =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                DirectJob j = hdj;
                j.sdir[0] = dv[0]; j.sdir[1] = dv[1]; j.sdir[2] = dv[2];
                j.grid = gs.dev;
                if (hipMemcpy(dj, &j, sizeof(j), hipMemcpyHostToDevice) != hipSuccess) { rc = -1; break; }
                const LaunchCfg def = default_launch_cfg(0);
                double ms = time_cfg(st, tm, def, [&](const LaunchCfg &c) {
                    return launch_direct(dj, 1, P, depth, counter, err, nullptr, c, ncu, st,
                                         scene_flags(S, ctx->sx) | RT_FL_GRIDS);
                });
                if (ms < 0) { rc = -1; break; }
                cost += ms + build_ms;
            }
            if (rc || !applies) break;
            count++;
            if (best_d < 0 || cost < 0.97 * best_d) { best_d = cost; density = dens; }
            use_grid = true;
        }
        if (rc == 0 && use_grid) {
            // the shape search below traces the request's (or the default) direction at the chosen density
            gs.density = (float)density;
            if (build_sun_grids(gs, S, hdj.sdir, 1, st) != 0) rc = -1;
            DirectJob j = hdj;
            j.grid = gs.dev;
            if (rc == 0 && hipMemcpy(dj, &j, sizeof(j), hipMemcpyHostToDevice) != hipSuccess) rc = -1;
        }
    }
    if (rc == 0)
        rc = search(st, tm, trace_c,
                    [&](const LaunchCfg &c) {
                        return launch_direct(dj, 1, sample, depth, counter, err, nullptr, c, ncu, st,
                                             scene_flags(S, ctx->sx) | (use_grid ? RT_FL_GRIDS : 0));
                    },
                    default_launch_cfg(0), best[0], best_ms[0], def_ms[0], count);
    best[0].grid_density = density;
    if (rc == 0)
        rc = search(st, tm, trace_c,
                    [&](const LaunchCfg &c) {
                        return launch_diffuse(fj, sample, depth, counter, err, nullptr, c, ncu, st, scene_flags(S, ctx->sx));
                    },
                    default_launch_cfg(1), best[1], best_ms[1], def_ms[1], count);
    if (rc == 0)
        rc = search(st, tm, gather_c,
                    [&](const LaunchCfg &c) { return launch_gather(gj, 1, sample, C, counter, c, ncu, st); },
                    default_launch_cfg(2), best[2], best_ms[2], def_ms[2], count);
    int herr = 0;
    if (rc == 0 && (hipMemcpy(&herr, err, 4, hipMemcpyDeviceToHost) != hipSuccess || herr != 0)) {
        set_error("rt_autotune: traversal error code " + std::to_string(herr));
        rc = -1;
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(mem);
    gs.release();
    if (rc != 0) return -1;
    for (int k = 0; k < 3; k++) {
        ctx->sx.cfg[k] = best[k];
        out->cfg[k].waves_per_block = best[k].waves_per_block;
        out->cfg[k].blocks_per_cu = best[k].blocks_per_cu;
        out->cfg[k].chunk = best[k].chunk;
        out->cfg[k].grid_density = best[k].grid_density;
        out->ms[k] = best_ms[k];
        out->default_ms[k] = def_ms[k];
    }
    out->candidates = count;
    out->sample_prims = sample;
    return 0;
}
