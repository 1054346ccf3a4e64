// rt_server.hip -- request server for radiation solves on one device: resident models, memory-footprint
// admission, continuous batching and overlapping execution.
//
// A model is a resident scene plus everything of a runBand request except the sun (bands, ray counts,
// scattering depths, optical properties in HBM, diffuse seeds).  A request is one sun for one model.
// Each of `slots` worker threads owns a HIP stream and a grow-only workspace.  When a slot is free it
// takes the oldest queued request and every other queued request of the same model (FIFO, up to the
// batch limit), shrinks that batch until its workspace fits the memory budget, and runs it as one batched
// solve (every diffuse pass traced once for all suns).  Requests that arrive while others run join the
// next batch a slot picks up; slots run concurrently on their own streams, so work of different batches
// overlaps on the device.  Each request's result is the same as a solve of that sun alone (the batched
// schedule keeps per-sun channel masks; same seeds give the same rays).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rt_ctx.h"

using namespace rt;

namespace {

using Clock = std::chrono::steady_clock;

struct Model {
    const Scene *S = nullptr;
    uint32_t P = 0, B = 0, Sn = 0;
    rt_band_request q{};  // host copy; pointers aim at the vectors below (rho/tau/eps/temp are null)
    std::vector<uint32_t> depth, emission, dseed, stype;
    std::vector<float> swidth, srot;
    std::vector<float> kext, kpeak, knorm, kprague;  // anisotropic sky of the model's bands
    std::vector<uint8_t> gis;                        // translucent covers [P][B] (empty: none)
    std::vector<float> gn, gkl;
    bool covers = false;   // some primitive is glass in every band: at most RT_GLASS_MAX_CH channels per batch
    void *dmem = nullptr;  // resident rho, tau[, eps][, temp]
    DevMaterials dm{};
    size_t resident_bytes = 0;
};

struct Request {
    uint64_t id = 0;
    uint32_t model = 0;
    std::vector<float> dir, flux, dflux;
    uint32_t seed = 0;
    float *out = nullptr;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = 0;
    std::string err;
};

struct Slot {
    SolveCtx sx;
    float *dout = nullptr;
    size_t dout_cap = 0;
    size_t accounted = 0;  // bytes of this slot counted against the budget
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct rt_server {
    rt_ctx *ctx = nullptr;
    int device = 0;
    uint32_t max_batch = 64;
    size_t budget = 0;
    size_t allocated = 0;
    size_t peak_allocated = 0;
    std::vector<std::unique_ptr<Model>> models;
    std::vector<std::unique_ptr<Slot>> slots;
    std::deque<Request *> queue;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::thread> workers;
    bool stop = false;
    uint32_t mem_waiters = 0;
    std::vector<double> lat_ms;
    uint64_t submitted = 0, completed = 0, failed = 0, batches = 0, batched_requests = 0;
    uint32_t max_batch_seen = 0;

    void worker(uint32_t w);
};

namespace {

void release_slot(rt_server &srv, Slot &sl) {
    sl.sx.ws.release();
    sl.sx.grids.release();
    if (sl.dout) (void)hipFree(sl.dout);
    sl.dout = nullptr;
    sl.dout_cap = 0;
    srv.allocated -= sl.accounted;
    sl.accounted = 0;
}

void finish(rt_server &srv, Request *r, int status, const std::string &err) {
    r->status = status;
    r->err = err;
    r->t_done = Clock::now();
    r->done = true;
    if (status == 0) {
        srv.completed++;
        srv.lat_ms.push_back(ms_between(r->t_submit, r->t_done));
    } else {
        srv.failed++;
    }
}

}  // namespace

void rt_server::worker(uint32_t w) {
    (void)hipSetDevice(device);
    Slot &sl = *slots[w];
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        cv_work.wait(lk, [&] { return stop || !queue.empty() || (mem_waiters > 0 && sl.accounted > 0); });
        if (queue.empty()) {
            if (stop) break;
            // woken only because another slot waits for memory: give back this slot's cache
            release_slot(*this, sl);
            cv_work.notify_all();
            continue;
        }
        Request *head = queue.front();
        Model &m = *models[head->model];
        uint32_t kmax = std::max(1u, std::min(max_batch, 256u / m.B));
        // the batch's direct pass must stay within one launch's work-item limit (a lone sun that exceeds it
        // fails in run_solve with that message)
        if (m.Sn) kmax = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(kmax, RT_MAX_WORK_ITEMS / ((uint64_t)m.Sn * m.P)));
        // a solve with translucent covers carries at most RT_GLASS_MAX_CH channels (suns x bands)
        if (m.covers) kmax = std::max(1u, std::min(kmax, (uint32_t)RT_GLASS_MAX_CH / m.B));
        std::vector<Request *> batch;
        for (Request *r : queue)
            if (r->model == head->model && batch.size() < kmax) batch.push_back(r);
        // shrink the batch until its footprint fits the budget
        size_t need_ws = 0, need_out = 0, new_acc = 0;
        bool fits = false;
        while (!batch.empty()) {
            const uint32_t K = (uint32_t)batch.size();
            need_ws = Workspace::capacity_for(solve_workspace_bytes(m.P, m.q, K, true)) +
                      Workspace::capacity_for(GridSet::budget(m.P, K * m.Sn)) + (3u << 20);
            need_out = (size_t)K * m.P * m.B * sizeof(float);
            const size_t have_ws = sl.sx.ws.cap + sl.sx.grids.small.cap + sl.sx.grids.cells.cap + sl.sx.grids.items.cap;
            new_acc = std::max(need_ws, have_ws) + std::max(need_out, sl.dout_cap);
            if (allocated - sl.accounted + new_acc <= budget) { fits = true; break; }
            if (K == 1) break;
            batch.pop_back();
        }
        if (!fits) {
            size_t alone = Workspace::capacity_for(solve_workspace_bytes(m.P, m.q, 1, true)) +
                           Workspace::capacity_for(GridSet::budget(m.P, m.Sn)) + (3u << 20) +
                           (size_t)m.P * m.B * sizeof(float);
            if (alone > budget) {
                queue.pop_front();
                finish(*this, head, -1,
                       "request needs " + std::to_string(alone) + " device bytes, more than the server's memory budget of " +
                           std::to_string(budget));
                cv_done.notify_all();
                continue;
            }
            if (sl.accounted > 0) {
                // this slot's own cache is what blocks it: drop it and retry
                release_slot(*this, sl);
                continue;
            }
            mem_waiters++;
            cv_work.notify_all();
            cv_work.wait_for(lk, std::chrono::milliseconds(2));
            mem_waiters--;
            continue;
        }
        for (Request *r : batch) queue.erase(std::find(queue.begin(), queue.end(), r));
        allocated += new_acc - sl.accounted;
        sl.accounted = new_acc;
        peak_allocated = std::max(peak_allocated, allocated);
        const Clock::time_point t_start = Clock::now();
        for (Request *r : batch) r->t_start = t_start;
        const uint32_t K = (uint32_t)batch.size();
        batches++;
        batched_requests += K;
        max_batch_seen = std::max(max_batch_seen, K);
        lk.unlock();

        // ---- run the batch (no lock held) ----
        std::string err;
        int rc = 0;
        const uint32_t Sn = m.Sn, B = m.B;
        std::vector<float> dir((size_t)K * Sn * 3), flux((size_t)K * Sn * B + 1), dfl((size_t)K * B);
        std::vector<uint32_t> seeds(K);
        for (uint32_t k = 0; k < K; k++) {
            const Request &r = *batch[k];
            std::copy(r.dir.begin(), r.dir.end(), dir.begin() + (size_t)k * Sn * 3);
            std::copy(r.flux.begin(), r.flux.end(), flux.begin() + (size_t)k * Sn * B);
            std::copy(r.dflux.begin(), r.dflux.end(), dfl.begin() + (size_t)k * B);
            seeds[k] = r.seed;
        }
        sl.sx.grids.cap = GridSet::budget(m.P, K * m.Sn);
        if (sl.sx.ws.reserve(solve_workspace_bytes(m.P, m.q, K, true))) rc = -1;
        if (rc == 0 && need_out > sl.dout_cap) {
            if (sl.dout) (void)hipFree(sl.dout);
            sl.dout = nullptr;
            sl.dout_cap = 0;
            if (hipMalloc(&sl.dout, need_out) != hipSuccess) {
                set_error("hipMalloc of the batch output failed");
                rc = -1;
            } else {
                sl.dout_cap = need_out;
            }
        }
        if (rc == 0) {
            static const float up[3] = {0.f, 0.f, 1.f};
            SunSet suns{K, Sn ? dir.data() : up, flux.data(), dfl.data(), seeds.data()};
            rc = run_solve(sl.sx, *m.S, m.q, suns, sl.dout, 1, &m.dm);
        }
        if (rc == 0) {
            const size_t one = (size_t)m.P * B;
            for (uint32_t k = 0; k < K && rc == 0; k++)
                if (hipMemcpyAsync(batch[k]->out, sl.dout + k * one, one * sizeof(float), hipMemcpyDeviceToHost,
                                   sl.sx.stream) != hipSuccess) {
                    set_error("copy of a result to the host failed");
                    rc = -1;
                }
            if (rc == 0 && hipStreamSynchronize(sl.sx.stream) != hipSuccess) {
                set_error("stream synchronisation failed");
                rc = -1;
            }
        }
        if (rc != 0) err = rt_last_error();

        lk.lock();
        for (Request *r : batch) finish(*this, r, rc, err);
        if (rc != 0 || mem_waiters > 0) release_slot(*this, sl);
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

int rt_server_create(rt_ctx *ctx, const rt_server_cfg *cfg, rt_server **out) {
    NN(ctx, "ctx");
    NN(out, "out");
    *out = nullptr;
    uint32_t nslots = cfg && cfg->slots ? cfg->slots : 4;
    if (nslots > 16) { set_error("rt_server_create: at most 16 slots"); return -1; }
    CKR(hipSetDevice(ctx->device));
    size_t free_b = 0, total_b = 0;
    CKR(hipMemGetInfo(&free_b, &total_b));
    std::unique_ptr<rt_server> s(new rt_server());
    s->ctx = ctx;
    s->device = ctx->device;
    s->max_batch = cfg && cfg->max_batch_suns ? cfg->max_batch_suns : 64;
    s->budget = cfg && cfg->memory_budget ? (size_t)cfg->memory_budget : free_b / 2;
    for (uint32_t w = 0; w < nslots; w++) {
        std::unique_ptr<Slot> sl(new Slot());
        sl->sx.ncu = ctx->sx.ncu;
        for (int k = 0; k < 3; k++) sl->sx.cfg[k] = ctx->sx.cfg[k];
        if (hipStreamCreateWithFlags(&sl->sx.stream, hipStreamNonBlocking) != hipSuccess) {
            for (auto &o : s->slots) (void)hipStreamDestroy(o->sx.stream);
            set_error("rt_server_create: hipStreamCreate failed");
            return -1;
        }
        s->slots.push_back(std::move(sl));
    }
    for (uint32_t w = 0; w < nslots; w++) s->workers.emplace_back([p = s.get(), w] { p->worker(w); });
    *out = s.release();
    return 0;
}

int rt_server_destroy(rt_server *srv) {
    if (!srv) return 0;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &t : srv->workers) t.join();
    (void)hipSetDevice(srv->device);
    for (auto &sl : srv->slots) {
        release_slot(*srv, *sl);
        (void)hipStreamDestroy(sl->sx.stream);
    }
    for (auto &m : srv->models)
        if (m->dmem) (void)hipFree(m->dmem);
    delete srv;
    return 0;
}

int rt_server_add_model(rt_server *srv, const rt_scene *scene, const rt_band_request *q, uint32_t *model_id) {
    NN(srv, "server");
    NN(scene, "scene");
    NN(q, "request");
    NN(model_id, "model_id");
    const uint32_t P = scene->s.h.P, B = q->nbands;
    if (B == 0 || B > 256) { set_error("rt_server_add_model: nbands must be 1..256"); return -1; }
    NN(q->band_depth, "band_depth");
    NN(q->emission, "emission");
    NN(q->rho, "rho");
    NN(q->tau, "tau");
    NN(q->diffuse_seed, "diffuse_seed");
    std::unique_ptr<Model> m(new Model());
    m->S = &scene->s;
    m->P = P;
    m->B = B;
    m->Sn = q->nsources;
    m->depth.assign(q->band_depth, q->band_depth + B);
    m->emission.assign(q->emission, q->emission + B);
    uint32_t maxdepth = 0;
    for (uint32_t d : m->depth) maxdepth = std::max(maxdepth, d);
    m->dseed.assign(q->diffuse_seed, q->diffuse_seed + maxdepth + 1);
    m->q = *q;
    m->q.band_depth = m->depth.data();
    m->q.emission = m->emission.data();
    m->q.diffuse_seed = m->dseed.data();
    m->q.source_dir = nullptr;
    m->q.source_flux = nullptr;
    m->q.diffuse_flux = nullptr;
    {
        std::vector<SourceShape> shapes;  // validated here so that a bad shape fails at registration
        if (source_shapes(q->nsources, q->source_type, q->source_width, q->source_rotation, shapes,
                          "rt_server_add_model"))
            return -1;
        m->q.source_type = nullptr;
        m->q.source_width = nullptr;
        m->q.source_rotation = nullptr;
        if (q->source_type) {
            m->stype.assign(q->source_type, q->source_type + q->nsources);
            m->q.source_type = m->stype.data();
            if (q->source_width) {
                m->swidth.assign(q->source_width, q->source_width + (size_t)2 * q->nsources);
                m->q.source_width = m->swidth.data();
            }
            if (q->source_rotation) {
                m->srot.assign(q->source_rotation, q->source_rotation + (size_t)3 * q->nsources);
                m->q.source_rotation = m->srot.data();
            }
        }
    }
    // anisotropic sky: copied with the model (every request of the model shares its peak directions)
    {
        const uint32_t Bq = q->nbands;
        auto keep = [](std::vector<float> &v, const float *src, size_t n) -> const float * {
            if (!src) return nullptr;
            v.assign(src, src + n);
            return v.data();
        };
        m->q.diffuse_extinction = keep(m->kext, q->diffuse_extinction, Bq);
        m->q.diffuse_peak_dir = keep(m->kpeak, q->diffuse_peak_dir, (size_t)3 * Bq);
        m->q.diffuse_dist_norm = keep(m->knorm, q->diffuse_dist_norm, Bq);
        m->q.sky_params = keep(m->kprague, q->sky_params, (size_t)4 * Bq);
    }
    m->q.rho = m->q.tau = m->q.eps = m->q.temperature = nullptr;
    // translucent covers: kept with the model, uploaded by each solve with its other per-solve tables
    m->q.is_glass = nullptr;
    m->q.glass_n = m->q.glass_KL = nullptr;
    if (q->is_glass && P) {
        NN(q->glass_n, "glass_n");
        const size_t PBg = (size_t)P * B;
        m->gis.assign(q->is_glass, q->is_glass + PBg);
        m->gn.assign(q->glass_n, q->glass_n + PBg);
        if (q->glass_KL) m->gkl.assign(q->glass_KL, q->glass_KL + PBg);
        else m->gkl.assign(PBg, 0.0f);
        m->q.is_glass = m->gis.data();
        m->q.glass_n = m->gn.data();
        m->q.glass_KL = m->gkl.data();
        for (uint32_t i = 0; i < P && !m->covers; i++) {
            bool all = true;
            for (uint32_t b = 0; b < B && all; b++) all = m->gis[(size_t)i * B + b] != 0;
            m->covers = all;
        }
    }
    // resident optical properties
    const size_t PB = (size_t)P * B;
    size_t bytes = PB * 2 * 4 + (q->eps ? PB * 4 : 0) + (q->temperature ? (size_t)P * 4 : 0);
    CKR(hipSetDevice(srv->device));
    CKR(hipMalloc(&m->dmem, bytes));
    float *p = (float *)m->dmem;
    m->dm.rho = p;
    m->dm.tau = p + PB;
    float *next = p + 2 * PB;
    m->dm.eps = q->eps ? next : nullptr;
    if (q->eps) next += PB;
    m->dm.temp = q->temperature ? next : nullptr;
    int rc = 0;
    if (hipMemcpy(p, q->rho, PB * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p + PB, q->tau, PB * 4, hipMemcpyHostToDevice) != hipSuccess ||
        (q->eps && hipMemcpy((void *)m->dm.eps, q->eps, PB * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (q->temperature && hipMemcpy((void *)m->dm.temp, q->temperature, (size_t)P * 4, hipMemcpyHostToDevice) != hipSuccess))
        rc = -1;
    if (rc) {
        (void)hipFree(m->dmem);
        set_error("rt_server_add_model: upload of optical properties failed");
        return -1;
    }
    m->resident_bytes = bytes;
    std::lock_guard<std::mutex> lk(srv->mu);
    *model_id = (uint32_t)srv->models.size();
    srv->models.push_back(std::move(m));
    return 0;
}

int rt_server_submit(rt_server *srv, uint32_t model_id, const rt_solve_sun *sun, float *R_out, uint64_t *ticket) {
    NN(srv, "server");
    NN(sun, "sun");
    NN(R_out, "R_out");
    NN(ticket, "ticket");
    std::unique_ptr<Request> r(new Request());
    const Model *mp = nullptr;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (model_id >= srv->models.size()) {
            set_error("rt_server_submit: unknown model " + std::to_string(model_id));
            return -1;
        }
        if (srv->stop) { set_error("rt_server_submit: server is shutting down"); return -1; }
        mp = srv->models[model_id].get();  // models are never removed while the server lives
    }
    const Model &m = *mp;
    if (m.Sn) {
        NN(sun->source_dir, "sun.source_dir");
        NN(sun->source_flux, "sun.source_flux");
    }
    NN(sun->diffuse_flux, "sun.diffuse_flux");
    r->model = model_id;
    r->dir.assign(sun->source_dir, sun->source_dir + (size_t)m.Sn * 3);
    r->flux.assign(sun->source_flux, sun->source_flux + (size_t)m.Sn * m.B);
    r->dflux.assign(sun->diffuse_flux, sun->diffuse_flux + m.B);
    r->seed = sun->direct_seed;
    r->out = R_out;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        r->id = srv->next_id++;
        *ticket = r->id;
        srv->queue.push_back(r.get());
        srv->submitted++;
        srv->reqs.emplace(r->id, std::move(r));
    }
    srv->cv_work.notify_one();
    return 0;
}

int rt_server_wait(rt_server *srv, uint64_t ticket, double *latency_ms) {
    NN(srv, "server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) {
        set_error("rt_server_wait: unknown or already collected ticket " + std::to_string(ticket));
        return -1;
    }
    Request *r = it->second.get();
    srv->cv_done.wait(lk, [&] { return r->done; });
    if (latency_ms) *latency_ms = ms_between(r->t_submit, r->t_done);
    int status = r->status;
    std::string err = r->err;
    srv->reqs.erase(it);
    if (status != 0) {
        set_error(err);
        return -1;
    }
    return 0;
}

int rt_server_get_stats(rt_server *srv, rt_server_stats *out) {
    NN(srv, "server");
    NN(out, "out");
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = rt_server_stats{};
    out->submitted = srv->submitted;
    out->completed = srv->completed;
    out->failed = srv->failed;
    out->batches = srv->batches;
    out->batched_requests = srv->batched_requests;
    out->max_batch = srv->max_batch_seen;
    out->queued = (uint32_t)srv->queue.size();
    out->memory_budget = srv->budget;
    out->peak_workspace_bytes = srv->peak_allocated;
    size_t resident = 0;
    for (auto &m : srv->models) resident += m->resident_bytes;
    out->resident_model_bytes = resident;
    if (!srv->lat_ms.empty()) {
        std::vector<double> v = srv->lat_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            double x = p / 100.0 * (double)(v.size() - 1);
            size_t i = (size_t)std::floor(x);
            size_t j = std::min(i + 1, v.size() - 1);
            return v[i] + (v[j] - v[i]) * (x - (double)i);
        };
        out->p50_ms = pct(50);
        out->p99_ms = pct(99);
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_ms = sum / (double)v.size();
    }
    return 0;
}

int rt_solve_footprint(uint32_t nprims, const rt_band_request *req, uint32_t nsuns, uint64_t *bytes) {
    NN(req, "request");
    NN(bytes, "bytes");
    NN(req->band_depth, "band_depth");
    if (nsuns == 0 || req->nbands == 0) { set_error("rt_solve_footprint: no suns or no bands"); return -1; }
    if (check_work_items((uint64_t)nsuns * req->nsources * nprims,
                         "rt_solve_footprint: direct pass (suns x sources x primitives)"))
        return -1;
    *bytes = Workspace::capacity_for(solve_workspace_bytes(nprims, *req, nsuns, true)) +
             Workspace::capacity_for(GridSet::budget(nprims, nsuns * req->nsources)) + (3u << 20) +
             (uint64_t)nsuns * nprims * req->nbands * sizeof(float);
    return 0;
}

int rt_server_reset_stats(rt_server *srv) {
    NN(srv, "server");
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->lat_ms.clear();
    srv->submitted = srv->completed = srv->failed = srv->batches = srv->batched_requests = 0;
    srv->max_batch_seen = 0;
    srv->peak_allocated = srv->allocated;
    return 0;
}

}  // extern "C"
