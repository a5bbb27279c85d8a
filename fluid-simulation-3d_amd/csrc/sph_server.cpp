// sph_server.cpp -- continuous batching of bake requests on one GPU.
//
// No reference counterpart: the reference runs one scene per process
// (physicsWorld.cc:32-37).  Each request is a scene (settings, particle
// count, table size, initial state) to advance by a number of steps.
// Requests wait in FIFO order and are admitted while a slot is free and their
// device footprint fits the HBM budget; every tick advances all running
// requests together -- each kernel of an Update is one launch over all of
// them (sph::Engine with many scenes) -- and finished requests leave with
// their results copied to host memory.  Membership changes repack the device
// layout in place, so the captured graphs of the engine stay valid.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/sph_amd.h"
#include "sph_engine.h"

namespace sph {
int set_last_error(int code, const std::string &msg);  // sph_capi.cpp
}

namespace {

// Device bytes per particle of an engine: seven float4 arrays (pos, vel, ps,
// xps, vs, vs2, dens), thirteen u32 words (bucket key and ticket, the (slot,
// index) pair of the bucket order, resident and sorted ids, two work-order
// key/value pairs, rank, scene map, row count), the 64-entry neighbour row
// and two float4 staging arrays.
constexpr uint64_t kBytesPerParticle = 7 * 16 + 13 * 4 + 64 * 4 + 2 * 16;
constexpr uint64_t kBytesPerBucket = 8;  // bucket bounds + counts

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Internal state of a request whose record must outlive it until the next
// repack (its id is still listed as a scene of the device layout).
constexpr int32_t kReqGone = 100;

struct Request {
    uint64_t id = 0;
    sph_params params{};
    uint32_t n = 0, table = 0;
    std::vector<float> pos, vel, dens;  // input until admitted, results once done
    bool generated = false;              // initial state generated on the device at admission
    sph::GenPlan gen{};
    uint64_t updates_total = 0, updates_done = 0;
    int32_t state = SPH_REQ_QUEUED;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t bytes = 0;
};

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q / 100.0 * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

struct sph_server_t {
    sph_server_config cfg{};
    sph::Engine *eng = nullptr;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> queue;
    std::vector<uint64_t> running;  // in engine scene order (after the last repack)
    std::vector<uint64_t> leaving;  // finished or cancelled since the last repack
    bool dirty = false;             // layout must be repacked before the next step
    uint64_t next_id = 1;
    uint64_t used_bytes = 0;
    // statistics
    std::vector<double> tick_ms;        // GPU time of each tick with work
    std::vector<uint64_t> tick_particles;  // particles it advanced
    std::vector<uint32_t> tick_layout;     // layout_no it ran on
    uint32_t layout_no = 0;                // repacks so far (membership changes)
    std::vector<double> request_ms;     // submit -> done of finished requests
    std::vector<double> wait_ms;        // submit -> admit
    uint64_t particle_updates = 0;      // particles x updates advanced
    double busy_s = 0;                  // wall time spent in ticks with work
    uint64_t finished = 0, admitted = 0, rejected = 0, cancelled = 0, admission_failures = 0;
    // GPU timing of ticks: event pairs in a ring; at most kInFlight ticks are
    // enqueued ahead of the GPU, so host-side admission overlaps device work
    static constexpr int kRing = 8;
    static constexpr int kInFlight = 2;
    hipEvent_t ev_start[kRing] = {}, ev_end[kRing] = {};
    std::deque<int> pending;  // ring slots of enqueued, not yet harvested ticks
    uint64_t slot_particles[kRing] = {};
    uint32_t slot_layout[kRing] = {};
    uint64_t tick_no = 0;
};

namespace {

int fail(int code, const std::string &msg) { return sph::set_last_error(code, msg); }

// A completed tick's GPU time joins the statistics and the tick log.
void record_tick(sph_server_t *s, int slot, float ms)
{
    s->tick_ms.push_back(ms);
    s->tick_particles.push_back(s->slot_particles[slot]);
    s->tick_layout.push_back(s->slot_layout[slot]);
    s->busy_s += ms * 1e-3;
}

uint64_t request_bytes(uint32_t n, uint32_t table)
{
    return kBytesPerParticle * n + kBytesPerBucket * table + sizeof(SceneDesc);
}

// Engine sizing: particle slots for the budget, but start at the configured
// capacity and grow (doubling) when admissions need it.
int ensure_capacity(sph_server_t *s, uint64_t need_n, uint64_t need_t, std::string *err)
{
    sph::Engine *e = s->eng;
    if (e && need_n <= e->capacity() && need_t <= e->table_capacity()) return SPH_OK;
    uint64_t cap = e ? e->capacity() : std::max<uint64_t>(s->cfg.initial_capacity, 1u << 16);
    uint64_t tcap = e ? e->table_capacity() : 2 * cap;
    while (cap < need_n) cap *= 2;
    while (tcap < need_t) tcap *= 2;
    const uint64_t cap_max = s->cfg.max_capacity ? s->cfg.max_capacity : (1ull << 31);
    cap = std::min<uint64_t>(cap, cap_max);
    tcap = std::min<uint64_t>(tcap, 1ull << 31);
    if (cap < need_n || tcap < need_t) {
        *err = "growing the device layout beyond its particle or bucket limit (max_capacity " +
               std::to_string(cap_max) + ")";
        return SPH_ERR_OOM;
    }
    sph::Engine *ne = nullptr;
    int rc = sph::Engine::create({}, (uint32_t)cap, (uint32_t)tcap, s->cfg.max_running, &ne, err,
                                 true, true);
    if (rc) return rc;
    if (e) {
        // carry the running scenes over through host memory (rare: growth only)
        std::vector<sph::RepackItem> items;
        std::vector<std::vector<float>> pos(e->n_scenes()), vel(e->n_scenes());
        for (uint32_t q = 0; q < e->n_scenes(); ++q) {
            const uint32_t n = e->scene(q).n;
            pos[q].resize(4 * (size_t)n);
            vel[q].resize(4 * (size_t)n);
            rc = e->download(q, pos[q].data(), vel[q].data(), nullptr, err);
            if (rc) {
                delete ne;
                return rc;
            }
            items.push_back({{e->params(q), n, e->scene(q).table}, -1, pos[q].data(), vel[q].data()});
        }
        rc = ne->repack(items, err);
        if (rc) {
            delete ne;
            return rc;
        }
        ne->set_launch_cfg(e->launch_cfg());
        delete e;
    }
    s->eng = ne;
    return SPH_OK;
}

// Applies departures and admissions to the device layout.
int repack(sph_server_t *s, const std::vector<uint64_t> &admit, std::string *err)
{
    sph::Engine *e = s->eng;
    std::vector<sph::RepackItem> items;
    std::vector<uint64_t> order;
    uint64_t need_n = 0, need_t = 0;
    for (size_t q = 0; q < s->running.size(); ++q) {
        const Request &r = s->reqs.at(s->running[q]);
        if (r.state != SPH_REQ_RUNNING) continue;
        items.push_back({{r.params, r.n, r.table}, (int)q, nullptr, nullptr});
        order.push_back(r.id);
        need_n += r.n;
        need_t += r.table;
    }
    for (uint64_t id : admit) {
        Request &r = s->reqs.at(id);
        items.push_back({{r.params, r.n, r.table}, -1, r.pos.data(), r.vel.data(),
                         r.generated ? &r.gen : nullptr});
        order.push_back(id);
        need_n += r.n;
        need_t += r.table;
    }
    // (growth carries every scene of the old engine over in the same order,
    // so old_scene indices stay valid)
    int rc = ensure_capacity(s, need_n, need_t, err);
    if (rc) return rc;
    (void)e;
    rc = s->eng->repack(items, err);
    if (rc) return rc;
    // records that only the old layout still referenced
    for (uint64_t id : s->running) {
        const int32_t st = s->reqs.at(id).state;
        if (st == SPH_REQ_CANCELLED || st == kReqGone) s->reqs.erase(id);
    }
    s->running = order;
    s->leaving.clear();
    for (uint64_t id : admit) {
        Request &r = s->reqs.at(id);
        r.pos.clear();
        r.pos.shrink_to_fit();
        r.vel.clear();
        r.vel.shrink_to_fit();
    }
    s->dirty = false;
    ++s->layout_no;
    return SPH_OK;
}

}  // namespace

#define SRV_GUARD_BEGIN try {
#define SRV_GUARD_END                                                                      \
    }                                                                                      \
    catch (const std::bad_alloc &) { return fail(SPH_ERR_OOM, "host allocation failed"); } \
    catch (const std::exception &ex) { return fail(SPH_ERR_INVALID, ex.what()); }          \
    catch (...) { return fail(SPH_ERR_INVALID, "unknown exception"); }

extern "C" {

void sph_server_default_config(sph_server_config *c)
{
    if (!c) return;
    c->hbm_budget_bytes = 200ull << 30;
    c->max_running = 64;
    c->initial_capacity = 1u << 23;
    c->dt = 1.0f / 120.0f;
    c->substeps = 3;
    c->updates_per_tick = 1;
    c->max_capacity = 0;
}

uint64_t sph_server_request_bytes(const sph_params *p, uint32_t n, uint32_t table_size)
{
    if (!p || sph::check_params(*p)) return 0;
    return request_bytes(n, table_size ? table_size : sph::grid_table_size(*p, n));
}

int sph_server_create(const sph_server_config *cfg, sph_server *out)
{
    SRV_GUARD_BEGIN
    if (!cfg || !out) return fail(SPH_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->max_running == 0 || cfg->substeps < 1 || cfg->updates_per_tick < 1 ||
        !(cfg->dt > 0.0f) || cfg->hbm_budget_bytes == 0)
        return fail(SPH_ERR_INVALID, "bad server configuration");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(SPH_ERR_HIP, "no HIP device visible");
    sph_server_t *s = new sph_server_t();
    s->cfg = *cfg;
    for (int i = 0; i < sph_server_t::kRing; ++i)
        // tick timers and in-flight throttling only: no system-scope fence
        if (hipEventCreateWithFlags(&s->ev_start[i], hipEventDisableSystemFence) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_end[i], hipEventDisableSystemFence) != hipSuccess) {
            sph_server_destroy(s);
            return fail(SPH_ERR_HIP, "event creation failed");
        }
    std::string err;
    int rc = ensure_capacity(s, 1, 1, &err);
    if (rc) {
        delete s;
        return fail(rc, err);
    }
    *out = s;
    return SPH_OK;
    SRV_GUARD_END
}

int sph_server_destroy(sph_server s)
{
    SRV_GUARD_BEGIN
    if (!s) return SPH_OK;
    if (s->eng) (void)hipStreamSynchronize(s->eng->stream());
    for (int i = 0; i < sph_server_t::kRing; ++i) {
        if (s->ev_start[i]) (void)hipEventDestroy(s->ev_start[i]);
        if (s->ev_end[i]) (void)hipEventDestroy(s->ev_end[i]);
    }
    delete s->eng;
    delete s;
    return SPH_OK;
    SRV_GUARD_END
}

int sph_server_submit(sph_server s, const sph_params *p, uint32_t n, uint32_t table_size,
                      const float *pos4, const float *vel4, uint32_t steps, uint64_t *id)
{
    SRV_GUARD_BEGIN
    if (!s || !p || !pos4 || !vel4 || !id) return fail(SPH_ERR_INVALID, "null argument");
    if (n == 0 || steps == 0) return fail(SPH_ERR_INVALID, "n and steps must be > 0");
    if (const char *m = sph::check_params(*p)) return fail(SPH_ERR_INVALID, m);
    Request r;
    r.id = s->next_id++;
    r.params = *p;
    r.n = n;
    r.table = table_size ? table_size : sph::grid_table_size(*p, n);
    r.bytes = request_bytes(r.n, r.table);
    if (r.bytes > s->cfg.hbm_budget_bytes) {
        ++s->rejected;
        return fail(SPH_ERR_BUDGET, "request exceeds the HBM budget on its own");
    }
    r.pos.assign(pos4, pos4 + 4 * (size_t)n);
    r.vel.assign(vel4, vel4 + 4 * (size_t)n);
    r.updates_total = (uint64_t)steps * (uint64_t)s->cfg.substeps;
    r.t_submit = now_s();
    *id = r.id;
    s->queue.push_back(r.id);
    s->reqs.emplace(r.id, std::move(r));
    return SPH_OK;
    SRV_GUARD_END
}

int sph_server_submit_generated(sph_server s, const sph_params *p, uint32_t n,
                                uint32_t table_size, int kind, uint64_t seed, const float box[3],
                                float spacing, float jitter, float speed, uint32_t steps,
                                uint64_t *id)
{
    SRV_GUARD_BEGIN
    if (!s || !p || !id) return fail(SPH_ERR_INVALID, "null argument");
    if (n == 0 || steps == 0) return fail(SPH_ERR_INVALID, "n and steps must be > 0");
    if (const char *m = sph::check_params(*p)) return fail(SPH_ERR_INVALID, m);
    Request r;
    if (sph::plan_generate(kind, seed, n, box, spacing, jitter, speed, &r.gen))
        return fail(SPH_ERR_INVALID, "invalid scene generator arguments");
    r.generated = true;
    r.id = s->next_id++;
    r.params = *p;
    r.n = n;
    r.table = table_size ? table_size : sph::grid_table_size(*p, n);
    r.bytes = request_bytes(r.n, r.table);
    if (r.bytes > s->cfg.hbm_budget_bytes) {
        ++s->rejected;
        return fail(SPH_ERR_BUDGET, "request exceeds the HBM budget on its own");
    }
    r.updates_total = (uint64_t)steps * (uint64_t)s->cfg.substeps;
    r.t_submit = now_s();
    *id = r.id;
    s->queue.push_back(r.id);
    s->reqs.emplace(r.id, std::move(r));
    return SPH_OK;
    SRV_GUARD_END
}

int sph_server_tick(sph_server s, sph_server_tick_info *info)
{
    SRV_GUARD_BEGIN
    if (!s) return fail(SPH_ERR_INVALID, "null server");
    const double t0 = now_s();
    std::string err;
    // 1. admit in arrival order while a slot is free and the budget holds;
    // the requests change state only once the device layout holds them (a
    // failed repack or growth leaves them queued, the budget uncharged)
    std::vector<uint64_t> admit;
    size_t slots = 0;
    uint64_t charge = 0;
    for (uint64_t id : s->running)
        if (s->reqs.at(id).state == SPH_REQ_RUNNING) ++slots;
    for (size_t q = 0; q < s->queue.size() && slots + admit.size() < s->cfg.max_running; ++q) {
        const Request &r = s->reqs.at(s->queue[q]);
        if (s->used_bytes + charge + r.bytes > s->cfg.hbm_budget_bytes) break;  // head of line waits
        charge += r.bytes;
        admit.push_back(r.id);
    }
    if (!admit.empty() || s->dirty) {
        int rc = repack(s, admit, &err);
        if (rc == SPH_ERR_OOM && !admit.empty()) {
            // the layout cannot grow to hold them (device memory, or
            // max_capacity): they wait at the head of the queue, as for the
            // budget, while the running requests go on
            ++s->admission_failures;
            admit.clear();
            charge = 0;
            bool any_running = false;
            for (uint64_t id : s->running)
                any_running = any_running || s->reqs.at(id).state == SPH_REQ_RUNNING;
            rc = s->dirty ? repack(s, admit, &err) : SPH_OK;  // departures still apply
            if (!rc && !any_running)
                return fail(SPH_ERR_OOM, "the request at the head of the queue cannot be admitted "
                                         "even with nothing else running: " + err);
        }
        if (rc) return fail(rc, err);
    }
    const double t_admit = now_s();
    for (uint64_t id : admit) {
        Request &r = s->reqs.at(id);
        r.state = SPH_REQ_RUNNING;
        r.t_admit = t_admit;
        s->wait_ms.push_back((r.t_admit - r.t_submit) * 1e3);
        s->queue.pop_front();
        ++s->admitted;
    }
    s->used_bytes += charge;
    // 2. advance every running request
    uint32_t finished = 0;
    uint64_t particles = 0;
    const float sub = s->cfg.dt / (float)s->cfg.substeps;
    // harvest the GPU durations of completed ticks
    auto harvest = [&](bool wait) -> int {
        while (!s->pending.empty()) {
            const int slot = s->pending.front();
            hipError_t q = wait ? hipEventSynchronize(s->ev_end[slot]) : hipEventQuery(s->ev_end[slot]);
            if (q == hipErrorNotReady) return SPH_OK;
            if (q != hipSuccess) return fail(SPH_ERR_HIP, "tick event");
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, s->ev_start[slot], s->ev_end[slot]) != hipSuccess)
                return fail(SPH_ERR_HIP, "tick timer");
            record_tick(s, slot, ms);
            s->pending.pop_front();
            wait = false;
        }
        return SPH_OK;
    };
    if (!s->running.empty()) {
        // keep at most kInFlight ticks ahead of the device
        while ((int)s->pending.size() >= sph_server_t::kInFlight) {
            const int slot = s->pending.front();
            if (hipEventSynchronize(s->ev_end[slot]) != hipSuccess) return fail(SPH_ERR_HIP, "tick wait");
            int hr = harvest(false);
            if (hr) return hr;
        }
        const int slot = (int)(s->tick_no++ % sph_server_t::kRing);
        hipStream_t st = s->eng->stream();
        if (hipEventRecord(s->ev_start[slot], st) != hipSuccess) return fail(SPH_ERR_HIP, "tick event");
        for (int u = 0; u < s->cfg.updates_per_tick; ++u) {
            int rc = s->eng->update(sub, &err);
            if (rc) return fail(rc, err);
        }
        if (hipEventRecord(s->ev_end[slot], st) != hipSuccess) return fail(SPH_ERR_HIP, "tick event");
        s->pending.push_back(slot);
        s->slot_layout[slot] = s->layout_no;
        s->slot_particles[slot] = 0;  // set below, once the running requests are counted
        int rc = SPH_OK;
        // 3. retire finished requests: results to host, slot freed at the next repack
        for (size_t q = 0; q < s->running.size(); ++q) {
            Request &r = s->reqs.at(s->running[q]);
            if (r.state != SPH_REQ_RUNNING) continue;
            particles += r.n;
            const uint64_t adv = std::min<uint64_t>(s->cfg.updates_per_tick,
                                                    r.updates_total - r.updates_done);
            r.updates_done += adv;
            s->particle_updates += (uint64_t)r.n * adv;
            if (r.updates_done >= r.updates_total) {
                r.pos.resize(4 * (size_t)r.n);
                r.vel.resize(4 * (size_t)r.n);
                r.dens.resize(2 * (size_t)r.n);
                // (download waits for the enqueued ticks)
                rc = s->eng->download((uint32_t)q, r.pos.data(), r.vel.data(), r.dens.data(), &err);
                if (rc) return fail(rc, err);
                r.state = SPH_REQ_DONE;
                r.t_done = now_s();
                s->request_ms.push_back((r.t_done - r.t_submit) * 1e3);
                s->used_bytes -= r.bytes;
                s->leaving.push_back(r.id);
                s->dirty = true;
                ++finished;
                ++s->finished;
            }
        }
        s->slot_particles[slot] = particles;
    }
    {
        int hr = harvest(false);
        if (hr) return hr;
    }
    if (info) {
        uint32_t run = 0;
        for (uint64_t id : s->running)
            if (s->reqs.at(id).state == SPH_REQ_RUNNING) ++run;
        info->running = run;
        info->queued = (uint32_t)s->queue.size();
        info->admitted = (uint32_t)admit.size();
        info->finished = finished;
        info->particles = particles;
        info->ms = (now_s() - t0) * 1e3;
    }
    return SPH_OK;
    SRV_GUARD_END
}

int sph_server_status(sph_server s, uint64_t id, int32_t *state, uint64_t *updates_done,
                      double *latency_ms)
{
    SRV_GUARD_BEGIN
    if (!s) return fail(SPH_ERR_INVALID, "null server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second.state == kReqGone)
        return fail(SPH_ERR_NOT_FOUND, "unknown request id");
    const Request &r = it->second;
    if (state) *state = r.state;
    if (updates_done) *updates_done = r.updates_done;
    if (latency_ms) *latency_ms = r.state == SPH_REQ_DONE ? (r.t_done - r.t_submit) * 1e3 : -1.0;
    return SPH_OK;
    SRV_GUARD_END
}

int sph_server_fetch(sph_server s, uint64_t id, uint32_t n, float *pos4, float *vel4, float *dens2)
{
    SRV_GUARD_BEGIN
    if (!s) return fail(SPH_ERR_INVALID, "null server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second.state == kReqGone)
        return fail(SPH_ERR_NOT_FOUND, "unknown request id");
    Request &r = it->second;
    if (r.state != SPH_REQ_DONE) return fail(SPH_ERR_STATE, "request not finished");
    if (n != r.n)
        return fail(SPH_ERR_INVALID, "n must equal the request's particle count (" +
                                         std::to_string(r.n) + ")");
    if (pos4) memcpy(pos4, r.pos.data(), sizeof(float) * r.pos.size());
    if (vel4) memcpy(vel4, r.vel.data(), sizeof(float) * r.vel.size());
    if (dens2) memcpy(dens2, r.dens.data(), sizeof(float) * r.dens.size());
    if (std::find(s->running.begin(), s->running.end(), id) != s->running.end()) {
        // still a scene of the device layout until the next repack
        r.state = kReqGone;
        std::vector<float>().swap(r.pos);
        std::vector<float>().swap(r.vel);
        std::vector<float>().swap(r.dens);
    } else {
        s->reqs.erase(it);
    }
    return SPH_OK;
    SRV_GUARD_END
}

int sph_server_cancel(sph_server s, uint64_t id)
{
    SRV_GUARD_BEGIN
    if (!s) return fail(SPH_ERR_INVALID, "null server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second.state == kReqGone)
        return fail(SPH_ERR_NOT_FOUND, "unknown request id");
    Request &r = it->second;
    if (r.state == SPH_REQ_DONE) return fail(SPH_ERR_STATE, "request already finished");
    if (r.state == SPH_REQ_CANCELLED) return fail(SPH_ERR_STATE, "request already cancelled");
    if (r.state == SPH_REQ_RUNNING) {
        s->used_bytes -= r.bytes;
        s->leaving.push_back(r.id);
        s->dirty = true;
    }
    ++s->cancelled;
    if (r.state == SPH_REQ_QUEUED) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), id));
        s->reqs.erase(it);
        return SPH_OK;
    }
    // a running request leaves the device layout at the next repack
    r.state = SPH_REQ_CANCELLED;
    return SPH_OK;
    SRV_GUARD_END
}

int sph_server_tick_log(sph_server s, uint64_t first, uint64_t cap, float *ms, uint64_t *particles,
                        uint32_t *layout, uint64_t *count)
{
    SRV_GUARD_BEGIN
    if (!s || !count) return fail(SPH_ERR_INVALID, "null argument");
    const uint64_t have = s->tick_ms.size();
    const uint64_t m = first < have ? std::min<uint64_t>(cap, have - first) : 0;
    for (uint64_t i = 0; i < m; ++i) {
        if (ms) ms[i] = (float)s->tick_ms[first + i];
        if (particles) particles[i] = s->tick_particles[first + i];
        if (layout) layout[i] = s->tick_layout[first + i];
    }
    *count = m;
    return SPH_OK;
    SRV_GUARD_END
}

int sph_server_get_stats(sph_server s, sph_server_stats *st)
{
    SRV_GUARD_BEGIN
    if (!s || !st) return fail(SPH_ERR_INVALID, "null argument");
    // finish timing the enqueued ticks
    while (!s->pending.empty()) {
        const int slot = s->pending.front();
        float ms = 0.0f;
        if (hipEventSynchronize(s->ev_end[slot]) != hipSuccess ||
            hipEventElapsedTime(&ms, s->ev_start[slot], s->ev_end[slot]) != hipSuccess)
            return fail(SPH_ERR_HIP, "tick timer");
        record_tick(s, slot, ms);
        s->pending.pop_front();
    }
    memset(st, 0, sizeof(*st));
    st->admitted = s->admitted;
    st->finished = s->finished;
    st->rejected = s->rejected;
    st->cancelled = s->cancelled;
    st->queued = s->queue.size();
    uint32_t run = 0;
    for (uint64_t id : s->running)
        if (s->reqs.at(id).state == SPH_REQ_RUNNING) ++run;
    st->running = run;
    st->used_bytes = s->used_bytes;
    st->particle_updates = s->particle_updates;
    st->busy_s = s->busy_s;
    st->ticks = s->tick_ms.size();
    st->tick_ms_p50 = percentile(s->tick_ms, 50);
    st->tick_ms_p99 = percentile(s->tick_ms, 99);
    st->request_ms_p50 = percentile(s->request_ms, 50);
    st->request_ms_p99 = percentile(s->request_ms, 99);
    st->wait_ms_p50 = percentile(s->wait_ms, 50);
    st->wait_ms_p99 = percentile(s->wait_ms, 99);
    st->capacity = s->eng ? s->eng->capacity() : 0;
    st->admission_failures = s->admission_failures;
    return SPH_OK;
    SRV_GUARD_END
}

}  // extern "C"
