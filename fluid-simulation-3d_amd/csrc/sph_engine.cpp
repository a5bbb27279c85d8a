// sph_engine.cpp -- host side of the gfx950 SPH step (see sph_engine.h).
#include "sph_engine.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>

#include "sph_device.h"

namespace sph {

namespace {
// glm::pi<float>() (exts/glm/ext/scalar_constants.inl:13-17)
const float kPiF = (float)3.14159265358979323846264338327950288;

uint32_t bits_for(uint32_t max_value)
{
    uint32_t b = 1;
    while (b < 32 && (max_value >> b) != 0) ++b;
    return b;
}

bool finite_f(float x) { return isfinite(x) != 0; }
}  // namespace

float sqrt_threshold(float h)
{
    if (!(h > 0.0f)) return 0.0f;
    if (isinf(h)) return INFINITY;
    // sqrtf is correctly rounded, hence monotone: binary search the first
    // non-negative float whose root reaches h.
    uint32_t lo = 0, hi = 0x7f800000u;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        float x;
        memcpy(&x, &mid, 4);
        if (sqrtf(x) >= h) hi = mid;
        else lo = mid + 1;
    }
    float r;
    memcpy(&r, &lo, 4);
    return r;
}

const char *check_params(const sph_params &p)
{
    if (!finite_f(p.smoothing_radius) || !(p.smoothing_radius > 0.0f))
        return "smoothing_radius must be finite and > 0";
    if (!finite_f(p.target_density) || !finite_f(p.pressure_mult) ||
        !finite_f(p.near_pressure_mult) || !finite_f(p.viscosity) || !finite_f(p.gravity_scale))
        return "settings must be finite";
    for (int c = 0; c < 3; ++c)
        if (!finite_f(p.bounds[c]) || !(p.bounds[c] > 0.0f)) return "bounds must be finite and > 0";
    if (!finite_f(p.lookahead) || !finite_f(p.damp)) return "lookahead and damp must be finite";
    if (!finite_f(p.prefilter_r2) || p.prefilter_r2 < 0.0f) return "prefilter_r2 must be finite and >= 0";
    return nullptr;
}

namespace {
// Cell range [lo, hi] of the box [-half, half] along one axis, 4 cells of
// margin each side (predicted positions leave the box by |v| / 120), clamped
// far inside the int32 range.
void box_cells(float half, float h, int32_t *lo, int32_t *hi)
{
    const double lim = (double)(1 << 29);
    const double a = std::max(-lim, std::min(lim, floor((double)(-half / h))));
    const double b = std::max(-lim, std::min(lim, floor((double)(half / h))));
    *lo = (int32_t)a - 4;
    *hi = (int32_t)b + 4;
}

uint64_t box_cell_count(const sph_params &p)
{
    uint64_t c = 1;
    for (int a = 0; a < 3; ++a) {
        int32_t lo, hi;
        box_cells(p.bounds[a] * 0.5f, p.smoothing_radius, &lo, &hi);
        c *= (uint64_t)(hi - lo + 1);
        if (c > (1ull << 40)) c = 1ull << 40;
    }
    return c;
}
}  // namespace

uint32_t grid_table_size(const sph_params &p, uint32_t n)
{
    const uint64_t cap = std::min<uint64_t>(std::max<uint64_t>(8ull * n, 1ull << 20), 1ull << 28);
    return (uint32_t)std::max<uint64_t>(1, std::min(box_cell_count(p), cap));
}

// Periods of the cell grid within `table` buckets: the box (+margin) when it
// fits, else every axis shrunk by the same factor (cells then wrap; the
// float-hash test of the walk keeps results exact).
void grid_dims(const sph_params &p, uint32_t table, uint32_t g[3])
{
    uint64_t ideal[3], prod = 1;
    for (int a = 0; a < 3; ++a) {
        int32_t l, hi;
        box_cells(p.bounds[a] * 0.5f, p.smoothing_radius, &l, &hi);
        ideal[a] = (uint64_t)(hi - l + 1);
        prod *= ideal[a];
    }
    if (prod <= table) {
        for (int a = 0; a < 3; ++a) g[a] = (uint32_t)ideal[a];
        return;
    }
    const double f = cbrt((double)table / (double)prod);
    uint64_t q = 1;
    for (int a = 0; a < 3; ++a) {
        g[a] = (uint32_t)std::max<double>(1.0, floor((double)ideal[a] * f));
        q *= g[a];
    }
    while (q > table) {
        int big = 0;
        for (int a = 1; a < 3; ++a)
            if (g[a] > g[big]) big = a;
        q = q / g[big] * (g[big] - 1);
        --g[big];
    }
}

void fill_scene_desc(const sph_params &p, uint32_t n, uint32_t table, bool grid, SceneDesc *d)
{
    const float r = p.smoothing_radius;
    d->n = n;
    d->table = table;
    d->mod_m = fastmod_constant(table);
    d->h = r;
    d->r2 = p.prefilter_r2;
    d->h2_lt = sqrt_threshold(r);
    // one comparison for both tests of the walk (prefilter d2 <= r2, cc:357,
    // and the kernels' support d < h): valid for every non-NaN d2
    d->thr = fminf(d->h2_lt, nextafterf(p.prefilter_r2, INFINITY));
    d->rho0 = p.target_density;
    d->k = p.pressure_mult;
    d->k_near = p.near_pressure_mult;
    d->mu = p.viscosity;
    d->w2c = 15 / (2 * kPiF * powf(r, 5));               // kernels.h:29
    d->w3c = 15.0f / (kPiF * powf(r, 6));                // kernels.h:41
    d->dw2c = 15.0f / (powf(r, 5) * kPiF);               // kernels.h:53
    d->dw3c = 45 / (powf(r, 6) * kPiF);                  // kernels.h:65
    d->poly6c = 315 / (64 * kPiF * powf(fabsf(r), 9));   // kernels.h:77
    d->hh = r * r;                                       // kernels.h:78
    d->g[0] = 0.0f;                                      // cc:315-320
    d->g[1] = p.gravity_on ? -p.gravity_scale : 0.0f;
    d->g[2] = 0.0f;
    d->lookahead = p.lookahead;                          // cc:47
    for (int c = 0; c < 3; ++c) d->half[c] = p.bounds[c] * 0.5f;  // cc:88
    d->damp_neg = -1 * p.damp;                           // cc:94
    for (int c = 0; c < 3; ++c) {
        int32_t hi;
        box_cells(d->half[c], r, &d->cell_lo[c], &hi);
    }
    d->grid = grid ? 1u : 0u;
    if (grid) {
        grid_dims(p, table, d->gdim);
    } else {
        d->gdim[0] = d->gdim[1] = d->gdim[2] = 1;
    }
    for (int c = 0; c < 3; ++c) d->gmod[c] = fastmod_constant(d->gdim[c]);
}

void work_key_layout(uint32_t n_scenes, EngineDesc *e)
{
    if (n_scenes <= 1) {
        e->scene_shift = 32;
        e->morton_drop = 0;
        return;
    }
    const uint32_t sb = bits_for(n_scenes - 1);
    e->scene_shift = 32 - sb;
    e->morton_drop = e->scene_shift >= 30 ? 0 : 30 - e->scene_shift;
}

int Engine::check(hipError_t e, const char *what, std::string *err)
{
    if (e == hipSuccess) return SPH_OK;
    if (err) *err = std::string(what) + ": " + hipGetErrorString(e);
    return SPH_ERR_HIP;
}

#define CK(expr, what)                                   \
    do {                                                 \
        int rc_ = check((expr), (what), err);            \
        if (rc_) return rc_;                             \
    } while (0)

int Engine::create(const std::vector<SceneSpec> &scenes, uint32_t capacity,
                   uint32_t table_capacity, uint32_t scene_capacity, Engine **out,
                   std::string *err, bool multi, bool grid)
{
    if (scenes.empty() && (capacity == 0 || table_capacity == 0 || scene_capacity == 0)) {
        *err = "no scenes and no capacities";
        return SPH_ERR_INVALID;
    }
    uint64_t n_total = 0, t_total = 0;
    for (const SceneSpec &s : scenes) {
        if (s.n == 0 || s.table == 0) {
            *err = "scene needs n > 0 and table_size > 0";
            return SPH_ERR_INVALID;
        }
        if (const char *m = check_params(s.params)) {
            *err = m;
            return SPH_ERR_INVALID;
        }
        n_total += s.n;
        t_total += s.table;
    }
    if (n_total > (1ull << 31) || t_total > (1ull << 31)) {
        *err = "particle or bucket total exceeds 2^31";
        return SPH_ERR_INVALID;
    }
    capacity = std::max<uint32_t>(capacity, (uint32_t)n_total);
    table_capacity = std::max<uint32_t>(table_capacity, (uint32_t)t_total);

    Engine *e = new Engine();
    int rc = e->check(hipStreamCreateWithFlags(&e->stream_, hipStreamNonBlocking),
                      "hipStreamCreate", err);
    e->scene_cap_ = std::max<uint32_t>(scene_capacity, (uint32_t)scenes.size());
    if (!rc) rc = e->alloc(capacity, table_capacity, err);
    if (rc) {
        delete e;
        return rc;
    }
    uint32_t p_off = 0, k_off = 0;
    std::vector<uint32_t> scene_of(capacity, 0);
    for (uint32_t s = 0; s < scenes.size(); ++s) {
        SceneDesc d{};
        fill_scene_desc(scenes[s].params, scenes[s].n, scenes[s].table, grid, &d);
        d.p_off = p_off;
        d.key_off = k_off;
        d.index = s;
        d.dt = 0.0f;
        for (uint32_t i = 0; i < scenes[s].n; ++i) scene_of[p_off + i] = s;
        p_off += scenes[s].n;
        k_off += scenes[s].table;
        e->sd_h_.push_back(d);
        e->params_.push_back(scenes[s].params);
    }
    e->ed_h_->n_total = p_off;
    e->ed_h_->table_total = k_off;
    e->ed_h_->n_scenes = (uint32_t)scenes.size();
    {
        // FluidSimCPU's gradient stops Color1..Color4 (fluidSimCPU.h:26-29)
        const float stops[16] = {0.0f, 0.75f, 1.0f, 1.0f, 0.0f, 1.0f, 0.0f, 1.0f,
                                 1.0f, 1.0f, 0.0f, 1.0f, 1.0f, 0.0f, 0.0f, 1.0f};
        memcpy(e->ed_h_->stops, stops, sizeof(stops));
    }
    work_key_layout(e->ed_h_->n_scenes, e->ed_h_);
    e->ed_h_->scan_spin = kScanSpinDefault;
    e->invalid_.assign(scenes.size(), 0);
    e->d_.multi = multi || scenes.size() > 1;
    e->grid_ = grid;
    e->d_.grid = grid;
    e->cfg_ = default_launch_cfg(grid);
    {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) e->d_.fused_grid = fused_update_grid(e->d_.multi, dev);
    }
    rc = e->check(hipMemcpy(e->d_scene_of_, scene_of.data(), sizeof(uint32_t) * capacity,
                            hipMemcpyHostToDevice),
                  "upload scene map", err);
    if (!rc) rc = e->push_descs(err);
    if (!rc)
        rc = e->check(hipMemsetAsync(e->d_.pos, 0, sizeof(float4) * capacity, e->stream_),
                      "clear pos", err);
    if (!rc)
        rc = e->check(hipMemsetAsync(e->d_.vel, 0, sizeof(float4) * capacity, e->stream_),
                      "clear vel", err);
    if (!rc)
        rc = e->check(hipMemsetAsync(e->d_.dens, 0, sizeof(float4) * capacity, e->stream_),
                      "clear dens", err);
    // work order and rank hold valid indices (0) until the first rebuild
    for (uint32_t *b : {e->d_.wvals[0], e->d_.wvals[1], e->d_.rank})
        if (!rc)
            rc = e->check(hipMemsetAsync(b, 0, sizeof(uint32_t) * capacity, e->stream_),
                          "clear work order", err);
    e->d_.work = e->d_.wvals[0];
    // resident order: particle order
    if (!rc) rc = e->check(launch_iota(e->d_.ids, capacity, 0, e->stream_), "resident ids", err);
    if (!rc) rc = e->check(hipStreamSynchronize(e->stream_), "sync", err);
    // stage timers: timing only (no system-scope fence when recorded)
    for (int i = 0; i < 7 && !rc; ++i)
        rc = e->check(hipEventCreateWithFlags(&e->ev_[i], hipEventDisableSystemFence), "event", err);
    if (rc) {
        delete e;
        return rc;
    }
    *out = e;
    return SPH_OK;
}

int Engine::alloc(uint32_t cap, uint32_t tcap, std::string *err)
{
    auto dm = [&](void **p, size_t bytes, const char *what) {
        device_bytes_ += bytes;
        return check(hipMalloc(p, bytes), what, err);
    };
    const size_t f4 = sizeof(float4) * (size_t)cap, u4 = sizeof(uint32_t) * (size_t)cap;
    // radix blocks: at least 4 tiles of 256 keys each, at most 1024 blocks
    uint32_t nblk = (cap + 1023) / 1024;
    nblk = std::max<uint32_t>(1, std::min<uint32_t>(nblk, 1024));
    d_.radix_blocks = nblk;
    // counting-sort chunks
    const uint32_t cblk = (uint32_t)(((uint64_t)tcap + kCountChunkBuckets - 1) / kCountChunkBuckets);
    d_.count_blocks = cblk;
    d_.capacity = cap;
    d_.table_capacity = tcap;
    int rc = 0;
    if (!rc) rc = check(hipHostMalloc((void **)&ed_h_, sizeof(EngineDesc)), "pinned desc", err);
    if (!rc) rc = dm((void **)&d_ed_, sizeof(EngineDesc), "engine desc");
    if (!rc) rc = dm((void **)&d_sd_, sizeof(SceneDesc) * scene_cap_, "scene table");
    for (int q = 0; q < kDescRing && !rc; ++q) {
        rc = check(hipHostMalloc((void **)&desc_ring_[q], sizeof(EngineDesc) + sizeof(SceneDesc) * scene_cap_),
                   "pinned descriptor staging", err);
        if (!rc)
            rc = check(hipEventCreateWithFlags(&desc_ev_[q], hipEventDisableTiming), "descriptor event", err);
    }
    if (!rc) rc = dm((void **)&d_scene_of_, u4, "scene map");
    if (!rc) rc = dm((void **)&d_.pos, f4, "pos");
    if (!rc) rc = dm((void **)&d_.vel, f4, "vel");
    if (!rc) rc = dm((void **)&d_.ids, u4, "resident ids");
    if (!rc) rc = dm((void **)&d_.ps, f4, "ps");
    if (!rc) rc = dm((void **)&d_.xps, f4, "xps");
    if (!rc) rc = dm((void **)&d_.vs, f4, "vs");
    if (!rc) rc = dm((void **)&d_.vs2, f4, "vs2");
    if (!rc) rc = dm((void **)&d_.dens, f4, "dens");
    for (int b = 0; b < 2 && !rc; ++b) rc = dm((void **)&d_.keys[b], u4, "keys");
    if (!rc) rc = dm((void **)&d_.order, 2 * u4, "bucket order");
    if (!rc) rc = dm((void **)&d_.sids, u4, "sorted ids");
    if (!rc) rc = dm((void **)&d_.hist, sizeof(uint32_t) * 256 * ((size_t)nblk + 1), "radix hist");
    if (!rc) rc = dm((void **)&d_.bounds, sizeof(uint32_t) * ((size_t)tcap + 1), "bucket bounds");
    d_.count_size = tcap + scene_cap_;
    const size_t count_bytes = sizeof(uint32_t) * (size_t)d_.count_size;
    if (!rc) rc = dm((void **)&d_.count, count_bytes, "bucket counts");
    if (!rc) rc = check(hipMemset(d_.count, 0, count_bytes), "bucket counts", err);
    if (!rc) rc = dm((void **)&d_.chain, sizeof(uint64_t) * ((size_t)cblk + 1), "scan chain");
    if (!rc) rc = check(hipMemset(d_.chain, 0, sizeof(uint64_t) * ((size_t)cblk + 1)), "scan chain", err);
    for (int b = 0; b < 2 && !rc; ++b) {
        rc = dm((void **)&d_.wkeys[b], u4, "work keys");
        if (!rc) rc = dm((void **)&d_.wvals[b], u4, "work order");
    }
    if (!rc) rc = dm((void **)&d_.rank, u4, "rank");
    if (!rc) rc = dm((void **)&d_.nbr, u4 * 64, "neighbour rows");
    if (!rc) rc = dm((void **)&d_.ncnt, u4, "neighbour counts");
    if (!rc) rc = dm((void **)&d_.sflags, sizeof(uint32_t) * scene_cap_, "scene flags");
    if (!rc) rc = dm((void **)&d_.scan_err, sizeof(uint32_t), "scan error flag");
    if (!rc) rc = check(hipMemset(d_.scan_err, 0, sizeof(uint32_t)), "scan error flag", err);
    if (!rc) rc = dm((void **)&d_.fbar, sizeof(uint32_t) * kFusedBarWords, "grid barrier");
    if (!rc) rc = check(hipMemset(d_.fbar, 0, sizeof(uint32_t) * kFusedBarWords), "grid barrier", err);
    // wrapped until a predict pass says otherwise
    if (!rc) rc = check(hipMemset(d_.sflags, 0xff, sizeof(uint32_t) * scene_cap_), "scene flags", err);
    d_.ed = d_ed_;
    d_.sd = d_sd_;
    d_.scene_of = d_scene_of_;
    if (rc && rc == SPH_ERR_HIP && err) *err = "device allocation failed: " + *err;
    return rc == SPH_ERR_HIP ? SPH_ERR_OOM : rc;
}

Engine::~Engine()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    invalidate_graph();
    void *bufs[] = {d_ed_,       d_sd_,       d_scene_of_, d_.pos,     d_.vel,
                    d_.xps,      d_.vs,       d_.vs2,      d_.dens,    d_.keys[0], d_.keys[1],
                    d_.ids,      d_.ps,       d_.order,    d_.sids,
                    d_.hist,     d_.bounds,   d_.count,    d_.chain,
                    d_.wkeys[0], d_.wkeys[1], d_.wvals[0], d_.wvals[1], d_.rank,
                    d_.nbr,      d_.ncnt,     d_.sflags,   d_.scan_err, d_.fbar};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (tmp_pos_) (void)hipFree(tmp_pos_);
    if (tmp_vel_) (void)hipFree(tmp_vel_);
    if (colors_buf_) (void)hipFree(colors_buf_);
    if (rpos_) (void)hipFree(rpos_);
    if (ed_h_) (void)hipHostFree(ed_h_);
    for (int q = 0; q < kDescRing; ++q) {
        if (desc_ring_[q]) (void)hipHostFree(desc_ring_[q]);
        if (desc_ev_[q]) (void)hipEventDestroy(desc_ev_[q]);
    }
    for (hipEvent_t &e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::invalidate_graph()
{
    for (int i = 0; i < 4; ++i) {
        if (graph_exec_[i]) (void)hipGraphExecDestroy(graph_exec_[i]);
        if (graph_[i]) (void)hipGraphDestroy(graph_[i]);
        graph_exec_[i] = nullptr;
        graph_[i] = nullptr;
    }
    if (run_exec_) (void)hipGraphExecDestroy(run_exec_);
    if (run_graph_) (void)hipGraphDestroy(run_graph_);
    run_exec_ = nullptr;
    run_graph_ = nullptr;
    run_len_ = 0;
}

int Engine::push_descs(std::string *err)
{
    if (sd_h_.size() > scene_cap_) {
        *err = "scene capacity exceeded";
        return SPH_ERR_INVALID;
    }
    // the next staging slot, once the copies of its previous use have run
    // (kDescRing uploads ago: normally long done)
    const int q = (int)(desc_no_++ % kDescRing);
    if (desc_live_[q]) CK(hipEventSynchronize(desc_ev_[q]), "descriptor staging");
    unsigned char *slot = desc_ring_[q];
    const size_t sd_bytes = sizeof(SceneDesc) * sd_h_.size();
    memcpy(slot, ed_h_, sizeof(EngineDesc));
    if (sd_bytes) memcpy(slot + sizeof(EngineDesc), sd_h_.data(), sd_bytes);
    if (sd_bytes)
        CK(hipMemcpyAsync(d_sd_, slot + sizeof(EngineDesc), sd_bytes, hipMemcpyHostToDevice, stream_),
           "upload scene table");
    CK(hipMemcpyAsync(d_ed_, slot, sizeof(EngineDesc), hipMemcpyHostToDevice, stream_), "upload engine desc");
    CK(hipEventRecord(desc_ev_[q], stream_), "descriptor staging");
    desc_live_[q] = true;
    return SPH_OK;
}

int Engine::repack(const std::vector<RepackItem> &items, std::string *err)
{
    uint64_t n_total = 0, t_total = 0;
    for (const RepackItem &it : items) {
        if (it.spec.n == 0 || it.spec.table == 0) {
            *err = "scene needs n > 0 and table_size > 0";
            return SPH_ERR_INVALID;
        }
        if (const char *m = check_params(it.spec.params)) {
            *err = m;
            return SPH_ERR_INVALID;
        }
        if (it.old_scene >= (int)sd_h_.size() ||
            (it.old_scene >= 0 && sd_h_[it.old_scene].n != it.spec.n) ||
            (it.old_scene < 0 && !it.gen && (!it.pos4 || !it.vel4)) ||
            (it.gen && it.gen->n != it.spec.n)) {
            *err = "repack item does not match an existing scene or lacks its state";
            return SPH_ERR_INVALID;
        }
        n_total += it.spec.n;
        t_total += it.spec.table;
    }
    if (n_total > d_.capacity || t_total > d_.table_capacity || items.size() > scene_cap_) {
        *err = "repack exceeds the engine's capacity";
        return SPH_ERR_INVALID;
    }
    int src = staging(err);
    if (src) return src;
    recorded_ = false;
    // 1. moved scenes: old place -> staging, in particle order; 2. staging ->
    // new place; new scenes: host or generator -> new place.  Moved and new
    // scenes start in particle order; scenes that keep their offset do not
    // move.
    std::vector<uint32_t> new_off(items.size());
    uint32_t off = 0;
    for (size_t q = 0; q < items.size(); ++q) {
        new_off[q] = off;
        off += items[q].spec.n;
    }
    const size_t f4 = sizeof(float4);
    for (size_t q = 0; q < items.size(); ++q) {
        const RepackItem &it = items[q];
        if (it.old_scene < 0 || sd_h_[it.old_scene].p_off == new_off[q]) continue;
        const uint32_t o = sd_h_[it.old_scene].p_off, n = it.spec.n;
        CK(launch_unpermute_state(d_, o, n, tmp_pos_ + new_off[q], tmp_vel_ + new_off[q], stream_),
           "repack move");
    }
    for (size_t q = 0; q < items.size(); ++q) {
        const RepackItem &it = items[q];
        const uint32_t n = it.spec.n, o = new_off[q];
        if (it.old_scene < 0 && it.gen) {
            CK(launch_generate(*it.gen, d_.pos + o, d_.vel + o, stream_), "repack generate");
        } else if (it.old_scene < 0) {
            CK(hipMemcpyAsync(d_.pos + o, it.pos4, f4 * n, hipMemcpyHostToDevice, stream_), "repack upload");
            CK(hipMemcpyAsync(d_.vel + o, it.vel4, f4 * n, hipMemcpyHostToDevice, stream_), "repack upload");
        } else if (sd_h_[it.old_scene].p_off != o) {
            CK(hipMemcpyAsync(d_.pos + o, tmp_pos_ + o, f4 * n, hipMemcpyDeviceToDevice, stream_), "repack move");
            CK(hipMemcpyAsync(d_.vel + o, tmp_vel_ + o, f4 * n, hipMemcpyDeviceToDevice, stream_), "repack move");
        } else {
            continue;  // in place, in its resident order
        }
        CK(launch_iota(d_.ids + o, n, o, stream_), "repack order");
    }
    // descriptors of the new layout
    std::vector<SceneDesc> sd(items.size());
    std::vector<sph_params> params(items.size());
    uint32_t k_off = 0, max_n = 0;
    bool host_sources = false;
    for (size_t q = 0; q < items.size(); ++q) {
        const RepackItem &it = items[q];
        fill_scene_desc(it.spec.params, it.spec.n, it.spec.table, grid_, &sd[q]);
        sd[q].p_off = new_off[q];
        sd[q].index = (uint32_t)q;
        sd[q].key_off = k_off;
        sd[q].dt = dt_cached_ < 0.0f ? 0.0f : dt_cached_;
        k_off += it.spec.table;
        params[q] = it.spec.params;
        max_n = std::max(max_n, it.spec.n);
        host_sources = host_sources || (it.old_scene < 0 && !it.gen);
    }
    // caller buffers must stay alive until their copies end; device-only
    // repacks (moves, generated scenes) leave the stream running
    if (host_sources) CK(hipStreamSynchronize(stream_), "repack sync");
    std::vector<uint8_t> invalid(items.size(), 0);
    for (size_t q = 0; q < items.size(); ++q)
        if (items[q].old_scene >= 0) invalid[q] = invalid_[items[q].old_scene];
    invalid_.swap(invalid);
    sd_h_.swap(sd);
    params_.swap(params);
    ed_h_->n_total = (uint32_t)n_total;
    ed_h_->table_total = (uint32_t)t_total;
    ed_h_->n_scenes = (uint32_t)items.size();
    work_key_layout(ed_h_->n_scenes, ed_h_);
    int rc = push_descs(err);
    if (rc) return rc;
    // the slot -> scene map of the new layout, from the scene table just queued
    CK(launch_scene_map(d_sd_, d_scene_of_, n_scenes(), max_n, stream_), "repack scene map");
    work_valid_ = false;
    have_tables_ = false;
    for (uint32_t q = 0; q < n_scenes(); ++q) CK(refresh_render(q), "repack render positions");
    return SPH_OK;
}

int Engine::staging(std::string *err)
{
    const size_t cap_bytes = sizeof(float4) * (size_t)d_.capacity;
    if (!tmp_pos_) CK(hipMalloc((void **)&tmp_pos_, cap_bytes), "staging");
    if (!tmp_vel_) CK(hipMalloc((void **)&tmp_vel_, cap_bytes), "staging");
    return SPH_OK;
}

hipError_t Engine::normalize_scene(uint32_t s)
{
    const SceneDesc &d = sd_h_[s];
    const size_t bytes = sizeof(float4) * d.n;
    hipError_t e = launch_unpermute_state(d_, d.p_off, d.n, tmp_pos_ + d.p_off, tmp_vel_ + d.p_off,
                                          stream_);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_.pos + d.p_off, tmp_pos_ + d.p_off, bytes, hipMemcpyDeviceToDevice, stream_);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_.vel + d.p_off, tmp_vel_ + d.p_off, bytes, hipMemcpyDeviceToDevice, stream_);
    if (e == hipSuccess) e = launch_iota(d_.ids + d.p_off, d.n, d.p_off, stream_);
    return e;
}

hipError_t Engine::refresh_render(uint32_t s)
{
    if (!rpos_) return hipSuccess;
    const SceneDesc &d = sd_h_[s];
    return launch_unpermute_state(d_, d.p_off, d.n, rpos_ + d.p_off, nullptr, stream_);
}

const float4 *Engine::render_positions(std::string *err)
{
    if (!rpos_) {
        if (check(hipMalloc((void **)&rpos_, sizeof(float4) * (size_t)d_.capacity), "render positions",
                  err))
            return nullptr;
        device_bytes_ += sizeof(float4) * (size_t)d_.capacity;
        for (uint32_t q = 0; q < n_scenes(); ++q)
            if (check(refresh_render(q), "render positions", err)) return nullptr;
        if (check(hipStreamSynchronize(stream_), "render positions", err)) return nullptr;
        d_.rpos = rpos_;
        invalidate_graph();
    }
    return rpos_;
}

int Engine::set_dt(float dt, std::string *err)
{
    if (dt == dt_cached_) return SPH_OK;
    recorded_ = false;  // recorded with the previous dt
    for (SceneDesc &d : sd_h_) d.dt = dt;
    dt_cached_ = dt;
    return push_descs(err);
}

int Engine::set_params(uint32_t s, const sph_params &p, std::string *err)
{
    if (s >= sd_h_.size()) {
        *err = "scene index out of range";
        return SPH_ERR_INVALID;
    }
    if (const char *m = check_params(p)) {
        *err = m;
        return SPH_ERR_INVALID;
    }
    recorded_ = false;
    SceneDesc &d = sd_h_[s];
    const float dt = d.dt;
    const uint32_t p_off = d.p_off, k_off = d.key_off;
    fill_scene_desc(p, d.n, d.table, grid_, &d);
    d.p_off = p_off;
    d.key_off = k_off;
    d.index = s;
    d.dt = dt;
    params_[s] = p;
    return push_descs(err);
}

int Engine::upload(uint32_t s, const float *pos4, const float *vel4, std::string *err)
{
    if (s >= sd_h_.size()) {
        *err = "scene index out of range";
        return SPH_ERR_INVALID;
    }
    const SceneDesc &d = sd_h_[s];
    const size_t bytes = sizeof(float4) * d.n;
    if (!pos4 && !vel4) return SPH_OK;
    recorded_ = false;
    if (!pos4 || !vel4) {
        // the part not uploaded moves into particle order first
        int rc = staging(err);
        if (rc) return rc;
        CK(normalize_scene(s), "upload reorder");
    } else {
        CK(launch_iota(d_.ids + d.p_off, d.n, d.p_off, stream_), "upload order");
    }
    if (pos4)
        CK(hipMemcpyAsync(d_.pos + d.p_off, pos4, bytes, hipMemcpyHostToDevice, stream_),
           "upload positions");
    if (vel4)
        CK(hipMemcpyAsync(d_.vel + d.p_off, vel4, bytes, hipMemcpyHostToDevice, stream_),
           "upload velocities");
    CK(refresh_render(s), "upload render positions");
    CK(hipStreamSynchronize(stream_), "sync after upload");
    if (pos4 && vel4) invalid_[s] = 0;  // (a partial upload keeps the rest of a failed state)
    return SPH_OK;
}

int Engine::set_colors(const float *stops, bool enable, std::string *err)
{
    if (stops) {
        memcpy(ed_h_->stops, stops, sizeof(ed_h_->stops));
        int rc = push_descs(err);
        if (rc) return rc;
    }
    if (enable && !colors_buf_) {
        CK(hipMalloc((void **)&colors_buf_, sizeof(float4) * (size_t)d_.capacity), "colour buffer");
        device_bytes_ += sizeof(float4) * (size_t)d_.capacity;
        CK(hipMemsetAsync(colors_buf_, 0, sizeof(float4) * (size_t)d_.capacity, stream_), "clear colours");
        CK(hipStreamSynchronize(stream_), "sync colours");
    }
    float4 *want = enable ? colors_buf_ : nullptr;
    if (want != d_.colors) {
        d_.colors = want;
        invalidate_graph();
    }
    return SPH_OK;
}

int Engine::download_colors(uint32_t s, float *rgba4, std::string *err)
{
    if (s >= sd_h_.size() || !rgba4) {
        *err = "scene index out of range or null output";
        return SPH_ERR_INVALID;
    }
    if (!colors_buf_) {
        *err = "colours are not enabled";
        return SPH_ERR_STATE;
    }
    const SceneDesc &d = sd_h_[s];
    CK(hipMemcpyAsync(rgba4, colors_buf_ + d.p_off, sizeof(float4) * d.n, hipMemcpyDeviceToHost, stream_),
       "download colours");
    CK(hipStreamSynchronize(stream_), "sync colours");
    return SPH_OK;
}

int Engine::generate(uint32_t s, const GenPlan &g, std::string *err)
{
    if (s >= sd_h_.size() || g.n != sd_h_[s].n) {
        *err = "generate: scene index or particle count does not match";
        return SPH_ERR_INVALID;
    }
    const SceneDesc &d = sd_h_[s];
    recorded_ = false;
    CK(launch_generate(g, d_.pos + d.p_off, d_.vel + d.p_off, stream_), "generate");
    CK(launch_iota(d_.ids + d.p_off, d.n, d.p_off, stream_), "generate order");
    CK(refresh_render(s), "generate render positions");
    CK(hipStreamSynchronize(stream_), "sync after generate");
    invalid_[s] = 0;
    return SPH_OK;
}

hipError_t Engine::enqueue_update(hipEvent_t *ev, bool rebuild_work, bool fresh, bool mark_predict)
{
    // ev (optional): 6 events bracketing predict | spatial (+ work order) |
    // density | pressure | viscosity+integrate; ev[0] only when mark_predict.
    // fresh: the previous update did not record this one's buckets (its
    // viscosity pass does otherwise)
    hipError_t e;
#define LK(x)                              \
    if ((e = (x)) != hipSuccess) return e;
    if (ev && mark_predict) LK(hipEventRecord(ev[0], stream_));
    if (fresh) LK(launch_predict(d_, cfg_, stream_));
    if (ev) LK(hipEventRecord(ev[1], stream_));
    if (rebuild_work) {
        int wb = 0;
        LK(launch_radix_sort(d_, d_.wkeys, d_.wvals, d_.ids, 32, stream_, &wb));
        d_.work = d_.wvals[wb];
        work_valid_ = true;
    }
    if (fused_ok(rebuild_work, fresh) && !ev) {
        // one persistent launch for the whole steady update
        LK(launch_fused_update(d_, stream_));
        return hipSuccess;
    }
    LK(launch_bucket_sort(d_, cfg_, stream_));
    if (ev) LK(hipEventRecord(ev[2], stream_));
    LK(launch_density(d_, cfg_, stream_));
    if (ev) LK(hipEventRecord(ev[3], stream_));
    LK(launch_pressure(d_, cfg_, stream_));
    if (ev) LK(hipEventRecord(ev[4], stream_));
    LK(launch_viscosity_integrate(d_, cfg_, stream_));
    if (ev) LK(hipEventRecord(ev[5], stream_));
#undef LK
    return hipSuccess;
}

int Engine::update(float dt, std::string *err)
{
    if (!isfinite(dt)) {
        *err = "dt must be finite";
        return SPH_ERR_INVALID;
    }
    int rc = set_dt(dt, err);
    if (rc) return rc;
    const bool rebuild = rebuild_due();
    const bool fresh = !recorded_;
    recorded_ = false;  // until this update is enqueued
    if (prof_active_ && prof_used_ * 6 < prof_ev_.size()) {
        // without predict the first event is left out (one packet fewer)
        prof_fresh_[prof_used_] = fresh ? 1 : 0;
        CK(enqueue_update(&prof_ev_[6 * prof_used_], rebuild, fresh, fresh), "launch profiled update");
        ++prof_used_;
    } else if (timing_) {
        CK(enqueue_update(ev_, rebuild, fresh), "launch update");
        CK(hipEventSynchronize(ev_[5]), "sync update");
        for (int i = 0; i < 5; ++i) {
            float ms = 0.0f;
            CK(hipEventElapsedTime(&ms, ev_[i], ev_[i + 1]), "stage timer");
            stage_ms_[i] = ms;
        }
        stage_ms_[5] = 0.0;  // position + collision runs inside the viscosity kernel
    } else if (use_graph_) {
        const int gi = (rebuild ? 1 : 0) + (fresh ? 2 : 0);
        if (!graph_exec_[gi]) {
            CK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
            hipError_t le = enqueue_update(nullptr, rebuild, fresh);
            hipError_t ce = hipStreamEndCapture(stream_, &graph_[gi]);
            CK(le, "capture update");
            CK(ce, "end capture");
            CK(hipGraphInstantiate(&graph_exec_[gi], graph_[gi], nullptr, nullptr, 0),
               "instantiate graph");
        }
        CK(hipGraphLaunch(graph_exec_[gi], stream_), "graph launch");
    } else {
        CK(enqueue_update(nullptr, rebuild, fresh), "launch update");
    }
    ++work_updates_;
    have_tables_ = true;
    recorded_ = true;  // for the next update, if nothing changes and dt stays
    return SPH_OK;
}

int Engine::autotune(float dt, int updates, double ms_out[3], std::string *err)
{
    if (updates < 1 || updates > 64 || !isfinite(dt)) {
        *err = "autotune needs finite dt and 1..64 updates per candidate";
        return SPH_ERR_INVALID;
    }
    if (prof_active_) {
        *err = "autotune while profiling";
        return SPH_ERR_STATE;
    }
    const uint32_t n = ed_h_->n_total;
    float4 *snap = nullptr;
    // positions, velocities and their resident order
    CK(hipMalloc((void **)&snap, sizeof(float4) * 2 * (size_t)n + sizeof(uint32_t) * (size_t)n),
       "autotune snapshot");
    uint32_t *snap_ids = reinterpret_cast<uint32_t *>(snap + 2 * (size_t)n);
    auto save = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(snap, d_.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(snap + n, d_.vel, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(snap_ids, d_.ids, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream_);
        return e;
    };
    auto restore = [&]() -> hipError_t {
        recorded_ = false;  // the state goes back: the next update starts with predict
        hipError_t e = hipMemcpyAsync(d_.pos, snap, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_.vel, snap + n, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_.ids, snap_ids, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream_);
        return e;
    };
    const LaunchCfg keep = cfg_;
    const bool keep_timing = timing_;
    timing_ = false;
    // Candidates: every (work, list, block) for the walk form of all three
    // kernels, then the neighbour-row mode (density writes rows, pressure and
    // viscosity read them) for each block size.  Per-kernel times combine:
    // walk mode = best walk variant of each kernel; row mode = best row-writing
    // density + for pressure and viscosity the better of rows and walk.
    struct Cand {
        GatherVariant v;
        bool nbr;
    };
    std::vector<Cand> cands;
    // (the cell-grid key has no work order; its row mode uses neighbour masks)
    for (int w = 0; w < (grid_ ? 1 : 2); ++w)
        for (int l = 0; l < 2; ++l)
            for (int b : {64, 128, 256}) cands.push_back({GatherVariant{w != 0, l != 0, b, false}, false});
    for (int b : {64, 128, 256}) cands.push_back({GatherVariant{!grid_, false, b, true}, true});
    double walk_ms[3] = {1e30, 1e30, 1e30}, row_ms[3] = {1e30, 1e30, 1e30};
    GatherVariant walk_v[3] = {keep.density, keep.pressure, keep.viscosity}, row_v[3];
    int rc = check(save(), "autotune snapshot", err);
    // two passes over the candidates, each kernel keeping its fastest run:
    // the block sizes of one form differ by about as much as one run's noise
    const size_t trials = 2 * cands.size();
    for (size_t ti = 0; ti < trials && !rc; ++ti) {
        const size_t ci = ti % cands.size();
        LaunchCfg c = keep;
        c.density = c.pressure = c.viscosity = cands[ci].v;
        c.nbr = cands[ci].nbr;
        cfg_ = c;
        invalidate_graph();
        work_valid_ = false;
        // one untimed update (work order, caches), then the timed ones
        rc = update(dt, err);
        if (!rc) rc = profile_begin((uint32_t)updates, err);
        for (int u = 0; u < updates && !rc; ++u) rc = update(dt, err);
        double ms[6] = {0, 0, 0, 0, 0, 0};
        uint32_t used = 0;
        if (!rc) rc = profile_end(ms, &used, err);
        if (!rc) rc = check(restore(), "autotune restore", err);
        for (uint32_t q = 0; q < n_scenes() && !rc; ++q)
            rc = check(refresh_render(q), "autotune restore", err);
        if (rc) break;
        double *best = cands[ci].nbr ? row_ms : walk_ms;
        GatherVariant *pick = cands[ci].nbr ? row_v : walk_v;
        for (int q = 0; q < 3; ++q)
            if (ms[2 + q] < best[q]) {
                best[q] = ms[2 + q];
                pick[q] = cands[ci].v;
            }
    }
    const double walk_total = walk_ms[0] + walk_ms[1] + walk_ms[2];
    GatherVariant rp = row_ms[1] < walk_ms[1] ? row_v[1] : walk_v[1];
    GatherVariant rv = row_ms[2] < walk_ms[2] ? row_v[2] : walk_v[2];
    const double row_total = row_ms[0] + std::min(row_ms[1], walk_ms[1]) + std::min(row_ms[2], walk_ms[2]);
    const bool use_nbr = (rp.rows || rv.rows) && row_total < walk_total;
    double best[3];
    GatherVariant pick[3];
    if (use_nbr) {
        pick[0] = row_v[0];
        pick[1] = rp;
        pick[2] = rv;
        best[0] = row_ms[0];
        best[1] = std::min(row_ms[1], walk_ms[1]);
        best[2] = std::min(row_ms[2], walk_ms[2]);
    } else {
        for (int q = 0; q < 3; ++q) {
            pick[q] = walk_v[q];
            pick[q].rows = false;
            best[q] = walk_ms[q];
        }
    }
    LaunchCfg chosen = keep;
    if (!rc) {
        chosen.density = pick[0];
        chosen.pressure = pick[1];
        chosen.viscosity = pick[2];
        chosen.nbr = use_nbr;
        chosen.fused = false;
    }
    // The fused update (one persistent launch per steady update) against the
    // separate launches of the chosen variants: `updates` steady updates
    // each, timed whole, two runs of each interleaved, the faster kept.
    cfg_ = chosen;
    cfg_.fused = true;
    const bool try_fused = !rc && fused_ok(false, false);
    cfg_ = chosen;
    if (try_fused) {
        hipEvent_t t0 = nullptr, t1 = nullptr;
        rc = check(hipEventCreate(&t0), "autotune timer", err);
        if (!rc) rc = check(hipEventCreate(&t1), "autotune timer", err);
        double best_ms[2] = {1e30, 1e30};
        for (int trial = 0; trial < 4 && !rc; ++trial) {
            const int f = trial & 1;
            cfg_ = chosen;
            cfg_.fused = f != 0;
            invalidate_graph();
            // the first update after the restore is fresh (predict runs); the
            // next one records the buckets, so the timed ones are steady
            rc = update(dt, err);
            if (!rc) rc = update(dt, err);
            if (!rc) rc = check(hipEventRecord(t0, stream_), "autotune timer", err);
            for (int u = 0; u < updates && !rc; ++u) rc = update(dt, err);
            if (!rc) rc = check(hipEventRecord(t1, stream_), "autotune timer", err);
            if (!rc) rc = check(hipEventSynchronize(t1), "autotune timer", err);
            float ms = 0.0f;
            if (!rc) rc = check(hipEventElapsedTime(&ms, t0, t1), "autotune timer", err);
            if (!rc) rc = device_status(-1, err);
            if (!rc) best_ms[f] = std::min(best_ms[f], (double)ms);
            if (!rc) rc = check(restore(), "autotune restore", err);
            for (uint32_t q = 0; q < n_scenes() && !rc; ++q)
                rc = check(refresh_render(q), "autotune restore", err);
        }
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (!rc) chosen.fused = best_ms[1] < best_ms[0];
    }
    if (!rc) rc = check(hipStreamSynchronize(stream_), "autotune sync", err);
    (void)hipFree(snap);
    timing_ = keep_timing;
    cfg_ = chosen;
    invalidate_graph();
    work_valid_ = false;
    have_tables_ = false;
    recorded_ = false;  // the trial updates' records belong to a state restored since
    if (ms_out)
        for (int q = 0; q < 3; ++q) ms_out[q] = best[q];
    return rc;
}

int Engine::profile_begin(uint32_t max_updates, std::string *err)
{
    if (prof_active_) {
        *err = "profiling already active";
        return SPH_ERR_STATE;
    }
    if (max_updates == 0 || max_updates > (1u << 20)) {
        *err = "max_updates must be in [1, 2^20]";
        return SPH_ERR_INVALID;
    }
    prof_ev_.assign(6 * (size_t)max_updates, nullptr);
    prof_fresh_.assign(max_updates, 0);
    // Timing-only events: no system-scope fence (cache writeback and
    // invalidate) when each is recorded, so the stage boundaries cost the
    // profiled updates no more than the kernel boundaries of a plain update.
    // profile_end synchronizes the stream before it reads them.
    for (hipEvent_t &e : prof_ev_)
        CK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "profile event");
    prof_used_ = 0;
    prof_active_ = true;
    return SPH_OK;
}

int Engine::profile_end(double mean_ms[6], uint32_t *updates, std::string *err)
{
    if (!prof_active_) {
        *err = "profiling not active";
        return SPH_ERR_STATE;
    }
    int rc = SPH_OK;
    double sum[6] = {0, 0, 0, 0, 0, 0};
    if (hipStreamSynchronize(stream_) != hipSuccess) rc = SPH_ERR_HIP;
    for (uint32_t u = 0; u < prof_used_ && !rc; ++u)
        for (int i = 0; i < 5; ++i) {
            float ms = 0.0f;
            if (i == 0 && !prof_fresh_[u]) continue;  // no predict: 0 ms
            if (hipEventElapsedTime(&ms, prof_ev_[6 * u + i], prof_ev_[6 * u + i + 1]) != hipSuccess) {
                rc = SPH_ERR_HIP;
                break;
            }
            sum[i] += ms;
        }
    for (int i = 0; i < 6; ++i) mean_ms[i] = prof_used_ ? sum[i] / prof_used_ : 0.0;
    if (updates) *updates = prof_used_;
    for (hipEvent_t &e : prof_ev_)
        if (e) (void)hipEventDestroy(e);
    prof_ev_.clear();
    prof_active_ = false;
    prof_used_ = 0;
    if (rc) *err = "profile readout failed";
    return rc;
}

int Engine::step(float dt, int substeps, bool sync_after, std::string *err)
{
    if (substeps < 1) {
        *err = "substeps must be >= 1";
        return SPH_ERR_INVALID;
    }
    const float sub = dt / (float)substeps;
    for (int i = 0; i < substeps; ++i) {
        const int left = substeps - i;
        // The remaining updates are steady (this dt, buckets recorded, no
        // work-order rebuild, no stage timers): one graph launch runs them
        // all, which saves the gaps between separate graph launches.  While
        // profiling, the first update of the step is the profiled one.
        if (left > 1 && use_graph_ && !timing_ && (!prof_active_ || i > 0) && recorded_ &&
            sub == dt_cached_ && (grid_ || !cfg_.needs_work())) {
            const int rc = run_steady(left, err);
            if (rc) return rc;
            break;
        }
        int rc = update(sub, err);
        if (rc) return rc;
    }
    return sync_after ? sync(err) : SPH_OK;
}

int Engine::run_steady(int n, std::string *err)
{
    if (!run_exec_ || run_len_ != n) {
        if (run_exec_) (void)hipGraphExecDestroy(run_exec_);
        if (run_graph_) (void)hipGraphDestroy(run_graph_);
        run_exec_ = nullptr;
        run_graph_ = nullptr;
        run_len_ = 0;
        CK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
        hipError_t le = hipSuccess;
        for (int u = 0; u < n && le == hipSuccess; ++u) le = enqueue_update(nullptr, false, false);
        hipError_t ce = hipStreamEndCapture(stream_, &run_graph_);
        CK(le, "capture updates");
        CK(ce, "end capture");
        CK(hipGraphInstantiate(&run_exec_, run_graph_, nullptr, nullptr, 0), "instantiate graph");
        run_len_ = n;
    }
    CK(hipGraphLaunch(run_exec_, stream_), "graph launch");
    work_updates_ += (uint64_t)n;
    have_tables_ = true;
    recorded_ = true;
    return SPH_OK;
}

int Engine::sync(std::string *err)
{
    CK(hipStreamSynchronize(stream_), "sync");
    return device_status(-1, err);
}

int Engine::device_status(int s, std::string *err)
{
    uint32_t flag = 0;
    CK(hipMemcpy(&flag, d_.scan_err, sizeof(uint32_t), hipMemcpyDeviceToHost), "read device status");
    if (flag) {
        for (uint8_t &v : invalid_) v = 1;
        CK(hipMemset(d_.scan_err, 0, sizeof(uint32_t)), "clear device status");
        // a fused update's barrier that gave up leaves its count behind
        CK(hipMemset(d_.fbar, 0, sizeof(uint32_t) * kFusedBarWords), "clear device status");
    }
    bool bad = false;
    if (s >= 0) bad = (size_t)s < invalid_.size() && invalid_[s];
    else
        for (uint8_t v : invalid_) bad = bad || v;
    if (!bad) return SPH_OK;
    *err = "a bounded wait of an update gave up (the count scan's look-back, or a grid barrier of "
           "the fused update; device busy beyond its bound): the state is invalid until new state "
           "is uploaded or generated";
    return SPH_ERR_DEVICE;
}

int Engine::set_scan_spin(uint32_t sleeps, std::string *err)
{
    ed_h_->scan_spin = sleeps;
    return push_descs(err);
}

int Engine::prime(std::string *err)
{
    // InitializeData (cc:144-145) runs UpdateSpatialLookup + updateDensities on
    // predictedPositions = positions (cc:550): stage A with dt = 0 and no
    // lookahead gives xp = x exactly, whatever the velocities.
    int rc = set_dt(0.0f, err);
    if (rc) return rc;
    recorded_ = false;
    std::vector<float> lookahead(sd_h_.size());
    for (size_t q = 0; q < sd_h_.size(); ++q) {
        lookahead[q] = sd_h_[q].lookahead;
        sd_h_[q].lookahead = 0.0f;
    }
    rc = push_descs(err);
    hipError_t e = rc ? hipSuccess : launch_predict(d_, cfg_, stream_);
    if (!rc && e == hipSuccess && !grid_ && cfg_.needs_work()) {
        int wb = 0;
        e = launch_radix_sort(d_, d_.wkeys, d_.wvals, d_.ids, 32, stream_, &wb);
        d_.work = d_.wvals[wb];
        work_valid_ = e == hipSuccess;
    }
    if (!rc && e == hipSuccess) e = launch_bucket_sort(d_, cfg_, stream_);
    if (!rc && e == hipSuccess) e = launch_density(d_, cfg_, stream_);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(stream_);
    // the settings' lookahead again for the updates
    for (size_t q = 0; q < sd_h_.size(); ++q) sd_h_[q].lookahead = lookahead[q];
    const int rc2 = push_descs(err);
    if (rc) return rc;
    CK(e, "prime");
    if (rc2) return rc2;
    have_tables_ = true;
    return SPH_OK;
}

int Engine::download(uint32_t s, float *pos4, float *vel4, float *dens2, std::string *err)
{
    if (s >= sd_h_.size()) {
        *err = "scene index out of range";
        return SPH_ERR_INVALID;
    }
    const SceneDesc &d = sd_h_[s];
    CK(hipStreamSynchronize(stream_), "sync before download");
    if (int rc = device_status((int)s, err)) return rc;
    if (pos4 || vel4) {
        // resident order -> particle order on the host
        std::vector<uint32_t> ids(d.n);
        std::vector<float4> buf(d.n);
        CK(hipMemcpy(ids.data(), d_.ids + d.p_off, sizeof(uint32_t) * d.n, hipMemcpyDeviceToHost),
           "download order");
        for (uint32_t k = 0; k < d.n; ++k)
            if (ids[k] - d.p_off >= d.n) {
                *err = "resident order is corrupt";
                return SPH_ERR_STATE;
            }
        for (int a = 0; a < 2; ++a) {
            float *out = a ? vel4 : pos4;
            if (!out) continue;
            CK(hipMemcpy(buf.data(), (a ? d_.vel : d_.pos) + d.p_off, sizeof(float4) * d.n,
                         hipMemcpyDeviceToHost), a ? "download velocities" : "download positions");
            float4 *o4 = reinterpret_cast<float4 *>(out);
            for (uint32_t k = 0; k < d.n; ++k) o4[ids[k] - d.p_off] = buf[k];
        }
    }
    if (dens2) {
        if (!have_tables_) {
            memset(dens2, 0, sizeof(float) * 2 * d.n);
            return SPH_OK;
        }
        // particle order through the repack staging where it exists (its
        // capacity float4s hold the capacity float2s), else a buffer of its own
        float2 *tmp = reinterpret_cast<float2 *>(tmp_pos_);
        if (!tmp) CK(hipMalloc((void **)&tmp, sizeof(float2) * d_.capacity), "density staging");
        hipError_t e1 = launch_unpermute_dens(d_, tmp, stream_);
        hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(stream_) : e1;
        hipError_t e3 = e2 == hipSuccess ? hipMemcpy(dens2, tmp + d.p_off, sizeof(float2) * d.n,
                                                     hipMemcpyDeviceToHost)
                                         : e2;
        if (tmp != reinterpret_cast<float2 *>(tmp_pos_)) (void)hipFree(tmp);
        CK(e3, "download densities");
    }
    return SPH_OK;
}

int Engine::debug_spatial(uint32_t s, uint32_t *hash_id, uint32_t *key_id, uint32_t *sorted_key,
                          uint32_t *sorted_idx, uint32_t *start, uint32_t *end,
                          std::string *err)
{
    if (s >= sd_h_.size()) {
        *err = "scene index out of range";
        return SPH_ERR_INVALID;
    }
    if (!have_tables_) {
        *err = "no update has run yet";
        return SPH_ERR_STATE;
    }
    // everything comes from the sorted side of the last update: the particle
    // index of each sorted slot, the exact hash in ps.w, the bucket bounds
    const SceneDesc &d = sd_h_[s];
    CK(hipStreamSynchronize(stream_), "sync");
    if (int rc = device_status((int)s, err)) return rc;
    std::vector<float4> ps(d.n);
    std::vector<uint32_t> si(d.n), bd((size_t)d.table + 1), sk(d.n, 0xffffffffu);
    CK(hipMemcpy(ps.data(), d_.ps + d.p_off, sizeof(float4) * d.n, hipMemcpyDeviceToHost), "ps");
    CK(hipMemcpy(si.data(), d_.sids + d.p_off, sizeof(uint32_t) * d.n, hipMemcpyDeviceToHost),
       "sorted idx");
    CK(hipMemcpy(bd.data(), d_.bounds + d.key_off, sizeof(uint32_t) * ((size_t)d.table + 1),
                 hipMemcpyDeviceToHost), "bucket bounds");
    for (uint32_t b = 0; b < d.table; ++b)
        for (uint32_t k = bd[b]; k < bd[b + 1]; ++k)
            if (k - d.p_off < d.n) sk[k - d.p_off] = b;
    for (uint32_t k = 0; k < d.n; ++k) si[k] -= d.p_off;
    for (uint32_t k = 0; k < d.n; ++k) {
        if (si[k] >= d.n) continue;
        if (hash_id) memcpy(&hash_id[si[k]], &ps[k].w, 4);
        if (key_id) key_id[si[k]] = sk[k];
    }
    if (sorted_key) memcpy(sorted_key, sk.data(), sizeof(uint32_t) * d.n);
    if (sorted_idx) memcpy(sorted_idx, si.data(), sizeof(uint32_t) * d.n);
    // the reference's table: INT_MAX start for an empty bucket (cc:481)
    for (uint32_t b = 0; b < d.table; ++b) {
        const bool empty = bd[b] == bd[b + 1];
        if (start) start[b] = empty ? SPH_START_EMPTY : bd[b] - d.p_off;
        if (end) end[b] = empty ? 0u : bd[b + 1] - d.p_off;
    }
    return SPH_OK;
}

int Engine::debug_predicted(uint32_t s, float *xp4, std::string *err)
{
    if (s >= sd_h_.size()) {
        *err = "scene index out of range";
        return SPH_ERR_INVALID;
    }
    if (!have_tables_) {
        *err = "no update has run yet";
        return SPH_ERR_STATE;
    }
    // the sorted copy of the last update's predicted positions
    const SceneDesc &d = sd_h_[s];
    CK(hipStreamSynchronize(stream_), "sync");
    if (int rc = device_status((int)s, err)) return rc;
    std::vector<float4> xps(d.n);
    std::vector<uint32_t> si(d.n);
    CK(hipMemcpy(xps.data(), d_.xps + d.p_off, sizeof(float4) * d.n, hipMemcpyDeviceToHost), "xps");
    CK(hipMemcpy(si.data(), d_.sids + d.p_off, sizeof(uint32_t) * d.n, hipMemcpyDeviceToHost),
       "sorted idx");
    for (uint32_t k = 0; k < d.n; ++k) {
        const uint32_t i = si[k] - d.p_off;
        if (i >= d.n) continue;
        xp4[4 * i + 0] = xps[k].x;
        xp4[4 * i + 1] = xps[k].y;
        xp4[4 * i + 2] = xps[k].z;
        xp4[4 * i + 3] = 0.0f;
    }
    return SPH_OK;
}

int Engine::debug_velocities(uint32_t s, float *v_ext4, float *v_press4, std::string *err)
{
    if (s >= sd_h_.size()) {
        *err = "scene index out of range";
        return SPH_ERR_INVALID;
    }
    if (!have_tables_) {
        *err = "no update has run yet";
        return SPH_ERR_STATE;
    }
    // vs (after the external force) and vs2 (after pressure) of the last
    // update, in its sorted order
    const SceneDesc &d = sd_h_[s];
    CK(hipStreamSynchronize(stream_), "sync");
    if (int rc = device_status((int)s, err)) return rc;
    std::vector<uint32_t> si(d.n);
    std::vector<float4> buf(d.n);
    CK(hipMemcpy(si.data(), d_.sids + d.p_off, sizeof(uint32_t) * d.n, hipMemcpyDeviceToHost),
       "sorted idx");
    for (int a = 0; a < 2; ++a) {
        float *out = a ? v_press4 : v_ext4;
        if (!out) continue;
        CK(hipMemcpy(buf.data(), (a ? d_.vs2 : d_.vs) + d.p_off, sizeof(float4) * d.n,
                     hipMemcpyDeviceToHost), "stage velocities");
        for (uint32_t k = 0; k < d.n; ++k) {
            const uint32_t i = si[k] - d.p_off;
            if (i >= d.n) continue;
            memcpy(out + 4 * (size_t)i, &buf[k], sizeof(float4));
        }
    }
    return SPH_OK;
}

int Engine::debug_neighbours(uint32_t s, uint64_t *offs, uint32_t *list, uint64_t cap,
                             uint64_t *total, std::string *err)
{
    if (s >= sd_h_.size()) {
        *err = "scene index out of range";
        return SPH_ERR_INVALID;
    }
    if (!have_tables_) {
        *err = "no update has run yet";
        return SPH_ERR_STATE;
    }
    const SceneDesc &d = sd_h_[s];
    const uint32_t n_all = ed_h_->n_total;
    uint32_t *d_counts = nullptr;
    uint64_t *d_offs = nullptr;
    uint32_t *d_list = nullptr;
    std::vector<uint32_t> counts(n_all);
    std::vector<uint64_t> offs_all(n_all + 1);
    int rc = SPH_OK;
    do {
        if ((rc = check(hipMalloc((void **)&d_counts, sizeof(uint32_t) * n_all), "counts", err)))
            break;
        if ((rc = check(launch_neighbour_count(d_, d_counts, stream_), "count", err)))
            break;
        if ((rc = check(hipStreamSynchronize(stream_), "sync", err))) break;
        if ((rc = check(hipMemcpy(counts.data(), d_counts, sizeof(uint32_t) * n_all,
                                  hipMemcpyDeviceToHost), "counts", err)))
            break;
        offs_all[0] = 0;
        for (uint32_t i = 0; i < n_all; ++i) offs_all[i + 1] = offs_all[i] + counts[i];
        const uint64_t tot_all = offs_all[n_all];
        if ((rc = check(hipMalloc((void **)&d_offs, sizeof(uint64_t) * (n_all + 1)), "offs", err)))
            break;
        if ((rc = check(hipMalloc((void **)&d_list, sizeof(uint32_t) * std::max<uint64_t>(tot_all, 1)),
                        "list", err)))
            break;
        if ((rc = check(hipMemcpy(d_offs, offs_all.data(), sizeof(uint64_t) * (n_all + 1),
                                  hipMemcpyHostToDevice), "offs", err)))
            break;
        if ((rc = check(launch_neighbour_fill(d_, d_offs, d_list, stream_), "fill", err)))
            break;
        if ((rc = check(hipStreamSynchronize(stream_), "sync", err))) break;
        const uint64_t base = offs_all[d.p_off];
        const uint64_t tot = offs_all[d.p_off + d.n] - base;
        if (offs)
            for (uint32_t i = 0; i <= d.n; ++i) offs[i] = offs_all[d.p_off + i] - base;
        if (total) *total = tot;
        if (list && cap) {
            const uint64_t m = std::min(cap, tot);
            if (m && (rc = check(hipMemcpy(list, d_list + base, sizeof(uint32_t) * m,
                                           hipMemcpyDeviceToHost), "list", err)))
                break;
        }
    } while (0);
    if (d_counts) (void)hipFree(d_counts);
    if (d_offs) (void)hipFree(d_offs);
    if (d_list) (void)hipFree(d_list);
    return rc;
}

int Engine::count_nonfinite(uint64_t *count, std::string *err)
{
    uint32_t *d_c = nullptr;
    CK(hipMalloc((void **)&d_c, sizeof(uint32_t)), "counter");
    uint32_t h = 0;
    hipError_t e = hipMemsetAsync(d_c, 0, sizeof(uint32_t), stream_);
    if (e == hipSuccess) e = launch_count_nonfinite(d_, d_c, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e == hipSuccess) e = hipMemcpy(&h, d_c, sizeof(uint32_t), hipMemcpyDeviceToHost);
    (void)hipFree(d_c);
    CK(e, "count non-finite");
    *count = h;
    return SPH_OK;
}

}  // namespace sph
