// sph_capi.cpp -- the C ABI of include/sph_amd.h over sph::Engine.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "../../include/sph_amd.h"
#include "sph_device.h"
#include "sph_engine.h"

struct sph_sim {
    sph::Engine *eng;
};

struct sph_batch_t {
    sph::Engine *eng;
};

namespace {
thread_local std::string g_err;
}  // namespace

namespace sph {
int set_last_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}
}  // namespace sph

namespace {

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int done(int rc, const std::string &msg)
{
    if (rc != SPH_OK) g_err = msg;
    return rc;
}

bool have_gpu()
{
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}
}  // namespace

#define SPH_GUARD_BEGIN try {
#define SPH_GUARD_END                                                     \
    }                                                                     \
    catch (const std::bad_alloc &) { return fail(SPH_ERR_OOM, "host allocation failed"); } \
    catch (const std::exception &ex) { return fail(SPH_ERR_INVALID, ex.what()); }          \
    catch (...) { return fail(SPH_ERR_INVALID, "unknown exception"); }

#define SPH_NEED(h)                                                       \
    if (!(h) || !(h)->eng) return fail(SPH_ERR_INVALID, "null handle")

extern "C" {

void sph_default_params(sph_params *p)
{
    if (!p) return;
    p->smoothing_radius = 0.35f;       // physicsWorld.h:97
    p->target_density = 99.7f;         // physicsWorld.h:98
    p->pressure_mult = 300.0f;         // physicsWorld.h:99
    p->near_pressure_mult = 20.0f;     // physicsWorld.h:100
    p->viscosity = 0.5f;               // physicsWorld.h:101
    p->gravity_on = 0;                 // physicsWorld.h:105
    p->gravity_scale = 10.0f;          // physicsWorld.h:106
    p->bounds[0] = p->bounds[1] = p->bounds[2] = 20.0f;  // physicsWorld.h:145
    p->lookahead = 1.0f / 120.0f;      // physicsWorld.cc:47
    p->damp = 0.95f;                   // physicsWorld.cc:87
    p->prefilter_r2 = 0.35f * 0.35f;   // physicsWorld.h:96
}

const char *sph_last_error(void) { return g_err.c_str(); }

const char *sph_version(void) { return "sph_amd 0.1 (gfx950)"; }

uint32_t sph_hash_cell(int32_t cx, int32_t cy, int32_t cz) { return sph::hash_cell(cx, cy, cz); }

int sph_debug_quotients(const float *a, const float *b, float *out, uint64_t n)
{
    if (n && (!a || !b || !out)) return fail(SPH_ERR_INVALID, "null argument");
    for (uint64_t i = 0; i < n; ++i) {
        const float y = sph::reciprocal_or_zero(b[i]);
        out[i] = (y != 0.0f && sph::quotients_ok(a[i])) ? sph::quotient(a[i], b[i], y) : a[i] / b[i];
    }
    return SPH_OK;
}

uint32_t sph_bucket_of(uint32_t hash, uint32_t table_size)
{
    if (table_size == 0) return 0;
    return sph::fastmod(hash, sph::fastmod_constant(table_size), table_size);
}

int sph_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int sph_debug_reciprocal_mismatches(uint32_t lo_bits, uint32_t hi_bits, uint64_t *count,
                                    uint32_t *first)
{
    SPH_GUARD_BEGIN
    if (!count || lo_bits > hi_bits) return fail(SPH_ERR_INVALID, "null count or empty range");
    if (!have_gpu()) return fail(SPH_ERR_HIP, "no HIP device visible");
    struct Out {
        unsigned long long bad;
        uint32_t first;
        uint32_t pad;
    } host{0ull, 0xffffffffu, 0u};
    Out *dev = nullptr;
    hipStream_t st = nullptr;
    hipError_t e = hipMalloc((void **)&dev, sizeof(Out));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemcpyAsync(dev, &host, sizeof(Out), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = sph::launch_recip_check(lo_bits, hi_bits, &dev->bad, &dev->first, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&host, dev, sizeof(Out), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (st) (void)hipStreamDestroy(st);
    if (dev) (void)hipFree(dev);
    if (e != hipSuccess) return fail(SPH_ERR_HIP, std::string("reciprocal check: ") + hipGetErrorString(e));
    *count = host.bad;
    if (first) *first = host.first;
    return SPH_OK;
    SPH_GUARD_END
}

int sph_copy_calibration(uint64_t bytes, int reps, double *ms_per_copy)
{
    SPH_GUARD_BEGIN
    if (bytes < 16 || reps < 1) return fail(SPH_ERR_INVALID, "bytes >= 16 and reps >= 1 required");
    if (!have_gpu()) return fail(SPH_ERR_HIP, "no HIP device visible");
    const size_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void **)&a, n * 16);
    if (e == hipSuccess) e = hipMalloc((void **)&b, n * 16);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * 16, st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = sph::launch_copy(a, b, n, st);  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = sph::launch_copy(a, b, n, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_copy) *ms_per_copy = ms / reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(SPH_ERR_HIP, std::string("copy calibration: ") + hipGetErrorString(e));
    return SPH_OK;
    SPH_GUARD_END
}

int sph_create(const sph_params *p, uint32_t n, uint32_t table_size, sph_handle *out)
{
    SPH_GUARD_BEGIN
    if (!p || !out) return fail(SPH_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n == 0) return fail(SPH_ERR_INVALID, "n must be > 0");
    if (!have_gpu()) return fail(SPH_ERR_HIP, "no HIP device visible");
    if (const char *m = sph::check_params(*p)) return fail(SPH_ERR_INVALID, m);
    const bool grid = table_size == SPH_TABLE_GRID;
    const uint32_t t = grid ? sph::grid_table_size(*p, n)
                            : (table_size == SPH_TABLE_REFERENCE ? n : table_size);
    sph::SceneSpec spec{*p, n, t};
    std::string err;
    sph::Engine *e = nullptr;
    int rc = sph::Engine::create({spec}, n, spec.table, 1, &e, &err, false, grid);
    if (rc) return fail(rc, err);
    *out = new sph_sim{e};
    return SPH_OK;
    SPH_GUARD_END
}

int sph_destroy(sph_handle h)
{
    SPH_GUARD_BEGIN
    if (!h) return SPH_OK;
    delete h->eng;
    delete h;
    return SPH_OK;
    SPH_GUARD_END
}

int sph_set_params(sph_handle h, const sph_params *p)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    if (!p) return fail(SPH_ERR_INVALID, "null params");
    std::string err;
    return done(h->eng->set_params(0, *p, &err), err);
    SPH_GUARD_END
}

int sph_get_params(sph_handle h, sph_params *p)
{
    SPH_NEED(h);
    if (!p) return fail(SPH_ERR_INVALID, "null params");
    *p = h->eng->params(0);
    return SPH_OK;
}

int sph_get_size(sph_handle h, uint32_t *n, uint32_t *table_size)
{
    SPH_NEED(h);
    if (n) *n = h->eng->scene(0).n;
    if (table_size) *table_size = h->eng->scene(0).table;
    return SPH_OK;
}

int sph_get_key(sph_handle h, int32_t *grid, uint32_t gdim[3], int32_t lo[3])
{
    SPH_NEED(h);
    const SceneDesc &d = h->eng->scene(0);
    if (grid) *grid = (int32_t)d.grid;
    for (int c = 0; c < 3; ++c) {
        if (gdim) gdim[c] = d.gdim[c];
        if (lo) lo[c] = d.cell_lo[c];
    }
    return SPH_OK;
}

int sph_upload(sph_handle h, const float *pos4, const float *vel4)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->upload(0, pos4, vel4, &err), err);
    SPH_GUARD_END
}

int sph_initialize(sph_handle h, int kind, uint64_t seed, const float box[3], float spacing,
                   float jitter, float speed)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    sph::GenPlan g;
    int rc = sph::plan_generate(kind, seed, h->eng->scene(0).n, box, spacing, jitter, speed, &g);
    if (rc) return fail(rc, "invalid scene generator arguments");
    std::string err;
    rc = h->eng->generate(0, g, &err);
    if (!rc) rc = h->eng->prime(&err);
    return done(rc, err);
    SPH_GUARD_END
}

int sph_initialize_data(sph_handle h)
{
    return sph_initialize(h, SPH_INIT_GRID_ARRANGEMENT, 0, nullptr, 0.0f, 0.0f, 0.0f);
}

int sph_set_colors(sph_handle h, const float stops[16], int enable)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    if (stops)
        for (int c = 0; c < 16; ++c)
            if (!isfinite(stops[c])) return fail(SPH_ERR_INVALID, "colour stops must be finite");
    std::string err;
    return done(h->eng->set_colors(stops, enable != 0, &err), err);
    SPH_GUARD_END
}

int sph_download_colors(sph_handle h, float *rgba4)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    int rc = h->eng->sync(&err);
    if (!rc) rc = h->eng->download_colors(0, rgba4, &err);
    return done(rc, err);
    SPH_GUARD_END
}

int sph_render_buffers(sph_handle h, const void **positions, const void **colors)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    const uint32_t off = h->eng->scene(0).p_off;
    int rc = SPH_OK;
    if (positions) {
        // kept in particle order from the first request on
        const float4 *rp = h->eng->render_positions(&err);
        if (!rp) rc = SPH_ERR_HIP;
        *positions = rp ? rp + off : nullptr;
    }
    if (colors) *colors = h->eng->colors_buffer() ? h->eng->colors_buffer() + off : nullptr;
    return done(rc, err);
    SPH_GUARD_END
}

int sph_update(sph_handle h, float dt)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    int rc = h->eng->update(dt, &err);
    if (!rc) rc = h->eng->sync(&err);
    return done(rc, err);
    SPH_GUARD_END
}

int sph_step(sph_handle h, float dt, int substeps)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->step(dt, substeps, true, &err), err);
    SPH_GUARD_END
}

int sph_step_async(sph_handle h, float dt, int substeps)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->step(dt, substeps, false, &err), err);
    SPH_GUARD_END
}

int sph_sync(sph_handle h)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->sync(&err), err);
    SPH_GUARD_END
}

int sph_download(sph_handle h, float *pos4, float *vel4, float *dens2)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->download(0, pos4, vel4, dens2, &err), err);
    SPH_GUARD_END
}

int sph_set_timing(sph_handle h, int on)
{
    SPH_NEED(h);
    h->eng->set_timing(on != 0);
    return SPH_OK;
}

int sph_stage_ms(sph_handle h, double ms[6])
{
    SPH_NEED(h);
    if (!ms) return fail(SPH_ERR_INVALID, "null output");
    memcpy(ms, h->eng->stage_ms(), sizeof(double) * 6);
    return SPH_OK;
}

int sph_profile_begin(sph_handle h, uint32_t max_updates)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->profile_begin(max_updates, &err), err);
    SPH_GUARD_END
}

int sph_profile_end(sph_handle h, double mean_ms[6], uint32_t *updates)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    if (!mean_ms) return fail(SPH_ERR_INVALID, "null output");
    std::string err;
    return done(h->eng->profile_end(mean_ms, updates, &err), err);
    SPH_GUARD_END
}

int sph_autotune(sph_handle h, float dt, int updates, double ms[3])
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->autotune(dt, updates, ms, &err), err);
    SPH_GUARD_END
}

namespace {
int get_variants(sph::Engine *e, int32_t v[14])
{
    if (!v) return fail(SPH_ERR_INVALID, "null output");
    const sph::LaunchCfg &c = e->launch_cfg();
    const sph::GatherVariant g[3] = {c.density, c.pressure, c.viscosity};
    for (int q = 0; q < 3; ++q) {
        v[4 * q + 0] = g[q].work;
        v[4 * q + 1] = g[q].list;
        v[4 * q + 2] = g[q].block;
        v[4 * q + 3] = g[q].rows;
    }
    v[12] = c.nbr;
    v[13] = c.fused;
    return SPH_OK;
}

int set_variants(sph::Engine *e, const int32_t v[14])
{
    if (!v) return fail(SPH_ERR_INVALID, "null input");
    sph::LaunchCfg c = e->launch_cfg();
    c.nbr = v[12] != 0;
    sph::GatherVariant *g[3] = {&c.density, &c.pressure, &c.viscosity};
    for (int q = 0; q < 3; ++q) {
        const int b = v[4 * q + 2];
        if (b != 64 && b != 128 && b != 192 && b != 256)
            return fail(SPH_ERR_INVALID, "block must be 64, 128, 192 or 256");
        g[q]->work = v[4 * q + 0] != 0;
        g[q]->list = v[4 * q + 1] != 0;
        g[q]->block = b;
        g[q]->rows = v[4 * q + 3] != 0;
    }
    if ((c.pressure.rows || c.viscosity.rows) && !c.nbr)
        return fail(SPH_ERR_INVALID, "reading neighbour rows needs the row mode (v[12])");
    c.fused = v[13] != 0;
    if (c.fused && !(c.nbr && c.density.rows && c.pressure.rows && c.viscosity.rows))
        return fail(SPH_ERR_INVALID, "the fused update (v[13]) needs the row mode in all three kernels");
    e->set_launch_cfg(c);
    return SPH_OK;
}
}  // namespace

int sph_get_variants(sph_handle h, int32_t v[14])
{
    SPH_NEED(h);
    return get_variants(h->eng, v);
}

int sph_set_variants(sph_handle h, const int32_t v[14])
{
    SPH_NEED(h);
    return set_variants(h->eng, v);
}

// ---- batches ----

int sph_batch_create(uint32_t n_scenes, const uint32_t *counts, const uint32_t *table_sizes,
                     const sph_params *params, sph_batch *out)
{
    SPH_GUARD_BEGIN
    if (!out || !counts || !params || n_scenes == 0)
        return fail(SPH_ERR_INVALID, "need n_scenes > 0, counts, params and out");
    *out = nullptr;
    if (!have_gpu()) return fail(SPH_ERR_HIP, "no HIP device visible");
    std::vector<sph::SceneSpec> specs(n_scenes);
    const bool grid = table_sizes == nullptr;
    for (uint32_t s = 0; s < n_scenes; ++s) {
        if (const char *m = sph::check_params(params[s])) return fail(SPH_ERR_INVALID, m);
        uint32_t t;
        if (grid) t = sph::grid_table_size(params[s], counts[s]);
        else if (table_sizes[s] == 0 || table_sizes[s] == SPH_TABLE_REFERENCE) t = counts[s];
        else t = table_sizes[s];
        specs[s] = sph::SceneSpec{params[s], counts[s], t};
    }
    std::string err;
    sph::Engine *e = nullptr;
    int rc = sph::Engine::create(specs, 0, 0, n_scenes, &e, &err, false, grid);
    if (rc) return fail(rc, err);
    *out = new sph_batch_t{e};
    return SPH_OK;
    SPH_GUARD_END
}

int sph_batch_destroy(sph_batch b)
{
    SPH_GUARD_BEGIN
    if (!b) return SPH_OK;
    delete b->eng;
    delete b;
    return SPH_OK;
    SPH_GUARD_END
}

#define SPH_BATCH_CALL(expr)                     \
    SPH_GUARD_BEGIN                              \
    SPH_NEED(b);                                 \
    std::string err;                             \
    return done((expr), err);                    \
    SPH_GUARD_END

int sph_batch_upload(sph_batch b, uint32_t scene, const float *pos4, const float *vel4)
{
    SPH_BATCH_CALL(b->eng->upload(scene, pos4, vel4, &err))
}

int sph_batch_initialize(sph_batch b, uint32_t scene, int kind, uint64_t seed, const float box[3],
                         float spacing, float jitter, float speed)
{
    SPH_GUARD_BEGIN
    SPH_NEED(b);
    if (scene >= b->eng->n_scenes()) return fail(SPH_ERR_INVALID, "scene index out of range");
    sph::GenPlan g;
    if (sph::plan_generate(kind, seed, b->eng->scene(scene).n, box, spacing, jitter, speed, &g))
        return fail(SPH_ERR_INVALID, "invalid scene generator arguments");
    std::string err;
    return done(b->eng->generate(scene, g, &err), err);
    SPH_GUARD_END
}

int sph_batch_set_params(sph_batch b, uint32_t scene, const sph_params *p)
{
    if (!p) return fail(SPH_ERR_INVALID, "null params");
    SPH_BATCH_CALL(b->eng->set_params(scene, *p, &err))
}

int sph_batch_step(sph_batch b, float dt, int substeps)
{
    SPH_BATCH_CALL(b->eng->step(dt, substeps, true, &err))
}

int sph_batch_step_async(sph_batch b, float dt, int substeps)
{
    SPH_BATCH_CALL(b->eng->step(dt, substeps, false, &err))
}

int sph_batch_sync(sph_batch b) { SPH_BATCH_CALL(b->eng->sync(&err)) }

int sph_batch_download(sph_batch b, uint32_t scene, float *pos4, float *vel4, float *dens2)
{
    SPH_BATCH_CALL(b->eng->download(scene, pos4, vel4, dens2, &err))
}

int sph_batch_autotune(sph_batch b, float dt, int updates, double ms[3])
{
    SPH_BATCH_CALL(b->eng->autotune(dt, updates, ms, &err))
}

int sph_batch_profile_begin(sph_batch b, uint32_t max_updates)
{
    SPH_BATCH_CALL(b->eng->profile_begin(max_updates, &err))
}

int sph_batch_profile_end(sph_batch b, double mean_ms[6], uint32_t *updates)
{
    if (!mean_ms) return fail(SPH_ERR_INVALID, "null output");
    SPH_BATCH_CALL(b->eng->profile_end(mean_ms, updates, &err))
}

int sph_batch_get_variants(sph_batch b, int32_t v[14])
{
    SPH_NEED(b);
    return get_variants(b->eng, v);
}

int sph_batch_set_variants(sph_batch b, const int32_t v[14])
{
    SPH_NEED(b);
    return set_variants(b->eng, v);
}

int sph_debug_set_scan_spin(sph_handle h, uint32_t sleeps)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->set_scan_spin(sleeps, &err), err);
    SPH_GUARD_END
}

int sph_count_nonfinite(sph_handle h, uint64_t *count)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    if (!count) return fail(SPH_ERR_INVALID, "null output");
    std::string err;
    return done(h->eng->count_nonfinite(count, &err), err);
    SPH_GUARD_END
}

int sph_debug_spatial(sph_handle h, uint32_t *hash_id, uint32_t *key_id, uint32_t *sorted_key,
                      uint32_t *sorted_idx, uint32_t *start, uint32_t *end)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->debug_spatial(0, hash_id, key_id, sorted_key, sorted_idx, start, end,
                                      &err),
                err);
    SPH_GUARD_END
}

int sph_debug_predicted(sph_handle h, float *xp4)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    if (!xp4) return fail(SPH_ERR_INVALID, "null output");
    std::string err;
    return done(h->eng->debug_predicted(0, xp4, &err), err);
    SPH_GUARD_END
}

int sph_debug_velocities(sph_handle h, float *v_ext4, float *v_press4)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->debug_velocities(0, v_ext4, v_press4, &err), err);
    SPH_GUARD_END
}

int sph_debug_neighbours(sph_handle h, uint64_t *offs, uint32_t *list, uint64_t cap,
                         uint64_t *total)
{
    SPH_GUARD_BEGIN
    SPH_NEED(h);
    std::string err;
    return done(h->eng->debug_neighbours(0, offs, list, cap, total, &err), err);
    SPH_GUARD_END
}

}  // extern "C"
