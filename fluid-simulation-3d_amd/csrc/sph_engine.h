// sph_engine.h -- host side of the gfx950 SPH step: owns the device buffers,
// the HIP stream and the captured graph of one Update(dt) over S scenes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/sph_amd.h"
#include "sph_launch.h"
#include "sph_types.h"

namespace sph {

struct SceneSpec {
    sph_params params;
    uint32_t n;
    uint32_t table;
};

// One scene of a repacked layout: an existing scene (old_scene >= 0, its
// state moves on the device) or a new one (state uploaded from host).
struct RepackItem {
    SceneSpec spec;
    int old_scene;
    const float *pos4;
    const float *vel4;
    const GenPlan *gen = nullptr;  // new scene generated on the device instead
};

// Fills the derived constants of a SceneDesc from user settings, with the
// reference's float expressions (kernels.h:25-82, physicsWorld.cc:47,87-88).
// grid: bucket by the cell-grid key (periods chosen within `table` buckets)
// instead of hash % table.
void fill_scene_desc(const sph_params &p, uint32_t n, uint32_t table, bool grid, SceneDesc *d);
// Bucket count of the cell-grid key for a scene: the box (+4 cells of margin
// per side) in cells, capped at max(8n, 2^20) and 2^28.
uint32_t grid_table_size(const sph_params &p, uint32_t n);
// Smallest float x >= 0 with sqrtf(x) >= h (so d < h <=> d*d-as-float < x).
float sqrt_threshold(float h);
// Validates settings; returns an error message or nullptr.
const char *check_params(const sph_params &p);

class Engine {
  public:
    // capacity: particle slots to allocate (>= total of scenes); table_capacity likewise.
    // multi: always run the multi-scene kernels (a layout that changes over
    // time may hold one scene now and many later).  An empty `scenes` is
    // allowed with explicit capacities (a server starts empty).
    static int create(const std::vector<SceneSpec> &scenes, uint32_t capacity,
                      uint32_t table_capacity, uint32_t scene_capacity, Engine **out,
                      std::string *err, bool multi = false, bool grid = false);
    // Replaces the scene list in place: buffers, grids and captured graphs stay
    // valid (the kernels read the counts from device memory).  Fails without
    // changes if the items exceed the capacities.
    int repack(const std::vector<RepackItem> &items, std::string *err);
    uint32_t capacity() const { return d_.capacity; }
    uint32_t table_capacity() const { return d_.table_capacity; }
    uint32_t scene_capacity() const { return scene_cap_; }
    ~Engine();

    uint32_t n_scenes() const { return (uint32_t)sd_h_.size(); }
    uint32_t n_total() const { return ed_h_->n_total; }
    const SceneDesc &scene(uint32_t s) const { return sd_h_[s]; }
    const sph_params &params(uint32_t s) const { return params_[s]; }
    hipStream_t stream() const { return stream_; }

    int set_params(uint32_t s, const sph_params &p, std::string *err);
    int upload(uint32_t s, const float *pos4, const float *vel4, std::string *err);
    // Initial state of scene s generated on the device (g.n must equal its count).
    int generate(uint32_t s, const GenPlan &g, std::string *err);
    // Render hand-off: gradient stops (16 floats, null keeps them) and
    // whether each update writes the speed colours (particle order, RGBA).
    int set_colors(const float *stops, bool enable, std::string *err);
    int download_colors(uint32_t s, float *rgba4, std::string *err);
    const float4 *colors_buffer() const { return colors_buf_; }
    // Render positions (OutPositions, particle order): allocated and filled
    // from the current state on the first call; every later update rewrites
    // them.  Null on failure (err set).
    const float4 *render_positions(std::string *err);
    // One Update(dt) for every scene (physicsWorld.cc:39-111).
    int update(float dt, std::string *err);
    // `substeps` Update(dt / substeps) calls; blocks until done when sync.
    int step(float dt, int substeps, bool sync, std::string *err);
    int sync(std::string *err);
    // Bound of the count scan's look-back wait in 64-clock sleeps (default
    // kScanSpinDefault).  A wait that reaches it makes the update's results
    // invalid, and the next synchronising call reports SPH_ERR_DEVICE.
    int set_scan_spin(uint32_t sleeps, std::string *err);
    static constexpr uint32_t kScanSpinDefault = 1u << 16;
    // InitializeData tail (cc:144-145): spatial lookup + densities of the current positions.
    int prime(std::string *err);
    int download(uint32_t s, float *pos4, float *vel4, float *dens2, std::string *err);
    int debug_spatial(uint32_t s, uint32_t *hash_id, uint32_t *key_id, uint32_t *sorted_key,
                      uint32_t *sorted_idx, uint32_t *start, uint32_t *end, std::string *err);
    int debug_predicted(uint32_t s, float *xp4, std::string *err);
    int debug_velocities(uint32_t s, float *v_ext4, float *v_press4, std::string *err);
    int debug_neighbours(uint32_t s, uint64_t *offs, uint32_t *list, uint64_t cap,
                         uint64_t *total, std::string *err);
    int count_nonfinite(uint64_t *count, std::string *err);

    void set_timing(bool on) { timing_ = on; }
    // Event pool for per-stage timing of up to `max_updates` updates without
    // synchronising: stage boundaries are recorded on the engine's stream.
    int profile_begin(uint32_t max_updates, std::string *err);
    // Mean per-stage milliseconds over the profiled updates (stage order of
    // stage_ms), their count, and the pool is released.
    int profile_end(double mean_ms[6], uint32_t *updates, std::string *err);
    const double *stage_ms() const { return stage_ms_; }
    const LaunchCfg &launch_cfg() const { return cfg_; }
    // Variants that do not exist under the engine's key are turned off
    // (the cell-grid key runs in sorted order: no work order, no rows).
    void set_launch_cfg(const LaunchCfg &c)
    {
        cfg_ = sanitize_launch_cfg(c, grid_);
        invalidate_graph();
    }
    bool grid() const { return grid_; }
    // Times every gather-kernel variant on the current state (`updates`
    // Update(dt) calls per candidate, twice over the candidates, from a
    // snapshot of positions, velocities and their order that is restored
    // afterwards) and keeps the fastest variant of each kernel.  ms_out (may be null) gets the chosen variants' mean ms of
    // density, pressure, viscosity.  Densities and the debug tables are
    // recomputed by the next update.
    int autotune(float dt, int updates, double ms_out[3], std::string *err);
    void invalidate_graph();
    // the fused update applies: chosen, grid key with neighbour masks in all
    // three kernels, a steady update, and a launch the device holds resident
    bool fused_ok(bool rebuild_work, bool fresh) const
    {
        return cfg_.fused && grid_ && cfg_.nbr && cfg_.density.rows && cfg_.pressure.rows &&
               cfg_.viscosity.rows && !rebuild_work && !fresh && d_.fused_grid != 0;
    }
    int set_use_graph(bool on) { use_graph_ = on; invalidate_graph(); return 0; }
    // Mean duration (ms) of each gather kernel over the last timed update.
    uint64_t device_bytes() const { return device_bytes_; }
    const DevState &dev() const { return d_; }
    void set_work_period(uint32_t p) { work_period_ = p ? p : 1; }

  private:
    Engine() = default;
    int alloc(uint32_t capacity, uint32_t table_capacity, std::string *err);
    int push_descs(std::string *err);
    int set_dt(float dt, std::string *err);
    hipError_t enqueue_update(hipEvent_t *ev, bool rebuild_work, bool fresh, bool mark_predict = true);
    // Resident slots of scene s into particle order (ids[k] = p_off + k).
    hipError_t normalize_scene(uint32_t s);
    // Render positions of scene s from its current state (if they exist).
    hipError_t refresh_render(uint32_t s);
    int run_steady(int n, std::string *err);  // n steady updates as one graph launch
    bool rebuild_due() const
    {
        return !grid_ && cfg_.needs_work() && (!work_valid_ || work_updates_ % work_period_ == 0);
    }
    int check(hipError_t e, const char *what, std::string *err);
    // After a stream synchronisation: a device-side failure of the updates
    // since the last check (the count scan's bounded wait gave up) marks every
    // scene's state invalid; returns SPH_ERR_DEVICE while scene s (any scene
    // when s < 0) is invalid.  Uploading or generating a scene's state makes
    // it valid again.
    int device_status(int s, std::string *err);
    std::vector<uint8_t> invalid_;  // per scene: state left by a failed update

    hipStream_t stream_ = nullptr;
    DevState d_{};
    LaunchCfg cfg_{};
    std::vector<SceneDesc> sd_h_;
    std::vector<sph_params> params_;
    EngineDesc *ed_h_ = nullptr;  // host copy of the engine descriptor
    // Descriptor uploads go through a ring of pinned staging slots (engine
    // descriptor, then the scene table), each reused only after the copies
    // from its previous use have run (its event), so an upload never waits
    // for the stream to drain.
    static constexpr int kDescRing = 4;
    unsigned char *desc_ring_[kDescRing] = {};
    hipEvent_t desc_ev_[kDescRing] = {};
    bool desc_live_[kDescRing] = {};
    uint32_t desc_no_ = 0;
    EngineDesc *d_ed_ = nullptr;
    SceneDesc *d_sd_ = nullptr;
    uint32_t *d_scene_of_ = nullptr;
    uint32_t scene_cap_ = 1;
    bool have_tables_ = false;
    bool grid_ = false;  // cell-grid bucket key (else the reference's hash % table)
    bool use_graph_ = true;
    // [rebuild work order? + 2 * fresh predict?]
    hipGraph_t graph_[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t graph_exec_[4] = {nullptr, nullptr, nullptr, nullptr};
    // run_len_ consecutive steady updates (buckets recorded, no work-order
    // rebuild) captured as one graph, for the substeps of a step
    hipGraph_t run_graph_ = nullptr;
    hipGraphExec_t run_exec_ = nullptr;
    int run_len_ = 0;
    // The last update's viscosity pass recorded the next update's buckets
    // (keys, tickets, counts) and nothing has changed since: state, settings,
    // layout and dt.  Otherwise the next update starts with predict.
    bool recorded_ = false;
    uint32_t work_period_ = 8;   // updates between work-order rebuilds
    uint64_t work_updates_ = 0;  // updates since creation
    bool work_valid_ = false;    // a work order has been built
    bool timing_ = false;
    hipEvent_t ev_[7] = {};
    double stage_ms_[6] = {0, 0, 0, 0, 0, 0};
    uint64_t device_bytes_ = 0;
    float dt_cached_ = -1.0f;
    float4 *colors_buf_ = nullptr;  // speed colours (allocated when first enabled)
    float4 *rpos_ = nullptr;        // render positions (allocated when first asked for)
    float4 *tmp_pos_ = nullptr;  // repack / reorder staging
    float4 *tmp_vel_ = nullptr;
    int staging(std::string *err);
    std::vector<hipEvent_t> prof_ev_;
    std::vector<uint8_t> prof_fresh_;  // per profiled update: predict ran (its events recorded)
    uint32_t prof_used_ = 0;
    bool prof_active_ = false;
};

}  // namespace sph
