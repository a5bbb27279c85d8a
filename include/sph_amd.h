/*
 * sph_amd.h -- C ABI of the MI355X (gfx950) SPH step.
 *
 * Drop-in for the SPH step of Allkams/Fluid-Simulation-3D: the singleton
 * Physics::Fluid::FluidSimulation (engine/physics/physicsWorld.h:28-155,
 * physicsWorld.cc:32-558) and its FluidSimBase backend slot
 * (projects/Simulation/code/simulations/fluidSimBase.h:5-13).
 *
 * Conventions
 *   - Every function returns SPH_OK (0) or a negative SPH_ERR_* code; the
 *     message of the last failure on the calling thread is sph_last_error().
 *     No exception crosses the ABI.
 *   - Handles are independent (no singleton); each owns its device buffers and
 *     one HIP stream.  Calls on one handle must be serialised by the caller;
 *     different handles may be used from different threads.
 *   - Host buffers are caller-owned.  Particle arrays are float4 per particle
 *     (x, y, z, w): pos4 / vel4 hold 4*n floats.  After an update pos4.w is
 *     0.34, so pos4 is also the reference's OutPositions (physicsWorld.h:80,
 *     physicsWorld.cc:107).
 *   - Results follow the reference's FP32 operation order, with two stated
 *     differences: the sort is stable (equal keys keep particle-index order,
 *     where std::sort at physicsWorld.cc:484 leaves it unspecified) and the
 *     viscosity stage reads a velocity snapshot (Jacobi) where the reference
 *     reads velocity[j] in place (physicsWorld.cc:460).
 */
#ifndef SPH_AMD_H
#define SPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_OK 0
#define SPH_ERR_INVALID (-1)   /* bad argument or settings */
#define SPH_ERR_HIP (-2)       /* HIP runtime failure (message has the HIP error) */
#define SPH_ERR_OOM (-3)       /* device or host allocation failed */
#define SPH_ERR_BUDGET (-4)    /* request can never fit the server's HBM budget */
#define SPH_ERR_STATE (-5)     /* call not valid in the current state */
#define SPH_ERR_NOT_FOUND (-6) /* unknown request id */
#define SPH_ERR_NONFINITE (-7) /* state holds NaN/Inf */
#define SPH_ERR_DEVICE (-8)    /* a bounded device-side wait gave up during an
                                  update: its results, and the state of every
                                  scene it advanced, are invalid until new
                                  state is uploaded or generated */

/* Simulation settings.  Field <-> reference member (physicsWorld.h):
 *   smoothing_radius   interactionRadius       :97  (setInteractionRadius :58)
 *   target_density     TargetDensity           :98  (setDensityTarget :61)
 *   pressure_mult      pressureMultiplier      :99  (setPressureMultiplier :64)
 *   near_pressure_mult nearPressureMultiplier  :100 (setNearPressureMultiplier :67)
 *   viscosity          viscosityStrength       :101 (setViscosityStrength :70)
 *   gravity_on         gravity                 :105 (setGravity :55)
 *   gravity_scale      gravityScale            :106 (setGravityScale :73)
 *   bounds             BoundScale              :145 (setBound :76)
 *   lookahead          1.0f/120.0f             physicsWorld.cc:47 (fixed in the reference)
 *   damp               dampFactor 0.95f        physicsWorld.cc:87 (fixed in the reference)
 *   prefilter_r2       sqrRadius 0.35f*0.35f   :96  (const in the reference) */
typedef struct sph_params {
    float smoothing_radius;
    float target_density;
    float pressure_mult;
    float near_pressure_mult;
    float viscosity;
    int32_t gravity_on;
    float gravity_scale;
    float bounds[3];
    float lookahead;
    float damp;
    float prefilter_r2;
} sph_params;

typedef struct sph_sim *sph_handle;

/* The reference's in-class defaults (physicsWorld.h:96-106, 145). */
void sph_default_params(sph_params *p);
/* Message of the last failure on this thread ("" if none). */
const char *sph_last_error(void);
const char *sph_version(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int sph_device_count(void);

/* Measurement support, no reference counterpart: `reps` launches of a plain
 * float4 copy kernel (k_copy) over `bytes` bytes; *ms_per_copy gets the mean
 * duration.  Its byte counters calibrate per-dispatch FETCH_SIZE/WRITE_SIZE. */
int sph_copy_calibration(uint64_t bytes, int reps, double *ms_per_copy);

/* ---- single scene: FluidSimulation (physicsWorld.h:28-155) ---- */

/* Bucket keys of the spatial lookup (UpdateSpatialLookup, cc:466-498).  The
 * results of every update are bit-identical under either key; only the lookup
 * tables (sph_debug_spatial) and the speed differ.
 *   SPH_TABLE_GRID (0): the cell-grid key -- buckets are the cells of a grid
 *     over the box (+4 cells per side), z fastest, so a neighbour row of three
 *     cells is one contiguous run of sorted slots; cells outside the grid
 *     wrap around it.  The bucket count is chosen from bounds and h.
 *   SPH_TABLE_REFERENCE: the reference's key hash % n (GetKeyFromHash with
 *     tableSize = numParticles, cc:479, 513-516).
 *   any other T: the reference's key with T buckets, hash % T. */
#define SPH_TABLE_GRID 0u
#define SPH_TABLE_REFERENCE 0xffffffffu

/* FluidSimulation::getInstance + InitializeData's sizing (h:32,36; cc:112-137):
 * a scene of n particles; table_size selects the bucket key (above).
 * All state starts at zero. */
int sph_create(const sph_params *p, uint32_t n, uint32_t table_size, sph_handle *out);
int sph_destroy(sph_handle h);
/* The setters / getters of physicsWorld.h:52-77 (take effect at the next update). */
int sph_set_params(sph_handle h, const sph_params *p);
int sph_get_params(sph_handle h, sph_params *p);
int sph_get_size(sph_handle h, uint32_t *n, uint32_t *table_size);
/* Bucket key in use: *grid = 1 for the cell-grid key, with its periods
 * gdim[3] and origin cell lo[3] (key = (wx*Gy + wy)*Gz + wz, w = (u32(cell) -
 * u32(lo)) mod G per axis); *grid = 0 for hash % table_size.  Any output may
 * be NULL.  No reference counterpart (parity tests). */
int sph_get_key(sph_handle h, int32_t *grid, uint32_t gdim[3], int32_t lo[3]);
/* Writes the positions vector (h:79) and the velocities; either may be NULL
 * (left unchanged).  Host float4 arrays of n entries. */
int sph_upload(sph_handle h, const float *pos4, const float *vel4);
/* InitializeData(n) (cc:112-147): the reference's GridArrangement lattice
 * (cc:518-557), zero velocity, then the spatial lookup and densities of that
 * state (cc:144-145). */
int sph_initialize_data(sph_handle h);
/* InitializeData for any synthetic scene (SPH_INIT_*, below), generated on
 * the device -- no host arrays cross PCIe -- with exactly the bytes
 * sph_generate writes for the same arguments; then the spatial lookup and
 * densities of that state (cc:144-145).  sph_initialize_data is this with
 * SPH_INIT_GRID_ARRANGEMENT (the reference's lattice, cc:518-557). */
int sph_initialize(sph_handle h, int kind, uint64_t seed, const float box[3], float spacing,
                   float jitter, float speed);
/* Render hand-off (FluidSimCPU::updateColors, fluidSimCPU.cc:100-125; the
 * GLSL path's UpdateColors, compute_particle.glsl:279-304): when enabled,
 * every update also writes each particle's speed colour -- the 4-stop
 * gradient of clamp(|v|, 0, 1.5) / 1.5 (getSpeedNormalzied,
 * physicsWorld.cc:178-182) -- fused into the integration kernel.  stops:
 * Color1..Color4 as RGBA (16 floats; NULL keeps the current ones, initially
 * fluidSimCPU.h:26-29). */
int sph_set_colors(sph_handle h, const float stops[16], int enable);
/* The colours of the last update, n float4 RGBA in particle order. */
int sph_download_colors(sph_handle h, float *rgba4);
/* Device pointers of the render buffers, laid out as the reference's shader
 * storage blocks (std430, fluidSimGPU.cc:98-104; particleRender.vs reads
 * bindings 0 and 1): positions = binding 0 ReadPosAndScale (vec4 x, y, z,
 * 0.34 -- OutPositions), colors = binding 1 ReadColors (vec4 RGBA; NULL
 * until sph_set_colors enables them).  n entries each, particle order,
 * valid until sph_destroy and rewritten by every update, so a renderer can
 * copy them into (or map them as) its SSBOs with no host round trip.  The
 * simulation keeps its own state in a spatially sorted order; the position
 * buffer is a particle-order copy, created and filled by the first call and
 * written by every update from then on. */
int sph_render_buffers(sph_handle h, const void **positions, const void **colors);
/* Update(deltatime) (cc:39-111), blocking. */
int sph_update(sph_handle h, float dt);
/* `substeps` calls of Update(dt / substeps); blocking. */
int sph_step(sph_handle h, float dt, int substeps);
/* As sph_step, but returns once the work is enqueued on the handle's stream. */
int sph_step_async(sph_handle h, float dt, int substeps);
int sph_sync(sph_handle h);
/* Reads positions (= OutPositions, w = 0.34), velocities and the (density,
 * near density) pairs of the last update, all in particle-index order; any
 * pointer may be NULL.  Replaces getPosition/getVelocity/getDensity/
 * getNearDensity (h:38-41, cc:149-170) and the OutPositions vector (h:80). */
int sph_download(sph_handle h, float *pos4, float *vel4, float *dens2);
/* The six stage timers of the last update in ms, physicsWorld.h:45-50 order:
 * gravity, spatial, density, pressure, viscosity, position+collision.
 * Filled only while timing is on (sph_set_timing); the viscosity and
 * position+collision stages run as one kernel, reported under viscosity
 * (position+collision reads 0). */
int sph_set_timing(sph_handle h, int on);
int sph_stage_ms(sph_handle h, double ms[6]);
/* Per-stage device timing of the next `max_updates` updates (up to 2^20)
 * without synchronising: timing-only events bracket each stage on the
 * handle's stream.  Within a step of several substeps (sph_step) only the
 * first update is profiled; the others run unprofiled as one graph launch.
 * sph_profile_end waits, returns the mean ms per update of each stage (order
 * of sph_stage_ms; the kernels are: 0 predict+hash (only when the previous
 * update did not record this one's buckets, else empty), 1 count scan +
 * scatter + reorder, 2 density, 3 pressure, 4 viscosity+integrate, which
 * also records the next update's buckets) and the number of
 * updates profiled.  No reference counterpart: measurement support for the
 * stage timers of physicsWorld.h:45-50. */
int sph_profile_begin(sph_handle h, uint32_t max_updates);
int sph_profile_end(sph_handle h, double mean_ms[6], uint32_t *updates);
/* Autotuner (no reference counterpart): times every variant of the three
 * gather kernels (work order on/off under the reference key, LDS neighbour
 * list on/off, workgroup size 64/128/256, and the row mode) on the current
 * state with `updates` Update(dt) calls per candidate, twice over the list of
 * candidates, restoring positions and velocities after each, and keeps the
 * fastest variant per kernel (its fastest run).  All
 * variants give bit-identical results.  ms (may be NULL) gets the chosen
 * variants' mean ms (density, pressure, viscosity).  Densities and debug
 * tables are recomputed by the next update.  Without it a handle runs the
 * defaults (under the grid key: neighbour masks for all three kernels). */
int sph_autotune(sph_handle h, float dt, int updates, double ms[3]);
/* The variant of each gather kernel, q = 0 density, 1 pressure, 2 viscosity:
 * v[4q+0] work order (reference key only; ignored under the grid key),
 * v[4q+1] LDS list, v[4q+2] workgroup size (64, 128, 192, 256), v[4q+3] read
 * the row-mode records (pressure, viscosity); v[12] row mode: the density
 * pass records each particle's neighbours -- 64-entry neighbour rows under
 * the reference key, the neighbour-mask record under the grid key; v[13]
 * fused: under the grid key with the row mode in all three kernels, each
 * steady update (buckets recorded by the previous one) is one persistent
 * launch with grid barriers between its stages instead of one launch per
 * stage (profiled updates keep the separate launches; a grid barrier's
 * bounded wait that gives up is reported as SPH_ERR_DEVICE).  The array
 * has 14 entries. */
int sph_get_variants(sph_handle h, int32_t v[14]);
int sph_set_variants(sph_handle h, const int32_t v[14]);
/* Test support, no reference counterpart: the bound of the count scan's
 * look-back wait (the prefix sum behind the bucket table of UpdateSpatialLookup,
 * cc:486-496), in sleeps of 64 clocks; 65536 by default, about 1.7 ms.  A
 * workgroup waits there only for chunks of the scan taken before its own by
 * workgroups already running, so a wait that reaches the bound means the
 * device was held up elsewhere (time-sliced with other work): the update's
 * bucket table is then incomplete, the next synchronising call (sph_sync,
 * sph_step, sph_update, sph_download, the debug readbacks) returns
 * SPH_ERR_DEVICE, and the state stays invalid until sph_upload (positions and
 * velocities) or sph_initialize replaces it.  0 makes every wait give up at
 * once, which forces that path from the host. */
int sph_debug_set_scan_spin(sph_handle h, uint32_t sleeps);
/* Number of non-finite position/velocity components in the current state. */
int sph_count_nonfinite(sph_handle h, uint64_t *count);

/* ---- one Update without allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops and harnesses that forbid hipMalloc): one Update(dt)
 * (cc:39-111) of one scene of n particles held in the caller's device memory,
 * as sph_create(SPH_TABLE_GRID) + sph_upload + sph_update + sph_download would
 * give it, bit for bit, with no device or host allocation: every intermediate
 * lives in a fixed device array of the library (SPH_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends).
 *   pos4_in, vel4_in    device float4 arrays of n particles, particle order
 *                       (read only)
 *   pos4_out, vel4_out  device float4 arrays of n: the state after the
 *                       Update (pos4.w = 0.34, OutPositions)
 *   dens2_out           device float2 array of n (density, near density), or NULL
 *   stream              hipStream_t to enqueue on (NULL: the null stream)
 *   wait                1: return once done, reporting a device-side failure
 *                       (SPH_ERR_DEVICE, as the handle path does); 0: return
 *                       once enqueued (a failure is reported by the next
 *                       waiting call)
 * n must be in [1, SPH_NOALLOC_MAX_PARTICLES] (SPH_ERR_INVALID otherwise).
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (from several threads, or on streams not ordered with each other) race. */
#define SPH_NOALLOC_MAX_PARTICLES 65536u
#define SPH_NOALLOC_SCRATCH_BYTES 24645376u
int sph_update_noalloc(const sph_params *p, uint32_t n, float dt, const float *pos4_in,
                       const float *vel4_in, float *pos4_out, float *vel4_out, float *dens2_out,
                       void *stream, int wait);

/* ---- parity / debugging: the spatial lookup of the last update ----
 * UpdateSpatialLookup (cc:466-498): per particle hash (HashCell) and key
 * (GetKeyFromHash); per sorted slot key and particle index; per bucket the
 * first sorted slot (startIndices, INT_MAX when empty) and one past the last.
 * Keys are the bucket key in use (sph_get_key).  Any pointer may be NULL.
 * hash_id/key_id/sorted_* hold n entries, start/end table_size entries. */
int sph_debug_spatial(sph_handle h, uint32_t *hash_id, uint32_t *key_id,
                      uint32_t *sorted_key, uint32_t *sorted_idx, uint32_t *start,
                      uint32_t *end);
/* predictedPositions of the last update (physicsWorld.h:118), float4, w = 0,
 * particle order; SPH_ERR_STATE before the first update. */
int sph_debug_predicted(sph_handle h, float *xp4);
/* The velocity vector of the last update after its external-force stage
 * (cc:45) and after its pressure stage (cc:62-69, 421), before viscosity:
 * float4 (w unused) in particle order; either pointer may be NULL.  So the
 * pressure term of each particle is v_press - v_ext and the viscosity term
 * (cc:463) is the downloaded velocity - v_press, for particles that did not
 * touch a wall (cc:90-105).  SPH_ERR_STATE before the first update. */
int sph_debug_velocities(sph_handle h, float *v_ext4, float *v_press4);
/* Neighbour sets of the last update as the reference's bucket walk sees them
 * (cc:331-363): for particle i, every j in the 27 cells whose hash matches and
 * whose squared distance is <= prefilter_r2, self included, in walk order.
 * CSR: offs[n+1]; list gets min(total, cap) entries; *total = offs[n]. */
int sph_debug_neighbours(sph_handle h, uint64_t *offs, uint32_t *list, uint64_t cap,
                         uint64_t *total);

/* ---- the spatial hash as the kernels compute it (host copies, for callers
 * and checks): HashCell (cc:505-511) on integer cell coordinates and
 * GetKeyFromHash (cc:513-516), i.e. hash % table_size. ---- */
uint32_t sph_hash_cell(int32_t cx, int32_t cy, int32_t cz);
uint32_t sph_bucket_of(uint32_t hash, uint32_t table_size);
/* Parity support, no reference counterpart: out[i] = a[i] / b[i] as the
 * pressure kernel computes it (the reciprocal-and-correction quotient where
 * its range conditions hold, IEEE division elsewhere), on the host from the
 * same source.  Equal to IEEE a / b except that -0 / b gives +0. */
int sph_debug_quotients(const float *a, const float *b, float *out, uint64_t n);
/* Parity support, no reference counterpart: on the GPU, counts the floats x
 * with bit patterns in [lo_bits, hi_bits] whose kernel reciprocal (the
 * estimate-and-correction RN(1/x) behind the pressure kernel's quotients)
 * differs from IEEE 1/x; *first (may be NULL) gets the smallest such bit
 * pattern, 0xffffffff when there is none. */
int sph_debug_reciprocal_mismatches(uint32_t lo_bits, uint32_t hi_bits, uint64_t *count,
                                    uint32_t *first);

/* ---- batching: many independent scenes per launch ----
 * No reference counterpart (the reference simulates one scene per process,
 * physicsWorld.cc:32-37).  A batch owns n_scenes scenes of uneven sizes, each
 * with its own settings and table size, packed into one set of device
 * arrays; every kernel of an update is ONE launch over all scenes (bucket keys
 * are offset per scene so scenes never interact), and the update is replayed
 * from a captured HIP graph.  Each scene's results are bit-identical to the
 * same scene run alone through sph_create/sph_update.
 * counts[s] particles and params[s] settings.  table_sizes NULL: every scene
 * uses the cell-grid key; else the reference key with table_sizes[s] buckets
 * (an entry 0 or SPH_TABLE_REFERENCE = counts[s]). */
typedef struct sph_batch_t *sph_batch;
int sph_batch_create(uint32_t n_scenes, const uint32_t *counts, const uint32_t *table_sizes,
                     const sph_params *params, sph_batch *out);
int sph_batch_destroy(sph_batch b);
int sph_batch_upload(sph_batch b, uint32_t scene, const float *pos4, const float *vel4);
/* Scene `scene`'s initial state generated on the device (as sph_initialize,
 * without the density pass; the next step computes it). */
int sph_batch_initialize(sph_batch b, uint32_t scene, int kind, uint64_t seed, const float box[3],
                         float spacing, float jitter, float speed);
int sph_batch_set_params(sph_batch b, uint32_t scene, const sph_params *p);
/* Every scene: `substeps` Update(dt / substeps) calls; blocking / async. */
int sph_batch_step(sph_batch b, float dt, int substeps);
int sph_batch_step_async(sph_batch b, float dt, int substeps);
int sph_batch_sync(sph_batch b);
int sph_batch_download(sph_batch b, uint32_t scene, float *pos4, float *vel4, float *dens2);
/* As sph_autotune / sph_profile_begin / sph_profile_end, over the whole batch. */
int sph_batch_autotune(sph_batch b, float dt, int updates, double ms[3]);
int sph_batch_profile_begin(sph_batch b, uint32_t max_updates);
int sph_batch_profile_end(sph_batch b, double mean_ms[6], uint32_t *updates);
int sph_batch_get_variants(sph_batch b, int32_t v[14]);
int sph_batch_set_variants(sph_batch b, const int32_t v[14]);

/* ---- serving: continuous batching of bake requests ----
 * No reference counterpart.  A request is a scene (settings, particle count,
 * bucket count, initial state) to advance by `steps` steps of `substeps`
 * Update(dt / substeps) calls.  Requests wait in arrival order and are
 * admitted while one of max_running slots is free and the device bytes of
 * the running requests (sph_server_request_bytes each) stay within
 * hbm_budget_bytes; a request larger than the budget on its own is refused
 * at submission (SPH_ERR_BUDGET).  Each sph_server_tick admits, advances
 * every running request by updates_per_tick Update calls -- one launch per
 * kernel for all of them -- and retires the finished ones, whose results are
 * kept in host memory until fetched.  Each result is bit-identical to the
 * same scene run alone.  Every request uses the cell-grid key; its
 * table_size is the bucket count (0: sized from bounds and h, as
 * SPH_TABLE_GRID; a smaller count wraps the grid, results are unchanged).
 * Calls on one server must be serialised by the caller. */
#define SPH_REQ_QUEUED 0
#define SPH_REQ_RUNNING 1
#define SPH_REQ_DONE 2
#define SPH_REQ_CANCELLED 3

typedef struct sph_server_config {
    uint64_t hbm_budget_bytes; /* device bytes of running requests (default 200 GiB) */
    uint32_t max_running;      /* slots: requests advanced together (default 64) */
    uint32_t initial_capacity; /* particle slots allocated at start; grows by doubling */
    float dt;                  /* frame dt of one step (default 1/120) */
    int32_t substeps;          /* Update calls per step (default 3) */
    int32_t updates_per_tick;  /* Update calls per tick (default 1: re-batch every substep) */
    uint32_t max_capacity;     /* particle slots the device layout may grow to (0: 2^31).
                                  Admissions that need more -- or that the device has no
                                  memory for -- wait at the head of the queue while the
                                  running requests go on (stats.admission_failures counts
                                  them); with nothing running the tick fails SPH_ERR_OOM */
} sph_server_config;

typedef struct sph_server_tick_info {
    uint32_t running, queued, admitted, finished;
    uint64_t particles; /* particles advanced by this tick */
    double ms;          /* wall time of the tick */
} sph_server_tick_info;

typedef struct sph_server_stats {
    uint64_t admitted, finished, rejected, cancelled, queued, running;
    uint64_t used_bytes, particle_updates, ticks, capacity;
    double busy_s; /* wall time of ticks that advanced requests */
    double tick_ms_p50, tick_ms_p99, request_ms_p50, request_ms_p99, wait_ms_p50, wait_ms_p99;
    uint64_t admission_failures; /* ticks whose admissions could not grow the layout (they waited) */
} sph_server_stats;

typedef struct sph_server_t *sph_server;
void sph_server_default_config(sph_server_config *c);
/* Device bytes a request holds while running (its HBM-budget charge). */
uint64_t sph_server_request_bytes(const sph_params *p, uint32_t n, uint32_t table_size);
int sph_server_create(const sph_server_config *cfg, sph_server *out);
int sph_server_destroy(sph_server s);
int sph_server_submit(sph_server s, const sph_params *p, uint32_t n, uint32_t table_size,
                      const float *pos4, const float *vel4, uint32_t steps, uint64_t *id);
/* A bake request whose initial state is generated on the device at admission
 * (SPH_INIT_* arguments as sph_generate, identical bytes): nothing is held
 * in host memory while it waits. */
int sph_server_submit_generated(sph_server s, const sph_params *p, uint32_t n,
                                uint32_t table_size, int kind, uint64_t seed, const float box[3],
                                float spacing, float jitter, float speed, uint32_t steps,
                                uint64_t *id);
int sph_server_tick(sph_server s, sph_server_tick_info *info);
int sph_server_status(sph_server s, uint64_t id, int32_t *state, uint64_t *updates_done,
                      double *latency_ms);
/* Copies a finished request's positions, velocities and densities (any may be
 * NULL; n float4, float4 and float2 entries) and forgets the request.  n must
 * be the request's particle count (SPH_ERR_INVALID otherwise, nothing
 * copied). */
int sph_server_fetch(sph_server s, uint64_t id, uint32_t n, float *pos4, float *vel4, float *dens2);
int sph_server_cancel(sph_server s, uint64_t id);
int sph_server_get_stats(sph_server s, sph_server_stats *st);
/* The per-tick log behind the tick percentiles: for the timed ticks
 * first .. first + cap - 1 (in the order they ran; sph_server_get_stats
 * finishes timing the enqueued ones), each tick's GPU time, the particles it
 * advanced and its layout number (the count of membership changes before it,
 * so ticks of one number advanced the same requests).  Any output may be
 * NULL; *count gets the number of entries written.  The server keeps every
 * tick since its creation (28 bytes each, with the tick percentiles'
 * samples) until it is destroyed. */
int sph_server_tick_log(sph_server s, uint64_t first, uint64_t cap, float *ms, uint64_t *particles,
                        uint32_t *layout, uint64_t *count);

/* ---- synthetic scenes: sph_generate writes host arrays; sph_initialize,
 * sph_batch_initialize and sph_server_submit_generated generate the same
 * bytes on the device (one per-particle function for both) ---- */
#define SPH_INIT_GRID_ARRANGEMENT 0 /* the reference's InitializeData lattice, cc:518-557 */
#define SPH_INIT_DAM_BREAK 1        /* block at the -x,-y,-z corner, dims nx:2nx:2nx, z fastest */
#define SPH_INIT_UNIFORM_LOWER 2    /* uniform in the lower half of the box */
#define SPH_INIT_TWO_BLOCKS 3       /* two blocks moving at +-speed along x */
#define SPH_INIT_CENTRED_BLOCK 4    /* cubic block centred in the box */
/* Writes n float4 positions (w = 0.34) and velocities (w = 0).  box: full
 * extent (the box is centred on the origin, as cc:88-106 assumes); spacing:
 * lattice spacing; jitter: half-width of the uniform per-component jitter;
 * speed: |v| of SPH_INIT_TWO_BLOCKS.  A counter-based generator (splitmix64 of
 * seed and particle index) makes the bytes independent of threading. */
int sph_generate(int kind, uint64_t seed, uint32_t n, const float box[3], float spacing,
                 float jitter, float speed, float *pos4, float *vel4);

#ifdef __cplusplus
}
#endif
#endif
