// sph_launch.h -- host-callable launchers of the gfx950 kernels in
// sph_kernels.hip.  Every launcher enqueues on `stream` and returns the
// launch status; none of them synchronises or allocates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_generate.h"
#include "sph_types.h"

namespace sph {

// buckets per workgroup of the one-pass count scan: a chain status word and a
// ticket (one atomic on a single counter) per chunk
constexpr uint32_t kCountChunkBuckets = 4096;
// words of the fused update's grid barrier (counters on cache lines of their own)
constexpr uint32_t kFusedBarWords = 64 + 16 * 64;

struct DevState {
    // engine descriptors
    const EngineDesc *ed;
    const SceneDesc *sd;
    const uint32_t *scene_of;  // [n] scene of each slot (unused when single)
    bool multi;                // more than one scene
    bool grid;                 // cell-grid bucket key (SceneDesc::grid of every scene)
    // particle state in the resident order: slot k holds particle ids[k]
    // (the sorted order of the last update; any order gives the same results)
    float4 *pos;    // x,y,z, w = 0.34 after an update (OutPositions layout, cc:107)
    float4 *vel;    // x,y,z, w unused
    uint32_t *ids;  // particle index (engine-wide) of each resident slot
    // bucket sort: keys[0] bucket key and keys[1] ticket by resident slot
    // (recorded by the previous update's viscosity pass, or by predict);
    // order the (resident slot, particle index) pairs of each bucket in
    // ticket order
    uint32_t *keys[2];
    uint2 *order;
    uint32_t *count;   // [count_size = table_capacity + scene capacity] bucket counts, then
                       // scene flags: the next update's records, or zero
    uint32_t count_size;
    uint64_t *chain;   // [count_blocks + 1] chunk status words of the one-pass scan, then
                       // its chunk counter; zero between updates
    uint32_t count_blocks;
    // radix sort of the Morton work order
    uint32_t *hist;        // [radix * radix_blocks] counts, then [radix] digit totals
    uint32_t radix_blocks;
    // sorted order
    float4 *xps;   // predicted position, w = float(hash) (the reference record's .y)
    float4 *vs;    // velocity after stage A, sorted
    float4 *vs2;   // velocity after the pressure stage, sorted
    float4 *dens;  // rho, rho_near, RN(1/rho), RN(1/rho_near) (reciprocals 0 out of range)
    float4 *ps;    // position before the update, w = bits of the exact u32 hash
    uint32_t *sids;  // particle index of each sorted slot
    uint32_t *bounds;  // [table_total + 1] bucket b holds sorted slots [bounds[b], bounds[b+1])
    // work order of the gather kernels
    uint32_t *wkeys[2];  // Morton ordering keys (sort buffers)
    uint32_t *wvals[2];  // particle ids (sort buffers); the sorted ones are the work order
    const uint32_t *work;  // = wvals[result] of the last rebuild
    uint32_t *rank;      // [capacity] sorted slot of each particle id
    float4 *colors;      // [capacity] speed colours, particle order (null: not written)
    float4 *rpos;        // [capacity] render positions, particle order (null: not written)
    uint32_t *sflags;    // [scenes] grid key: 1 when a particle's cell lies outside the grid's box
                         // (copied from the count flags by the bucket sort)
    uint32_t *nbr;       // [capacity * 64] neighbour rows by work index (nbr mode)
    uint32_t *ncnt;      // [capacity] neighbour count per row
    uint32_t *scan_err;  // [1] raised when a bounded wait gave up (count scan, grid barrier)
    uint32_t *fbar;      // [kFusedBarWords] grid barrier of the fused update (zero between launches)
    uint32_t fused_grid; // workgroups of the fused update that the device holds resident at once
    uint32_t capacity;       // particle slots allocated
    uint32_t table_capacity; // bucket slots allocated
};

// Result-identical variants of a gather kernel (sph_kernels.hip explains them).
struct GatherVariant {
    bool work = false;  // Morton work order + XCD-contiguous workgroups
    bool list = false;  // two phases through a per-lane LDS list
    int block = 256;    // work-items per workgroup (64..256)
    bool rows = false;  // pressure / viscosity: read the density pass's neighbour rows
};

struct LaunchCfg {
    // defaults: the autotuner's usual picks on the benchmark scenes
    GatherVariant density{false, false, 64};
    GatherVariant pressure{true, true, 64};
    GatherVariant viscosity{false, false, 64};
    // nbr mode: density (in the work order) stores neighbour rows; pressure
    // and viscosity read them when their `rows` is set (then in the work order)
    bool nbr = false;
    // grid key, neighbour masks in all three gathers: a steady update (buckets
    // recorded) is one persistent launch with grid barriers between its
    // stages (launch_fused_update) instead of one launch per stage
    bool fused = false;
    int block_stream = 256;  // threads per block of streaming kernels
    bool needs_work() const { return nbr || density.work || pressure.work || viscosity.work; }
};

// Defaults per bucket key.  Under the cell-grid key the sorted order is
// already spatially coherent: the kernels run in it (no Morton work order),
// and the row mode records neighbour masks instead of neighbour rows.
inline LaunchCfg sanitize_launch_cfg(LaunchCfg c, bool grid)
{
    if (grid)
        for (GatherVariant *v : {&c.density, &c.pressure, &c.viscosity}) v->work = false;
    return c;
}

inline LaunchCfg default_launch_cfg(bool grid)
{
    LaunchCfg c;
    if (grid) {
        // the autotuner's usual picks: neighbour masks for all three kernels
        c.density = GatherVariant{false, false, 256, true};
        c.pressure = GatherVariant{false, false, 256, true};
        c.viscosity = GatherVariant{false, false, 128, true};
        c.nbr = true;
    }
    return sanitize_launch_cfg(c, grid);
}

// Stage A + B1 when the previous update did not record the buckets: clears
// the counts, then predict, cell, hash, bucket key and ticket of every particle.
hipError_t launch_predict(const DevState &s, const LaunchCfg &c, hipStream_t st);
// A steady update (buckets recorded by the previous one) as one launch of
// s.fused_grid persistent workgroups: count scan, scatter, reorder, density,
// pressure, viscosity+integrate, separated by grid barriers.  Grid key with
// neighbour masks only (c.fused; see LaunchCfg).
hipError_t launch_fused_update(const DevState &s, hipStream_t st);
// Workgroups of the fused update that fit on the device at once (its
// occupancy x compute units, a multiple of 8), or 0 when none does.
uint32_t fused_update_grid(bool multi, int device);
// Stages B2 + B3: counting sort into (bucket, particle index) order, bucket
// bounds, gather of xp, v into sorted order; clears the counts.
hipError_t launch_bucket_sort(const DevState &s, const LaunchCfg &c, hipStream_t st);
// Stable LSD radix sort of (keys, vals) over `bits` key bits (Morton work
// order); the input values are first_vals (null: the input slot indices).
// Returns the index (0/1) of the buffer pair holding the result.
hipError_t launch_radix_sort(const DevState &s, uint32_t *const keys[2], uint32_t *const vals[2],
                             const uint32_t *first_vals, uint32_t bits, hipStream_t st,
                             int *result_buf);
// Stage C, D, E+F.
hipError_t launch_density(const DevState &s, const LaunchCfg &c, hipStream_t st);
hipError_t launch_pressure(const DevState &s, const LaunchCfg &c, hipStream_t st);
hipError_t launch_viscosity_integrate(const DevState &s, const LaunchCfg &c, hipStream_t st);

// Sorted -> particle order densities (rho, rho_near) for download.
hipError_t launch_unpermute_dens(const DevState &s, float2 *out, hipStream_t st);
// Resident slots [off, off + n) of one scene (particles off .. off + n - 1)
// into particle order: dst_pos[i] / dst_vel[i] = state of particle off + i.
hipError_t launch_unpermute_state(const DevState &s, uint32_t off, uint32_t n, float4 *dst_pos,
                                  float4 *dst_vel, hipStream_t st);
// a[k] = first + k for k < n.
hipError_t launch_iota(uint32_t *a, uint32_t n, uint32_t first, hipStream_t st);
// scene_of of every slot of the first n_scenes scenes, from the device scene
// table (max_n: the largest scene's particle count)
hipError_t launch_scene_map(const SceneDesc *sd, uint32_t *scene_of, uint32_t n_scenes, uint32_t max_n,
                            hipStream_t st);
// a[k] = 0 for k < n, by a fill kernel (no memset node when captured).
hipError_t launch_clear_words(uint32_t *a, uint32_t n, hipStream_t st);
// Neighbour lists for parity checks (particle order, walk order, self included).
hipError_t launch_neighbour_count(const DevState &s, uint32_t *counts, hipStream_t st);
hipError_t launch_neighbour_fill(const DevState &s, const uint64_t *offs,
                                 uint32_t *list, hipStream_t st);
// Initial state of one scene from its plan (sph_generate.h), n = g.n slots
// starting at pos / vel.
hipError_t launch_generate(const GenPlan &g, float4 *pos, float4 *vel, hipStream_t st);
// Count of floats with bit patterns in [lo, hi] whose kernel reciprocal
// (recip_rn) differs from IEEE 1/x, and the smallest such pattern (atomic
// min into *first_bad, which the caller sets first).
hipError_t launch_recip_check(uint32_t lo, uint32_t hi, unsigned long long *bad,
                              uint32_t *first_bad, hipStream_t st);
// Plain copy of n float4 (HBM counter calibration).
hipError_t launch_copy(const float4 *src, float4 *dst, size_t n, hipStream_t st);
// Count of non-finite position/velocity components (error flag).
hipError_t launch_count_nonfinite(const DevState &s, uint32_t *counter, hipStream_t st);

}  // namespace sph
