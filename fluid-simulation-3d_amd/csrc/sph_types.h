// sph_types.h -- data structures shared by host code and the gfx950 kernels.
//
// One "engine" holds S >= 1 independent scenes packed into flat SoA arrays.
//
// Bucket keys.  The reference buckets particles by hash % T (cc:513-516) and
// walks, for each of the 27 neighbour cells, that cell's bucket, keeping the
// records whose float-rounded hash equals the cell's.  A particle within h of
// particle i always lies in one of i's 27 cells, and no two of those cells
// have float-equal hashes: two of them differ by (dx, dy, dz) in [-2, 2]^3,
// so their hashes differ by dx*15823 + dy*9737333 + dz*440817757 mod 2^32,
// which is at least 15823 away from 0 mod 2^32, while u32 values that round
// to the same float are less than 256 apart.  Each sum of the reference is
// therefore: for each offset in the reference order, the particles of that
// cell in ascending index (the stable sort).  That holds for ANY key that is
// a function of the cell.  The engine
// can therefore use either the reference key or the cell-grid key
// ((wx * Gy + wy) * Gz + wz with w = (c - cell_lo) mod G per axis), under
// which the three cells dz = -1, 0, 1 of a neighbour row are consecutive
// buckets, i.e. one contiguous run of sorted slots.  Results are identical.
// Scene s owns particle slots [p_off, p_off + n) -- the same range in id order
// and in sorted order, because every key of scene s lies in
// [key_off, key_off + table) and the key ranges of scenes are disjoint and
// increasing with s.  A single-scene handle is an engine with S = 1.
#pragma once
#include <stdint.h>

#define SPH_START_EMPTY 0x7fffffffu  // INT_MAX: empty bucket (physicsWorld.cc:481)

// Per-scene constants, resident in device memory and read by every kernel.
// Everything derived from the user settings (kernel normalisations, the
// remainder constant, the d < h threshold on d^2) is computed once on the host
// with the same float expressions as the reference (kernels.h:25-82).
struct SceneDesc {
    uint32_t p_off;       // first particle slot
    uint32_t n;           // particle count
    uint32_t key_off;     // first bucket in the engine-wide bucket tables
    uint32_t table;       // bucket count T (reference: numParticles, cc:479)
    uint64_t mod_m;       // Lemire remainder constant: floor((2^64-1)/table)+1
    float h;              // interactionRadius
    float r2;             // distance prefilter sqrRadius (physicsWorld.h:96)
    float h2_lt;          // smallest float x with sqrtf(x) >= h: d < h  <=>  d2 < h2_lt
    float thr;            // min(h2_lt, next float above r2): d2 <= r2 && d < h  <=>  d2 < thr
    float rho0, k, k_near, mu;
    float w2c, w3c;       // SmoothingPow2/Pow3 volumes (kernels.h:29, 41)
    float dw2c, dw3c;     // derivative scales (kernels.h:53, 65)
    float poly6c;         // viscosity poly6 scale (kernels.h:77)
    float hh;             // h*h (kernels.h:78: radius * radius)
    float g[3];           // external acceleration (cc:313-323)
    float lookahead;      // 1/120 (cc:47)
    float half[3];        // BoundScale * 0.5f (cc:88)
    float damp_neg;       // -1 * dampFactor (cc:94)
    float dt;             // substep dt of the current Update
    int32_t cell_lo[3];   // origin of the scene's cell grid (grid key, work-order key)
    uint32_t grid;        // bucket key: 1 = cell-grid key, 0 = hash % table (cc:513-516)
    uint32_t gdim[3];     // grid key: period of the cell grid per axis, product <= table
    uint32_t index;       // this scene's position in the layout (per-scene flags)
    uint64_t gmod[3];     // grid key: remainder constants of gdim
};

// Whole-engine counters, resident in device memory so that a captured graph
// stays valid while the packed layout changes (continuous batching).
struct EngineDesc {
    uint32_t n_total;      // particles over all scenes
    uint32_t table_total;  // buckets over all scenes
    uint32_t n_scenes;
    uint32_t scene_shift;  // work-order key: scene index << scene_shift (32: one scene)
    uint32_t morton_drop;  // work-order key: Morton code >> morton_drop
    uint32_t scan_spin;    // bound of the count scan's look-back wait, in 64-clock sleeps
    uint32_t pad[2];
    float stops[16];       // speed-colour gradient, Color1..Color4 RGBA (fluidSimCPU.h:26-29)
};
