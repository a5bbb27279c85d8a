// sph_kernels.hip -- hand-written gfx950 kernels of the SPH step.
//
// Reference: Allkams/Fluid-Simulation-3D, engine/physics/physicsWorld.cc
// (FluidSimulation::Update, cc:39-111) and engine/physics/kernels.h.
//
// Pipeline of one Update(dt) over all scenes of an engine:
//   (k_count_clear, k_predict)  stage A + B1 (cc:42-48, 473-482) in resident
//                      order: each particle's bucket and a ticket in it.  Only
//                      when the previous update did not record them already
//   k_count_scan/scatter/reorder  stages B2 + B3 (cc:481-496): a
//                      counting sort over the buckets, ordered within a bucket
//                      by particle index (the stable (key, index) order);
//                      bucket bounds; stage A recomputed into sorted order
//   k_density          stage C (cc:304-311, 325-365)
//   k_pressure         stage D (cc:62-69, 367-422)
//   k_viscosity_integrate  stage E (cc:71-78, 424-464, Jacobi) + F (cc:80-108),
//                      the new state written in sorted order (the next
//                      resident order), and stage A + B1 of the next update
//                      recorded from it
//
// Every floating-point expression keeps the reference's operation order and
// rounding (the library is compiled with -ffp-contract=off and IEEE division
// and square root), so results match the reference's FP32 arithmetic; only the
// viscosity stage differs by design (Jacobi: it reads the velocity snapshot
// taken after the pressure stage, where the reference reads velocity[j] in
// place).
//
// Layout: SoA float4 arrays, one 16-byte load per particle per array, wave64
// coalesced.  Gather kernels run one work-item per particle in SORTED order, so
// the 64 lanes of a wave cover a handful of cells and read the same neighbour
// buckets, which stay resident in the per-XCD L2.  The particle state (pos,
// vel) stays in the sorted order of the last update (the resident order;
// ids[k] is the particle index held by slot k): particles move little per
// update, so the next sort's gathers and scatters are nearly sequential.
// Results never depend on the resident order: the sort orders by (bucket,
// particle index) whatever order its input is in.
#include <hip/hip_runtime.h>

#include "sph_device.h"
#include "sph_launch.h"

namespace sph {

namespace {

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortBlock = 256;  // 4 waves
constexpr int kSortWaves = kSortBlock / 64;

__device__ __forceinline__ uint32_t grid_stride() { return gridDim.x * blockDim.x; }
__device__ __forceinline__ uint32_t global_id() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Scene q's descriptor for a wave-uniform q, read through the constant
// address space: the table does not change while a kernel runs, so the reads
// are scalar loads into scalar registers wherever they are placed.
__device__ __forceinline__ SceneDesc scene_uniform(const SceneDesc *sd, uint32_t q)
{
#if __HIP_DEVICE_COMPILE__
    typedef const __attribute__((address_space(4))) uint32_t ConstWord;
    const ConstWord *w = (const ConstWord *)(sd + q);
    SceneDesc d;
    uint32_t *o = reinterpret_cast<uint32_t *>(&d);
#pragma unroll
    for (uint32_t j = 0; j < sizeof(SceneDesc) / 4; ++j) o[j] = w[j];
    return d;
#else
    return sd[q];  // (host pass: never called)
#endif
}

// Runs body(S) with slot k's scene.  MULTI: the wave takes its scenes one at
// a time (one pass unless it spans a scene boundary), the lanes of that scene
// active and its descriptor read with scalar loads into scalar registers.
// The loop runs on a wave-uniform mask, and a lane's membership is read back
// from a ballot, so the compiler keeps the body inside the loop with the
// uniform scene index (rather than substituting the per-lane one).
template <bool MULTI, typename Body>
__device__ __forceinline__ void run_scene(const SceneDesc *sd, const uint32_t *scene_of, uint32_t k,
                                          Body &body)
{
    if (!MULTI) {
        body(sd[0]);
        return;
    }
    const uint32_t q = scene_of[k], lane = threadIdx.x & 63u;
    uint64_t pending = __ballot(1);
    while (pending) {
        const uint32_t qq = (uint32_t)__builtin_amdgcn_readlane((int)q, (int)__builtin_ctzll(pending));
        const uint64_t mine = __ballot(q == qq);
        if ((mine >> lane) & 1u) body(scene_uniform(sd, qq));
        pending &= ~mine;
    }
}

template <bool MULTI>
__device__ __forceinline__ const SceneDesc &scene_of_slot(const SceneDesc *sd,
                                                          const uint32_t *scene_of, uint32_t k)
{
    return MULTI ? sd[scene_of[k]] : sd[0];
}

// PositionToCellCoord (cc:499-503): floor(pos / r) with IEEE division, then (int).
__device__ __forceinline__ int cell_coord(float x, float h) { return (int)floorf(x / h); }

// ---------------------------------------------------------------------------
// Stage A + B1
// ---------------------------------------------------------------------------
// 10-bit Morton spread: bit i of x moves to bit 3i.
__device__ __forceinline__ uint32_t spread3(uint32_t x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Ordering key of the gather kernels' work order: scene index in the top
// bits, then the Morton code of the cell (relative to the scene's box,
// clamped to 10 bits per axis).  It only decides WHICH work-item handles a
// particle, never a result, so clamping and truncation are harmless.
__device__ __forceinline__ uint32_t work_key(const EngineDesc &E, const SceneDesc &S, uint32_t s,
                                             int cx, int cy, int cz)
{
    const uint32_t ux = (uint32_t)min(max(cx - S.cell_lo[0], 0), 1023);
    const uint32_t uy = (uint32_t)min(max(cy - S.cell_lo[1], 0), 1023);
    const uint32_t uz = (uint32_t)min(max(cz - S.cell_lo[2], 0), 1023);
    const uint32_t m = spread3(ux) | (spread3(uy) << 1) | (spread3(uz) << 2);
    return E.scene_shift >= 32 ? m : ((s << E.scene_shift) | (m >> E.morton_drop));
}

// ---------------------------------------------------------------------------
// Stage A + B1
// ---------------------------------------------------------------------------
// Stage A of one particle (cc:42-48) and its cell (cc:499-503) and hash
// (cc:505-511).  Every kernel that needs them recomputes them from the same
// inputs with the same operations, so the values agree to the bit.
struct Predicted {
    float vx, vy, vz;  // velocity += CalculateExternalFoce(...) * deltatime (cc:45)
    float px, py, pz;  // predictedPositions = positions + velocity * (1/120) (cc:47)
    int cx, cy, cz;
    uint32_t hash;
};

__device__ __forceinline__ Predicted predict_particle(const SceneDesc &S, const float4 &x,
                                                      const float4 &v)
{
    Predicted P;
    const float dt = S.dt;
    P.vx = v.x + S.g[0] * dt;
    P.vy = v.y + S.g[1] * dt;
    P.vz = v.z + S.g[2] * dt;
    P.px = x.x + P.vx * S.lookahead;
    P.py = x.y + P.vy * S.lookahead;
    P.pz = x.z + P.vz * S.lookahead;
    P.cx = cell_coord(P.px, S.h);
    P.cy = cell_coord(P.py, S.h);
    P.cz = cell_coord(P.pz, S.h);
    P.hash = hash_cell(P.cx, P.cy, P.cz);
    return P;
}

// Where the next sort puts a particle: its bucket (GetKeyFromHash, cc:513-516,
// or the cell-grid key) and a ticket, a distinct place in the bucket in
// arbitrary order, from count ([table_total] records per bucket, zero before
// the first record; then [n_scenes] flags raised when a particle's cell lies
// outside the grid's box).  One atomic per run of equal keys in consecutive
// lanes: in the resident order, nearly sorted, a wave's 64 records fall in a
// few runs.  Every lane of the wave that has not exited calls this, the
// active lanes a prefix of the wave; S is read only when `valid`.
__device__ __forceinline__ void record_bucket(const EngineDesc &E, const SceneDesc &S,
                                              const Predicted &P, bool valid, uint32_t g,
                                              uint32_t *__restrict__ keys,
                                              uint32_t *__restrict__ ticks,
                                              uint32_t *__restrict__ count,
                                              uint32_t *__restrict__ wkeys)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t upto = (2ull << lane) - 1ull;  // lanes 0..lane
    uint32_t key = 0xffffffffu;
    if (valid) {
        key = S.key_off + (S.grid ? grid_key(P.cx, P.cy, P.cz, S.cell_lo, S.gdim, S.gmod)
                                  : fastmod(P.hash, S.mod_m, S.table));
        keys[g] = key;
        // grid key: a cell outside the grid's box wraps onto another cell's bucket
        if (S.grid && ((uint32_t)P.cx - (uint32_t)S.cell_lo[0] >= S.gdim[0] ||
                       (uint32_t)P.cy - (uint32_t)S.cell_lo[1] >= S.gdim[1] ||
                       (uint32_t)P.cz - (uint32_t)S.cell_lo[2] >= S.gdim[2]))
            count[E.table_total + S.index] = 1u;
        if (wkeys) wkeys[g] = work_key(E, S, S.index, P.cx, P.cy, P.cz);
    }
    // runs of equal keys: a run starts at lane 0, after a different key, and
    // after a lane without a record (those come after every lane with one)
    const uint32_t prev = __shfl_up(key, 1u, 64);
    const uint64_t heads = __ballot(valid && (lane == 0u || prev != key)) | ~__ballot(valid);
    const uint32_t first = 63u - (uint32_t)__builtin_clzll(heads & upto);  // this run's head
    const uint64_t later = heads & ~upto;
    const uint32_t len = (later ? (uint32_t)__builtin_ctzll(later) : 64u) - lane;
    uint32_t base = 0;
    if (valid && first == lane) base = atomicAdd(&count[key], len);
    base = __shfl(base, (int)first, 64);
    if (valid) ticks[g] = base + (lane - first);
}

// Stage A + B1 over resident slots g, for an update whose buckets the previous
// update did not record (the first one, or after a change of state, settings
// or dt); otherwise the viscosity pass of the previous update recorded them.
template <bool MULTI>
__global__ void k_predict(const EngineDesc *__restrict__ ed, const SceneDesc *__restrict__ sd,
                          const uint32_t *__restrict__ scene_of, const float4 *__restrict__ pos,
                          const float4 *__restrict__ vel, uint32_t *__restrict__ keys,
                          uint32_t *__restrict__ ticks, uint32_t *__restrict__ count,
                          uint32_t *__restrict__ wkeys)
{
    const EngineDesc E = *ed;
    // whole waves iterate together (record_bucket's ballots need every lane)
    for (uint32_t g0 = blockIdx.x * blockDim.x; g0 < E.n_total; g0 += grid_stride()) {
        const uint32_t g = g0 + threadIdx.x;
        const bool valid = g < E.n_total;
        const SceneDesc &S = sd[MULTI && valid ? scene_of[g] : 0u];
        Predicted P{};
        if (valid) P = predict_particle(S, pos[g], vel[g]);
        record_bucket(E, S, P, valid, g, keys, ticks, count, wkeys);
    }
}

// count[0, n) <- 0: the counts and flags of a recorded update that will not run.
__global__ void k_count_clear(uint32_t *__restrict__ count, uint32_t n)
{
    for (uint32_t i = global_id(); i < n; i += grid_stride()) count[i] = 0u;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort (the Morton work order).  A fixed grid of `nblk`
// blocks; block b owns the contiguous chunk [b*chunk, min((b+1)*chunk, n)) and
// walks it tile by tile in order, so equal digits keep their input order
// (stability).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void chunk_of(uint32_t n, uint32_t &beg, uint32_t &end)
{
    const uint32_t nblk = gridDim.x;
    const uint32_t chunk = (n + nblk - 1) / nblk;
    beg = min(n, blockIdx.x * chunk);
    end = min(n, beg + chunk);
}

__global__ __launch_bounds__(kSortBlock) void k_radix_hist(const EngineDesc *__restrict__ ed,
                                                            const uint32_t *__restrict__ keys,
                                                            uint32_t shift,
                                                            uint32_t *__restrict__ hist)
{
    // one private histogram per wave limits LDS atomic contention
    __shared__ uint32_t cnt[kSortWaves][kRadix];
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    for (uint32_t i = tid; i < kSortWaves * kRadix; i += kSortBlock) (&cnt[0][0])[i] = 0;
    __syncthreads();
    uint32_t beg, end;
    chunk_of(ed->n_total, beg, end);
    for (uint32_t e = beg + tid; e < end; e += kSortBlock)
        atomicAdd(&cnt[w][(keys[e] >> shift) & (kRadix - 1)], 1u);
    __syncthreads();
    for (uint32_t d = tid; d < kRadix; d += kSortBlock) {
        uint32_t s = 0;
        for (int q = 0; q < kSortWaves; ++q) s += cnt[q][d];
        hist[d * gridDim.x + blockIdx.x] = s;  // digit-major: scan gives global offsets
    }
}

// Block-wide exclusive scan of one value per work-item (kSortBlock items);
// returns the exclusive prefix and writes the block total to *total.
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t *wsum, uint32_t *total)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int q = 0; q < kSortWaves; ++q) {
        before += (uint32_t)q < w ? wsum[q] : 0u;
        all += wsum[q];
    }
    *total = all;
    __syncthreads();
    return before + x - v;
}

// One block per digit: exclusive scan of that digit's per-block counts (in
// place) and the digit's total count.
__global__ __launch_bounds__(kSortBlock) void k_radix_scan_digit(uint32_t *__restrict__ hist,
                                                                  uint32_t nblk,
                                                                  uint32_t *__restrict__ digit_total)
{
    __shared__ uint32_t wsum[kSortWaves];
    const uint32_t d = blockIdx.x, tid = threadIdx.x;
    uint32_t *row = hist + (size_t)d * nblk;
    const uint32_t per = (nblk + kSortBlock - 1) / kSortBlock;  // <= 4 for nblk <= 1024
    uint32_t v[4] = {0, 0, 0, 0}, sum = 0;
    for (uint32_t j = 0; j < per && j < 4; ++j) {
        const uint32_t i = tid * per + j;
        v[j] = i < nblk ? row[i] : 0u;
        sum += v[j];
    }
    uint32_t total;
    uint32_t run = block_scan_excl(sum, wsum, &total);
    for (uint32_t j = 0; j < per && j < 4; ++j) {
        const uint32_t i = tid * per + j;
        if (i < nblk) row[i] = run;
        run += v[j];
    }
    if (tid == 0) digit_total[d] = total;
}

__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(
    const EngineDesc *__restrict__ ed, const uint32_t *__restrict__ kin,
    const uint32_t *__restrict__ vin, uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
    uint32_t shift, const uint32_t *__restrict__ hist, const uint32_t *__restrict__ digit_total)
{
    static_assert(kRadix == kSortBlock, "one digit per work-item");
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t wcnt[kSortWaves][kRadix];
    __shared__ uint32_t wsum[kSortWaves];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    {
        // global offset of digit `tid` = (counts of smaller digits) + (this
        // digit's counts in earlier blocks)
        uint32_t total;
        const uint32_t digit_base = block_scan_excl(digit_total[tid], wsum, &total);
        base[tid] = digit_base + hist[tid * gridDim.x + blockIdx.x];
        for (int q = 0; q < kSortWaves; ++q) wcnt[q][tid] = 0;
    }
    __syncthreads();
    uint32_t beg, end;
    chunk_of(ed->n_total, beg, end);
    const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (uint32_t t0 = beg; t0 < end; t0 += kSortBlock) {
        const uint32_t e = t0 + tid;
        const bool valid = e < end;
        const uint32_t key = valid ? kin[e] : 0u;
        const uint32_t val = valid ? (vin ? vin[e] : e) : 0u;  // no input values: slot index
        const uint32_t d = (key >> shift) & (kRadix - 1);
        // lanes of this wave holding the same digit: intersect one ballot per bit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = base[d] + rank;
            for (uint32_t q = 0; q < w; ++q) off += wcnt[q][d];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        for (uint32_t dd = tid; dd < kRadix; dd += kSortBlock) {
            uint32_t s = 0;
            for (int q = 0; q < kSortWaves; ++q) {
                s += wcnt[q][dd];
                wcnt[q][dd] = 0;
            }
            base[dd] += s;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Stages B2 + B3 (cc:481-496): counting sort over the buckets.  The reference
// sorts (key, index) records with std::sort (cc:484) and takes each bucket's
// first slot (cc:486-496); here the sorted order is the stable one, ascending
// particle index within a bucket, and the table is the exclusive prefix sum of
// the bucket counts: bucket b holds slots [bounds[b], bounds[b+1]), empty when
// they are equal (the reference's INT_MAX start, cc:481).
//   k_predict        count[key]++ and a ticket per particle (above), or the
//                    previous update's viscosity pass
//   k_count_scan     bounds: one pass over the counts in chunks, with the
//                    chunk prefixes by decoupled look-back; publishes the
//                    scene flags
//   k_count_scatter  tmp[bounds[key] + ticket] = slot: each bucket's slots, in
//                    ticket order
//   k_count_reorder  a record's place in its bucket is the number of the
//                    bucket's slots with a smaller index; gather into sorted
//                    order; clear the counts for the next update
// ---------------------------------------------------------------------------
constexpr uint32_t kCountChunk = kCountChunkBuckets;      // buckets per block
constexpr uint32_t kCountPer = kCountChunk / kSortBlock;   // buckets per work-item
static_assert(kCountPer % 4 == 0 && kCountPer * kSortBlock == kCountChunk, "count chunk");

__device__ __forceinline__ void load_counts(const uint32_t *__restrict__ count, uint32_t b0,
                                            uint32_t T, uint32_t c[kCountPer])
{
    if (b0 + kCountPer <= T) {
        const uint4 *q = reinterpret_cast<const uint4 *>(count + b0);
#pragma unroll
        for (uint32_t j = 0; j < kCountPer / 4; ++j) {
            const uint4 u = q[j];
            c[4 * j] = u.x;
            c[4 * j + 1] = u.y;
            c[4 * j + 2] = u.z;
            c[4 * j + 3] = u.w;
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kCountPer; ++j) c[j] = b0 + j < T ? count[b0 + j] : 0u;
    }
}

// Bucket bounds in one pass (decoupled look-back).  Each workgroup takes the
// next chunk in launch order from a counter, publishes the chunk's count sum,
// then adds the sums of the chunks before it -- its first wave reads 64 status
// words at a time, back to the nearest chunk whose inclusive prefix is
// published -- and publishes its own inclusive prefix.  A workgroup waits only
// for chunks taken before its own, by workgroups already running, so every
// wait ends; the wait is bounded as well (EngineDesc::scan_spin sleeps of 64
// clocks).  A wait that reaches the bound counts the unpublished chunks as
// empty and raises *scan_err, which the host reads at its next
// synchronisation and reports as an error (the update's results are then
// invalid); every bound written is clamped to the particle count, so the
// kernels after it stay inside the particle arrays either way.  chain:
// [gridDim.x] status words, then the chunk counter; all zero on entry (the
// reorder clears them).  The workgroup of the last chunk writes bounds[T] and
// publishes the scene flags.
constexpr uint64_t kChainAgg = 1ull << 62;    // the chunk's count sum
constexpr uint64_t kChainInc = 2ull << 62;    // the inclusive prefix through the chunk
constexpr uint64_t kChainFlags = 3ull << 62;

__device__ __forceinline__ uint64_t chain_load(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void chain_store(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Chunks of the count scan holding the T buckets in use (at least one,
// which writes bounds[T] and the scene flags); nb = the chunks the table's
// capacity needs.
__device__ __forceinline__ uint32_t scan_chunks_live(const EngineDesc *ed, uint32_t nb)
{
    return min(nb, max(1u, (ed->table_total + kCountChunk - 1) / kCountChunk));
}

// Chunk c (< na) of the count scan, by one workgroup of kSortBlock.
__device__ __forceinline__ void scan_chunk(uint32_t c, uint32_t na, const EngineDesc *__restrict__ ed,
                                           const uint32_t *__restrict__ count, uint64_t *__restrict__ chain,
                                           uint32_t *__restrict__ bounds, uint32_t *__restrict__ sflags,
                                           uint32_t *__restrict__ scan_err)
{
    __shared__ uint32_t wsum[kSortWaves];
    __shared__ uint32_t base_s;
    const uint32_t T = ed->table_total, tid = threadIdx.x;
    const uint32_t n = ed->n_total, spin_max = ed->scan_spin;
    const uint32_t b0 = c * kCountChunk + tid * kCountPer;
    uint32_t cnt[kCountPer];
    load_counts(count, b0, T, cnt);
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kCountPer; ++j) sum += cnt[j];
    uint32_t total;
    const uint32_t local = block_scan_excl(sum, wsum, &total);
    if (tid < 64) {
        const uint32_t lane = tid;
        if (lane == 0) chain_store(&chain[c], (c == 0 ? kChainInc : kChainAgg) | total);
        uint32_t prefix = 0, spins = 0;
        for (int32_t j0 = (int32_t)c - 1; j0 >= 0; j0 -= 64) {
            const int32_t j = j0 - (int32_t)lane;  // lane 0: the nearest chunk
            uint64_t st = j >= 0 ? chain_load(&chain[j]) : kChainInc;  // before chunk 0: 0
            while (__ballot((st & kChainFlags) == 0) != 0 && spins < spin_max) {
                __builtin_amdgcn_s_sleep(1);
                ++spins;
                if ((st & kChainFlags) == 0) st = chain_load(&chain[j]);
            }
            // the bound is reached: the chunks still unpublished count as empty
            if (__ballot((st & kChainFlags) == 0) != 0 && lane == 0) *scan_err = 1u;
            const uint64_t inc = __ballot((st & kChainFlags) == kChainInc);
            // the chunks up to the nearest inclusive prefix (all 64 if none)
            const uint32_t stop = inc ? (uint32_t)__builtin_ctzll(inc) : 63u;
            uint32_t v = lane <= stop ? (uint32_t)st : 0u;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            prefix += v;
            if (inc) break;
        }
        if (lane == 0) {
            if (c > 0) chain_store(&chain[c], kChainInc | (uint64_t)(prefix + total));
            base_s = prefix;
        }
    }
    __syncthreads();
    // (min: an exact prefix never exceeds n; one after a given-up wait stays in range)
    uint32_t p = base_s + local;
    if (b0 + kCountPer <= T) {
        uint4 *q = reinterpret_cast<uint4 *>(bounds + b0);
        uint32_t o[kCountPer];
#pragma unroll
        for (uint32_t j = 0; j < kCountPer; ++j) {
            o[j] = min(p, n);
            p += cnt[j];
        }
#pragma unroll
        for (uint32_t j = 0; j < kCountPer / 4; ++j)
            q[j] = make_uint4(o[4 * j], o[4 * j + 1], o[4 * j + 2], o[4 * j + 3]);
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kCountPer; ++j) {
            if (b0 + j < T) bounds[b0 + j] = min(p, n);
            p += cnt[j];
        }
    }
    if (c == na - 1) {
        if (tid == 0) bounds[T] = min(base_s + total, n);
        for (uint32_t q = tid; q < ed->n_scenes; q += kSortBlock) sflags[q] = count[T + q];
    }
}

__global__ __launch_bounds__(kSortBlock) void k_count_scan(const EngineDesc *__restrict__ ed,
                                                            const uint32_t *__restrict__ count,
                                                            uint64_t *__restrict__ chain,
                                                            uint32_t *__restrict__ bounds,
                                                            uint32_t *__restrict__ sflags,
                                                            uint32_t *__restrict__ scan_err)
{
    __shared__ uint32_t chunk_s;
    // the grid covers the table's capacity; only the live chunks' workgroups
    // take tickets (tickets 0 .. na-1)
    const uint32_t nb = gridDim.x, na = scan_chunks_live(ed, nb);
    if (blockIdx.x >= na) return;
    if (threadIdx.x == 0) chunk_s = atomicAdd(reinterpret_cast<uint32_t *>(chain + nb), 1u);
    __syncthreads();
    const uint32_t c = chunk_s;
    if (c >= na) return;  // never: na workgroups take na tickets
    scan_chunk(c, na, ed, count, chain, bounds, sflags, scan_err);
}

// order[bounds[key] + ticket] = (resident slot, particle index)
// (work-items first, first + stride, ... of the grid-stride loop)
__device__ __forceinline__ void scatter_range(uint32_t first, uint32_t stride, const EngineDesc *__restrict__ ed,
                                              const uint32_t *__restrict__ keys,
                                              const uint32_t *__restrict__ ticks,
                                              const uint32_t *__restrict__ ids,
                                              const uint32_t *__restrict__ bounds, uint2 *__restrict__ order)
{
    const uint32_t n = ed->n_total;
    for (uint32_t g = first; g < n; g += stride) {
        const uint32_t slot = bounds[keys[g]] + ticks[g];  // < n: the counts started at zero
        if (slot < n) order[slot] = make_uint2(g, ids[g]);
    }
}

__global__ void k_count_scatter(const EngineDesc *__restrict__ ed,
                                const uint32_t *__restrict__ keys,
                                const uint32_t *__restrict__ ticks,
                                const uint32_t *__restrict__ ids,
                                const uint32_t *__restrict__ bounds, uint2 *__restrict__ order)
{
    scatter_range(global_id(), grid_stride(), ed, keys, ticks, ids, bounds, order);
}

// A record's sorted slot is its bucket's first slot plus the number of the
// bucket's records with a smaller particle index.  Recomputes stage A from the
// resident state and writes the sorted copies: xps (predicted position, w =
// float(hash), the reference record's .y, cc:480), vs (velocity after the
// external force), ps (position, w = the exact hash bits) and sids, the
// particle index of each sorted slot.
//
// The count: a work-item whose bucket holds at most kRankDirect records reads
// them itself (four loads in flight per round).  Records of larger buckets are
// ranked by the whole workgroup: it walks such a bucket once in tiles of
// blockDim.x particle indices staged in LDS, every member work-item counting
// the smaller ones with broadcast LDS reads.  So a bucket of k records costs
// k global reads per workgroup that holds some of its records, not per
// record, and the per-record work is k LDS comparisons: a degenerate scene
// (every particle in one cell, or a tiny reference table) stays bounded where
// per-record global reads would take k^2 of them.  The workgroup's records
// are consecutive sorted places, so few large buckets meet in one workgroup.
// Launched with kReorderBlock work-items per workgroup.
constexpr uint32_t kRankDirect = 64;
constexpr uint32_t kReorderBlock = 256;

// Workgroup b of nblk (of kReorderBlock work-items) of the reorder.
template <bool MULTI>
__device__ __forceinline__ void reorder_blocks(
    uint32_t b, uint32_t nblk, const EngineDesc *__restrict__ ed, const SceneDesc *__restrict__ sd,
    const uint32_t *__restrict__ scene_of, const uint32_t *__restrict__ keys,
    const uint32_t *__restrict__ bounds, const uint2 *__restrict__ order,
    uint32_t *__restrict__ sids, const float4 *__restrict__ vel, const float4 *__restrict__ pos,
    float4 *__restrict__ xps, float4 *__restrict__ vs, float4 *__restrict__ ps,
    uint32_t *__restrict__ rank, uint32_t *__restrict__ count, uint64_t *__restrict__ chain,
    uint32_t chain_words)
{
    __shared__ uint32_t tile[kReorderBlock];
    __shared__ uint32_t next_key;
    const uint32_t n = ed->n_total, B = blockDim.x, tid = threadIdx.x;
    // whole workgroups iterate together (the cooperative rank synchronises)
    for (uint32_t t0 = b * B; t0 < n; t0 += nblk * B) {
        const uint32_t t = t0 + tid;
        uint2 o = make_uint2(n, 0u);  // (resident slot, particle index)
        if (t < n) o = order[t];
        const bool have = o.x < n;    // always, while the counts start at zero
        float4 x0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v0 = x0;
        uint32_t key = 0, lo = 0, hi = 0;
        if (have) {
            x0 = pos[o.x];  // in flight during the bucket count
            v0 = vel[o.x];
            key = keys[o.x];
            lo = bounds[key];
            hi = min(bounds[key + 1], n);
        }
        const bool big = have && hi > lo && hi - lo > kRankDirect;
        uint32_t before = 0;
        if (have && !big) {
            // the bucket's records four at a time, all four loads issued
            // before any compare (past hi the load repeats record j and is not
            // counted)
            for (uint32_t j = lo; j < hi; j += 4) {
                uint32_t y[4];
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) y[q] = order[j + q < hi ? j + q : j].y;
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) before += (j + q < hi && y[q] < o.y) ? 1u : 0u;
            }
        }
        if (__syncthreads_or(big)) {
            bool pending = big;
            for (;;) {
                // the next large bucket of this workgroup (the smallest key left)
                if (tid == 0) next_key = 0xffffffffu;
                __syncthreads();
                if (pending) atomicMin(&next_key, key);
                __syncthreads();
                const uint32_t bk = next_key;
                if (bk == 0xffffffffu) break;
                const uint32_t blo = bounds[bk], bhi = min(bounds[bk + 1], n);
                const bool mine = pending && key == bk;
                for (uint32_t j0 = blo; j0 < bhi; j0 += B) {
                    const uint32_t m = min(B, bhi - j0);
                    __syncthreads();  // the previous tile is consumed
                    if (tid < m) tile[tid] = order[j0 + tid].y;
                    __syncthreads();
                    if (mine) {
                        uint32_t c0 = 0, c1 = 0;
                        uint32_t q = 0;
                        for (; q + 1 < m; q += 2) {
                            c0 += tile[q] < o.y ? 1u : 0u;
                            c1 += tile[q + 1] < o.y ? 1u : 0u;
                        }
                        if (q < m) c0 += tile[q] < o.y ? 1u : 0u;
                        before += c0 + c1;
                    }
                }
                pending = pending && !mine;
            }
        }
        if (!have) continue;
        const uint32_t slot = lo + before;
        if (slot >= hi) continue;  // as above
        const SceneDesc &S = scene_of_slot<MULTI>(sd, scene_of, o.x);
        const Predicted P = predict_particle(S, x0, v0);
        sids[slot] = o.y;
        xps[slot] = make_float4(P.px, P.py, P.pz, (float)P.hash);
        vs[slot] = make_float4(P.vx, P.vy, P.vz, v0.w);
        ps[slot] = make_float4(x0.x, x0.y, x0.z, __uint_as_float(P.hash));
        if (rank) rank[o.y] = slot;  // only the Morton work order reads it
    }
    const uint32_t nc = ed->table_total + ed->n_scenes;
    for (uint32_t i = b * B + tid; i < nc; i += nblk * B) count[i] = 0u;
    for (uint32_t i = b * B + tid; i < chain_words; i += nblk * B) chain[i] = 0u;
}

template <bool MULTI>
__global__ __launch_bounds__(kReorderBlock) void k_count_reorder(
    const EngineDesc *__restrict__ ed, const SceneDesc *__restrict__ sd,
    const uint32_t *__restrict__ scene_of, const uint32_t *__restrict__ keys,
    const uint32_t *__restrict__ bounds, const uint2 *__restrict__ order,
    uint32_t *__restrict__ sids, const float4 *__restrict__ vel, const float4 *__restrict__ pos,
    float4 *__restrict__ xps, float4 *__restrict__ vs, float4 *__restrict__ ps,
    uint32_t *__restrict__ rank, uint32_t *__restrict__ count, uint64_t *__restrict__ chain,
    uint32_t chain_words)
{
    reorder_blocks<MULTI>(blockIdx.x, gridDim.x, ed, sd, scene_of, keys, bounds, order, sids, vel, pos, xps,
                          vs, ps, rank, count, chain, chain_words);
}

// Resident slots [0, n) of one scene into particle order:
// dst_pos[ids[k] - id0] = pos[k] (and vel; either destination may be null).
__global__ void k_unpermute_state(const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                  const uint32_t *__restrict__ ids, uint32_t n, uint32_t id0,
                                  float4 *__restrict__ dst_pos, float4 *__restrict__ dst_vel)
{
    for (uint32_t k = global_id(); k < n; k += grid_stride()) {
        const uint32_t q = ids[k] - id0;
        if (q >= n) continue;  // ids of the scene lie in [id0, id0 + n)
        if (dst_pos) dst_pos[q] = pos[k];
        if (dst_vel) dst_vel[q] = vel[k];
    }
}

__global__ void k_iota(uint32_t *__restrict__ a, uint32_t n, uint32_t first)
{
    for (uint32_t k = global_id(); k < n; k += grid_stride()) a[k] = first + k;
}

// scene_of[p_off + i] = s for every scene s (blockIdx.y) and i < n, from the
// scene table on the device
__global__ void k_scene_map(const SceneDesc *__restrict__ sd, uint32_t *__restrict__ scene_of)
{
    const SceneDesc &S = sd[blockIdx.y];
    const uint32_t n = S.n, off = S.p_off;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        scene_of[off + i] = blockIdx.y;
}

// ---------------------------------------------------------------------------
// Neighbour walk (cc:331-363, 378-419, 432-462): 27 cells in the reference's
// offset order (physicsWorld.h:131-143: dx outer, dz inner); within a bucket
// in sorted order; records of other cells sharing the bucket are skipped by
// the float-rounded hash test.  `body(m, ox, oy, oz, d2, aux)` runs for every
// record with d2 <= r2 (the sqrRadius prefilter) and d < h, in that order;
// records with d >= h contribute exactly zero to every sum of the reference.
//
// Latency hiding without changing the summation order: the bucket range of
// the next cell is fetched while the current bucket is walked, and a bucket is
// walked KB records at a time (all loads of a batch issued before any test).
// AUX: also fetch aux[m] (the neighbour's density or velocity) for the
// records of a batch that pass the tests.
// ---------------------------------------------------------------------------
constexpr int KB = 4;

template <bool SKIP_SELF, bool AUX, typename Body>
__device__ __forceinline__ void walk_hash(const SceneDesc &S, const float4 *__restrict__ xps,
                                     const float4 *__restrict__ aux,
                                     const uint32_t *__restrict__ start,
                                     const uint32_t *__restrict__ endt, uint32_t k,
                                     const float4 &pi, Body &&body)
{
    const int cx = cell_coord(pi.x, S.h), cy = cell_coord(pi.y, S.h), cz = cell_coord(pi.z, S.h);
    const float thr = S.thr;
    const uint32_t key_off = S.key_off, table = S.table;
    const uint64_t mod_m = S.mod_m;
    uint32_t nb, ne;
    float nhf;
    auto fetch = [&](int o) {
        const int dx = o / 9 - 1, dy = (o / 3) % 3 - 1, dz = o % 3 - 1;
        const uint32_t hash = hash_cell(cx + dx, cy + dy, cz + dz);
        const uint32_t key = key_off + fastmod(hash, mod_m, table);
        nb = start[key];
        ne = endt[key];
        nhf = (float)hash;
    };
    fetch(0);
#pragma unroll 1
    for (int o = 0; o < 27; ++o) {
        const uint32_t b = nb, e = ne;
        const float hf = nhf;
        if (o + 1 < 27) fetch(o + 1);
#pragma unroll 1
        for (uint32_t m = b; m < e; m += KB) {
            float4 r[KB];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                r[j] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0x7fc00000));  // NaN: no match
                if (m + j < e) r[j] = xps[m + j];
            }
            bool pass[KB];
            float ox[KB], oy[KB], oz[KB], d2[KB];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                ox[j] = r[j].x - pi.x;
                oy[j] = r[j].y - pi.y;
                oz[j] = r[j].z - pi.z;
                d2[j] = ox[j] * ox[j] + oy[j] * oy[j] + oz[j] * oz[j];  // glm::dot order
                pass[j] = r[j].w == hf && !(SKIP_SELF && m + j == k) && d2[j] < thr;
            }
            float4 a[KB];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                a[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (AUX && pass[j]) a[j] = aux[m + j];
            }
#pragma unroll
            for (int j = 0; j < KB; ++j)
                if (pass[j]) body(m + j, ox[j], oy[j], oz[j], d2[j], a[j]);
        }
    }
}

// The same walk under the cell-grid key.  The 27 cells, in the reference's
// order, form 9 rows (dx, dy) of three cells dz = -1, 0, 1.  Where the three
// cells of a row are consecutive buckets (everywhere but the grid's wrap
// seam) their records are one contiguous run of sorted slots, walked as one;
// a record belongs to the bucket its slot falls in and is tested against that
// cell's float hash, exactly as in the hash walk.  At the seam the three
// buckets are walked one after the other.  Either way `body` sees the same
// records in the same order as under the reference key.  The bucket ranges of
// the next row are fetched while the current row is walked.
//
// Row staging (stage != nullptr): the work-items of a wave hold consecutive
// sorted slots -- a z-segment of one grid column -- so their runs in a row
// overlap inside one short slot range [lo, hi) (lo = the first active lane's
// run start, hi = the last one's end).  When every active lane's run lies in
// that range and it fits kStageCap records, the wave copies it into its LDS
// buffer with coalesced loads and the lanes walk their runs from LDS.  The
// check is exact per lane, so the range assumption never affects a result;
// rows that fail it are walked from global memory.  The buffer holds the
// records as four planes of kStageBuf floats (x, y, z, w): lanes reading
// nearby records then hit distinct banks, where 16-byte records at a 16-byte
// stride made lanes eight records apart collide.
constexpr int KBG = 4;                          // records per batch of the grid walk
constexpr uint32_t kStageCap = 156;             // staged records per wave and row
// floats per plane: the cap + padding (a batch may read KBG-1 past the end).
// Four planes and the density pass's 512-bit stream words (16 per lane) fill
// the 26 words per lane of the mask table exactly, so the cap costs no LDS;
// compressed fluid (16-30 particles a cell) needs ranges past 120 records.
constexpr uint32_t kStageBuf = 160;
static_assert(kStageBuf >= kStageCap + KBG, "staging planes");

// Slots [b, e) of a grid row: the buckets of the row's cells dz = -1, 0, 1
// (float hashes g0, g1, g2), in that order.  A record passes when d2 < thr
// and its float hash is one of g0, g1, g2.  That equals the reference's test
// (hash of the record's own bucket cell): a record with d2 < thr lies within
// h, so its cell is one of the 27 around the particle, whose float hashes are
// pairwise distinct; matching g_q therefore means it is a particle of cell q,
// which sits in bucket q.  The tests are combined without short-circuit
// branches.  STAGED: the records are read from the wave's LDS planes at
// slot - base, for a staged range containing [b, e) (the padded planes make
// the reads past e harmless; they never pass); else from xps in global memory.
template <bool STAGED, bool SKIP_SELF, bool AUX, bool MASKS, bool HASH, typename Body>
__device__ __forceinline__ void walk_run(const float4 *src, uint32_t base, const float4 *__restrict__ xps,
                                         const float4 *__restrict__ aux, uint32_t k,
                                         const float4 &pi, float thr, uint32_t b, uint32_t e,
                                         float g0, float g1, float g2, uint64_t &mask, Body &body)
{
#pragma unroll 1
    for (uint32_t m = b; m < e; m += KBG) {
        float4 rr[KBG];
#pragma unroll
        for (int j = 0; j < KBG; ++j) {
            if (STAGED) {
                const float *sp = reinterpret_cast<const float *>(src);
                const uint32_t q = m + j - base;
                rr[j] = make_float4(sp[q], sp[kStageBuf + q], sp[2 * kStageBuf + q], sp[3 * kStageBuf + q]);
            } else {
                rr[j] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0x7fc00000));  // NaN: no match
                if (m + j < e) rr[j] = xps[m + j];
            }
        }
        bool pass[KBG];
        float ox[KBG], oy[KBG], oz[KBG], d2[KBG];
#pragma unroll
        for (int j = 0; j < KBG; ++j) {
            const uint32_t mj = m + j;
            ox[j] = rr[j].x - pi.x;
            oy[j] = rr[j].y - pi.y;
            oz[j] = rr[j].z - pi.z;
            d2[j] = ox[j] * ox[j] + oy[j] * oy[j] + oz[j] * oz[j];  // glm::dot order
            const float w = rr[j].w;
            pass[j] = ((int)(!STAGED || mj < e) & (int)(d2[j] < thr) &
                       (int)(!HASH || (w == g0) | (w == g1) | (w == g2)) &
                       (int)(!SKIP_SELF || mj != k)) != 0;
        }
        float4 a[KBG];
#pragma unroll
        for (int j = 0; j < KBG; ++j) {
            a[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (AUX && pass[j]) a[j] = aux[m + j];
        }
#pragma unroll
        for (int j = 0; j < KBG; ++j) {
            if (MASKS && pass[j] && m + j - b < 64u) mask |= 1ull << (m + j - b);
            if (pass[j]) body(m + j, ox[j], oy[j], oz[j], d2[j], a[j]);
        }
    }
}

// The mask of a staged run of len <= 64 records (planes x, y, z at sp[0, len),
// kStageBuf floats apart) when no float-hash test is needed: bit o set when
// record o lies within h (d2 < thr, the walk's expression).  Every lane starts
// at offset 0 and steps by KBG, so the offset is one scalar for the wave: a
// batch's four results form a nibble that one shift by that scalar places,
// with no per-record branch or bound check.  The two 32-record halves are
// walked by two loops into 32-bit words (one shift-or per batch instead of a
// 64-bit shift and two ors).  Bits at and past len (a batch reads up to KBG-1
// records past the run, inside the padded buffer) are cleared at the end.
static_assert(32 % KBG == 0, "a batch never straddles the two halves");
template <uint32_t O0>
__device__ __forceinline__ uint32_t walk_half(const float *sp, uint32_t len, const float4 &pi, float thr)
{
    uint32_t bits = 0;
    const uint32_t end = len < O0 + 32u ? len : O0 + 32u;
#pragma unroll 1
    for (uint32_t o = O0; o < end; o += KBG) {
        float4 rr[KBG];
#pragma unroll
        for (int j = 0; j < KBG; ++j)
            rr[j] = make_float4(sp[o + j], sp[kStageBuf + o + j], sp[2 * kStageBuf + o + j], 0.0f);
        uint32_t nib = 0;
#pragma unroll
        for (int j = 0; j < KBG; ++j) {
            const float ox = rr[j].x - pi.x, oy = rr[j].y - pi.y, oz = rr[j].z - pi.z;
            const float d2 = ox * ox + oy * oy + oz * oz;  // glm::dot order
            nib |= (d2 < thr ? 1u : 0u) << j;
        }
        bits |= nib << (o - O0);
    }
    return bits;
}

__device__ __forceinline__ uint64_t walk_bits(const float *sp, uint32_t len, const float4 &pi, float thr)
{
    const uint32_t lo = walk_half<0>(sp, len, pi, thr);
    const uint32_t hi = walk_half<32>(sp, len, pi, thr);
    const uint64_t mask = ((uint64_t)hi << 32) | lo;
    return len >= 64u ? mask : mask & ((1ull << len) - 1ull);
}

// Neighbour masks (grid key, row mode).  The density pass records, for
// sorted slot k, which slots of its nine row runs passed the walk's tests
// (self included), as one bitstream: row r's run [rb_r, rb_r + L_r) gives L_r
// bits, bit o set when slot rb_r + o passed, the rows in walk order, so row r
// starts at bit off_r = L_0 + ... + L_{r-1}.  The neighbour-row buffer holds
// bits[4][n] (uint4: the 512-bit stream; word w > 1 only written, and read,
// when the stream is longer than 128 w bits) and then meta[n] (uint2: the
// nine lengths, 7 bits each, as a 64-bit word whose top bit, kMaskUnusable,
// marks an unusable record).  A record is unusable -- the consumers walk
// that particle instead -- when a row's three buckets are not consecutive
// (the z wrap seam), a run exceeds 64 slots or the runs exceed 512 bits in
// all.  Otherwise a consumer re-derives each rb_r as the walk found it, from
// the particle's cell (its position is unchanged between the passes) and the
// bucket bounds (a table the wave's lanes share in cache), so the record is
// 40 bytes (56 past 256 bits, 72 past 384).  A lattice of spacing h / 2 gives
// 216 bits; fluid compressed to twice that density (a dam break's impact)
// about 350.
constexpr uint32_t kMaskBitsMax = 512;
constexpr uint32_t kMaskWords = kMaskBitsMax / 32;
constexpr uint32_t kMaskMetaAt = kMaskWords / 2;  // meta[k] is uint2 [kMaskMetaAt n + k]
constexpr uint64_t kMaskUnusable = 1ull << 63;

struct MaskOut {
    uint32_t *buf;    // neighbour-row buffer (>= 18 n words)
    uint32_t n;
    uint32_t *words;  // this lane's stream words in LDS (stride 64)
    uint64_t lens;    // out: the nine run lengths, 7 bits each
};

// stream |= mask << off on this lane's LDS words (stride 64).  The mask has
// no bit at or past its run's length L and off + L <= kMaskBitsMax, so every
// word touched lies in the stream.
__device__ __forceinline__ void put_bits(uint32_t *mw, uint32_t off, uint64_t mask)
{
    const uint32_t q = off >> 5, s = off & 31u;
    const uint32_t lo = (uint32_t)mask, hi = (uint32_t)(mask >> 32);
    const uint32_t p0 = lo << s;
    const uint32_t p1 = s ? (hi << s) | (lo >> (32u - s)) : hi;
    const uint32_t p2 = s ? hi >> (32u - s) : 0u;
    if (p0) mw[q * 64] |= p0;
    if (p1) mw[(q + 1) * 64] |= p1;
    if (p2) mw[(q + 2) * 64] |= p2;
}

// DEFER (density): where a row's run fits the 64-bit mask, the walk over the
// run only records the mask, and the body then runs over the mask's slots
// in ascending order (the walk order), reading the records from the wave's
// LDS copy of the row when it was staged.  The body then costs one pass per
// neighbour instead of one per tested candidate batch.
// MASK_ONLY (with MASKS): the walk records the masks and calls no body; seam
// rows are not walked (they make the particle's masks unusable anyway).
// MASKS: the record is written to mo->buf, and mo->lens is set;
// returns 1 when the record is usable (0 without MASKS).
template <bool SKIP_SELF, bool AUX, bool MASKS = false, bool DEFER = false, bool MASK_ONLY = false,
          typename Body>
__device__ __forceinline__ uint32_t walk_grid(const SceneDesc &S, const float4 *__restrict__ xps,
                                          const float4 *__restrict__ aux,
                                          const uint32_t *__restrict__ start,
                                          const uint32_t *__restrict__ endt, uint32_t k,
                                          const float4 &pi, float4 *stage, bool clean, Body &&body,
                                          MaskOut *mo = nullptr)
{
    const int cx = cell_coord(pi.x, S.h), cy = cell_coord(pi.y, S.h), cz = cell_coord(pi.z, S.h);
    const float thr = S.thr;
    const uint32_t gy = S.gdim[1], gz = S.gdim[2], key_off = S.key_off;
    bool masks_ok = true;
    uint32_t moff = 0;   // MASKS: bits recorded so far
    uint64_t lens = 0;   // MASKS: the run lengths, 7 bits per row
    uint32_t *mw = MASKS ? mo->words : nullptr;
    if (MASKS)
#pragma unroll
        for (uint32_t w = 0; w < kMaskWords; ++w) mw[w * 64] = 0u;
    const uint32_t ux = (uint32_t)cx - (uint32_t)S.cell_lo[0];
    const uint32_t uy = (uint32_t)cy - (uint32_t)S.cell_lo[1];
    const uint32_t uz = (uint32_t)cz - (uint32_t)S.cell_lo[2];
    // the particle's 27 cells lie inside the grid without wrapping
    const bool interior = ux - 1u < S.gdim[0] - 2u && uy - 1u < gy - 2u && uz - 1u < gz - 2u;
    // No hash test needed: when no particle of the scene has its cell outside
    // the grid's box (clean) and this particle's 27 cells are all inside it,
    // the key is injective on the occupied cells, so each of those cells owns
    // its bucket and every record of a row's run belongs to the row.
    const bool inner = clean && interior;
    const bool nohash = __ballot(inner) == __ballot(1);
    uint32_t wz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) wz[d] = grid_wrap(uz + (uint32_t)(d - 1), gz, S.gmod[2]);
    const bool run = wz[1] == wz[0] + 1u && wz[2] == wz[1] + 1u;
    const uint32_t h0 = hash_cell(cx, cy, cz);
    // bucket bounds of the next row: the run [start of its first bucket, end of
    // its last) when its buckets are consecutive, else each bucket's.  The
    // row's x and y cells are wrapped when fetched (fewer live registers).
    uint32_t ns[3] = {0u, 0u, 0u}, ne[3] = {0u, 0u, 0u};
    auto fetch = [&](int r) {
        const uint32_t wx = grid_wrap(ux + (uint32_t)(r / 3 - 1), S.gdim[0], S.gmod[0]);
        const uint32_t wy = grid_wrap(uy + (uint32_t)(r % 3 - 1), gy, S.gmod[1]);
        const uint32_t base = key_off + (wx * gy + wy) * gz;
        if (run) {
            ns[0] = start[base + wz[0]];
            ne[2] = endt[base + wz[2]];
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                ns[q] = start[base + wz[q]];
                ne[q] = endt[base + wz[q]];
            }
        }
    };
    fetch(0);
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        const uint32_t s0 = ns[0], s1 = ns[1], s2 = ns[2], e0 = ne[0], e1 = ne[1], e2 = ne[2];
        if (r + 1 < 9) fetch(r + 1);
        // float hashes of the row's cells (hash_cell is linear mod 2^32)
        const uint32_t hr = h0 + (uint32_t)(r / 3 - 1) * 15823u + (uint32_t)(r % 3 - 1) * 9737333u;
        const float f0 = (float)(hr - 440817757u), f1 = (float)hr, f2 = (float)(hr + 440817757u);
        if (run) {
            // the row's merged run
            const uint32_t rb = s0, re = e2;
            bool staged = false;
            uint32_t lo = 0;
            if (stage) {
                // staging decision, uniform over the wave's active lanes
                const uint64_t act = __ballot(1);
                const int first = __builtin_ctzll(act), last = 63 - __builtin_clzll(act);
                lo = (uint32_t)__builtin_amdgcn_readlane((int)rb, first);
                const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)re, last);
                const bool inside = rb == re || (rb >= lo && re <= hi);
                if (__ballot(inside) == act && hi >= lo && hi - lo <= kStageCap) {
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(act >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
                    const uint32_t cnt = (uint32_t)__builtin_popcountll(act);
                    __builtin_amdgcn_wave_barrier();  // the previous row's LDS reads are issued
                    // two records per lane in flight before either is stored
                    // (past the range the second is the first again)
                    const uint32_t len = hi - lo;
                    for (uint32_t t = rank; t < len; t += 2 * cnt) {
                        const uint32_t u = t + cnt < len ? t + cnt : t;
                        const float4 a0 = xps[lo + t];
                        const float4 a1 = xps[lo + u];
                        float *sp = reinterpret_cast<float *>(stage);
                        sp[t] = a0.x;
                        sp[kStageBuf + t] = a0.y;
                        sp[2 * kStageBuf + t] = a0.z;
                        sp[u] = a1.x;
                        sp[kStageBuf + u] = a1.y;
                        sp[2 * kStageBuf + u] = a1.z;
                        if (!nohash) {  // only the float-hash test reads w
                            sp[3 * kStageBuf + t] = a0.w;
                            sp[3 * kStageBuf + u] = a1.w;
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    staged = true;
                }
            }
            uint64_t mask = 0;
            const bool defer = DEFER && re - rb <= 64u;
            auto inline_body = [&](uint32_t m, float ox, float oy, float oz, float d2, const float4 &a) {
                if (!MASK_ONLY && !defer) body(m, ox, oy, oz, d2, a);
            };
            // MASK_ONLY: a lane whose run passes 64 slots gets the first 64
            // bits (its reads stay inside its staged run) and its record is
            // unusable below, so the wave keeps the branch-free walk
            if ((DEFER || MASK_ONLY) && !SKIP_SELF && !AUX && staged && nohash &&
                (MASK_ONLY || __ballot(re - rb <= 64u) == __ballot(1)))
                mask = walk_bits(reinterpret_cast<const float *>(stage) + (rb - lo), re - rb, pi, thr);
            else if (staged && nohash)
                walk_run<true, SKIP_SELF, AUX, MASKS || DEFER, false>(stage, lo, xps, aux, k, pi, thr, rb, re,
                                                                     f0, f1, f2, mask, inline_body);
            else if (staged)
                walk_run<true, SKIP_SELF, AUX, MASKS || DEFER, true>(stage, lo, xps, aux, k, pi, thr, rb, re,
                                                                    f0, f1, f2, mask, inline_body);
            else
                walk_run<false, SKIP_SELF, AUX, MASKS || DEFER, true>(xps, 0, xps, aux, k, pi, thr, rb, re,
                                                                     f0, f1, f2, mask, inline_body);
            if (DEFER && defer) {
                uint64_t mm = mask;
                const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (staged) {
#pragma unroll 1
                    while (mm) {
                        const uint32_t m = rb + (uint32_t)__builtin_ctzll(mm);
                        mm &= mm - 1;
                        const float *sp = reinterpret_cast<const float *>(stage);
                        const float4 rec = make_float4(sp[m - lo], sp[kStageBuf + m - lo], sp[2 * kStageBuf + m - lo], 0.0f);
                        const float ox = rec.x - pi.x, oy = rec.y - pi.y, oz = rec.z - pi.z;
                        body(m, ox, oy, oz, ox * ox + oy * oy + oz * oz, zero);
                    }
                } else {
#pragma unroll 1
                    while (mm) {
                        const uint32_t m = rb + (uint32_t)__builtin_ctzll(mm);
                        mm &= mm - 1;
                        const float4 rec = xps[m];
                        const float ox = rec.x - pi.x, oy = rec.y - pi.y, oz = rec.z - pi.z;
                        body(m, ox, oy, oz, ox * ox + oy * oy + oz * oz, zero);
                    }
                }
            }
            if (MASKS) {
                const uint32_t L = re - rb;
                if (L <= 64u && moff + L <= kMaskBitsMax) {
                    put_bits(mw, moff, mask);
                    lens |= (uint64_t)L << (7 * r);
                    moff += L;
                } else {
                    masks_ok = false;
                }
            }
        } else if (!MASK_ONLY) {
            masks_ok = false;
            // wrap seam: the three buckets one by one
#pragma unroll 1
            for (int q = 0; q < 3; ++q) {
                const uint32_t b = q == 0 ? s0 : (q == 1 ? s1 : s2);
                const uint32_t e = q == 0 ? e0 : (q == 1 ? e1 : e2);
                const float fq = q == 0 ? f0 : (q == 1 ? f1 : f2);
                uint64_t unused = 0;
                walk_run<false, SKIP_SELF, AUX, false, true>(xps, 0, xps, aux, k, pi, thr, b, e, fq, fq, fq, unused,
                                                             body);
            }
        } else {
            masks_ok = false;
        }
    }
    if (!MASKS) return 0u;
    const bool usable = masks_ok;
    uint4 *bits = reinterpret_cast<uint4 *>(mo->buf);
    mo->lens = lens;
    const uint64_t meta = lens | (usable ? 0ull : kMaskUnusable);
    reinterpret_cast<uint2 *>(mo->buf)[kMaskMetaAt * (size_t)mo->n + k] =
        make_uint2((uint32_t)meta, (uint32_t)(meta >> 32));
    if (usable) {
        bits[k] = make_uint4(mw[0], mw[64], mw[128], mw[192]);
        bits[(size_t)mo->n + k] = make_uint4(mw[256], mw[320], mw[384], mw[448]);
        if (moff > 256u) bits[2 * (size_t)mo->n + k] = make_uint4(mw[512], mw[576], mw[640], mw[704]);
        if (moff > 384u) bits[3 * (size_t)mo->n + k] = make_uint4(mw[768], mw[832], mw[896], mw[960]);
    }
    return usable ? 1u : 0u;
}

// Consumer side of the neighbour masks: body(m, ox, oy, oz, d2, aux[m]) for
// every recorded slot m, rows in order and slots ascending within a row --
// the order of the walk that recorded them.  d2 is recomputed with the walk's
// expression.  next_row(mask, base) gives the next row that holds a slot, or
// leaves mask = 0 when none is left.  A row is taken as its two 32-slot halves
// (empty halves skipped), and slots are taken from two registers, the current
// half's remaining bits and the next half's, with 32-bit operations and no
// branches; NB are taken per round (a round can end short only when two
// halves run out within it), their records loaded together, and the two
// registers refilled once per round.  AUXN: the components of aux[m] the
// body reads (0, 3 or 4); only those are loaded.
template <int NB, int AUXN, typename NextRow, typename Body>
__device__ __forceinline__ void take_masks(const float4 *__restrict__ xps, const float4 *__restrict__ aux,
                                           uint32_t k, const float4 &pi, NextRow &&next_row, Body &&body)
{
    uint32_t cur = 0, nxt = 0;  // remaining bits of the current and the next half
    uint32_t cb = 0, nb = 0;    // their first slots
    uint32_t hi = 0, hb = 0;    // upper half of the row last taken, not yet used
    auto next_half = [&](uint32_t &m, uint32_t &b) {
        if (hi != 0) {
            m = hi;
            b = hb;
            hi = 0;
            return;
        }
        uint64_t row = 0;
        uint32_t base = 0;
        next_row(row, base);
        const uint32_t lo32 = (uint32_t)row, hi32 = (uint32_t)(row >> 32);
        m = lo32 != 0 ? lo32 : hi32;
        b = lo32 != 0 ? base : base + 32u;
        hi = lo32 != 0 ? hi32 : 0u;
        hb = base + 32u;
    };
    auto refill = [&]() {
        // cur == 0 implies nxt == 0: a half is moved to cur only from nxt
        if (cur == 0) next_half(cur, cb);
        if (nxt == 0) next_half(nxt, nb);
    };
    refill();
#pragma unroll 1
    while (cur != 0) {
        uint32_t m[NB];
        bool v[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            v[q] = cur != 0;
            m[q] = v[q] ? cb + (uint32_t)__builtin_ctz(cur) : k;
            cur &= cur - 1;
            const bool e = cur == 0;
            cur = e ? nxt : cur;
            cb = e ? nb : cb;
            nxt = e ? 0u : nxt;
        }
        float4 rr[NB], a[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            rr[q] = xps[m[q]];
            if (AUXN == 4) {
                a[q] = aux[m[q]];
            } else if (AUXN == 3) {
                const float3 t = *reinterpret_cast<const float3 *>(aux + m[q]);
                a[q] = make_float4(t.x, t.y, t.z, 0.0f);
            } else {
                a[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
        refill();
#pragma unroll
        for (int q = 0; q < NB; ++q)
            if (v[q]) {
                const float ox = rr[q].x - pi.x, oy = rr[q].y - pi.y, oz = rr[q].z - pi.z;
                body(m[q], ox, oy, oz, ox * ox + oy * oy + oz * oz, a[q]);
            }
    }
}

// Decodes a usable record and runs the body over its slots with take_masks.
// The stream words are in this lane's LDS column (mw, stride 64).  Each
// row's first slot rb_r is the start of the row's first bucket, found from
// the particle's cell as walk_grid finds it (a usable record has no z seam,
// so the row's buckets are consecutive).  All nine rows are cut from the
// stream first (row r: L_r bits from bit off_r), then the non-empty ones go
// to this lane's LDS column of the mask table (tab: entry t's mask at words
// 2t, 2t + 1 and its first slot at word 18 + t, stride 64), which may lie
// over the stream words: they are all read by then.  A ninth entry can only
// be row 8 (entries are the non-empty rows in order), so its first slot stays
// in a register (rb[8]) and the table is 26 words per lane: 6,656 bytes per
// wave, six 256-work-item workgroups in a CU's 160 KiB of LDS (five at 27).
// The take loop's next_row is then one table read.  With SKIP_SELF the
// particle's own slot k is cleared from row 4 (dx = dy = 0), its own cell's
// run.
constexpr uint32_t kMaskTabWords = 26 * 64;  // per wave
template <int NB, bool SKIP_SELF, int AUXN, typename Body>
__device__ __forceinline__ void gather_bits(const SceneDesc &S, const float4 *__restrict__ xps,
                                            const float4 *__restrict__ aux,
                                            const uint32_t *__restrict__ bounds, uint32_t k,
                                            const float4 &pi, uint64_t lens, const uint32_t *mw,
                                            uint32_t *tab, Body &&body)
{
    const uint32_t gy = S.gdim[1], gz = S.gdim[2];
    const uint32_t ux = (uint32_t)cell_coord(pi.x, S.h) - (uint32_t)S.cell_lo[0];
    const uint32_t uy = (uint32_t)cell_coord(pi.y, S.h) - (uint32_t)S.cell_lo[1];
    const uint32_t uz = (uint32_t)cell_coord(pi.z, S.h) - (uint32_t)S.cell_lo[2];
    const uint32_t wz0 = grid_wrap(uz - 1u, gz, S.gmod[2]);
    uint32_t wx[3], wy[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        wx[d] = grid_wrap(ux + (uint32_t)(d - 1), S.gdim[0], S.gmod[0]);
        wy[d] = grid_wrap(uy + (uint32_t)(d - 1), gy, S.gmod[1]);
    }
    uint32_t rb[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) rb[r] = bounds[S.key_off + (wx[r / 3] * gy + wy[r % 3]) * gz + wz0];
    uint64_t rm[9];
    uint32_t off = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const uint32_t L = (uint32_t)(lens >> (7 * r)) & 127u;
        // words past the stream's end only feed bits that the length mask
        // drops (off + L <= kMaskBitsMax, and off may be kMaskBitsMax when
        // L = 0): clamp
        const uint32_t q = off >> 5, sh = off & 31u;
        const uint32_t w0 = mw[min(q, kMaskWords - 1u) * 64];
        const uint32_t w1 = mw[min(q + 1u, kMaskWords - 1u) * 64];
        const uint32_t w2 = mw[min(q + 2u, kMaskWords - 1u) * 64];
        const uint32_t lo = sh ? (w0 >> sh) | (w1 << (32u - sh)) : w0;
        const uint32_t hi = sh ? (w1 >> sh) | (w2 << (32u - sh)) : w1;
        rm[r] = (((uint64_t)hi << 32) | lo) & (L >= 64u ? ~0ull : (1ull << L) - 1ull);
        off += L;
    }
    if (SKIP_SELF) rm[4] &= ~(1ull << (k - rb[4]));
    uint32_t rows = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        tab[(2 * rows) * 64] = (uint32_t)rm[r];
        tab[(2 * rows + 1) * 64] = (uint32_t)(rm[r] >> 32);
        if (r < 8) tab[(18 + rows) * 64] = rb[r];  // rows <= r here
        else if (rows < 8) tab[(18 + rows) * 64] = rb[8];
        rows += rm[r] != 0 ? 1u : 0u;
    }
    const uint32_t rb8 = rb[8];
    uint32_t t = 0;  // next table entry
    take_masks<NB, AUXN>(xps, aux, k, pi,
                         [&](uint64_t &mask, uint32_t &base) {
                             if (t < rows) {
                                 mask = (uint64_t)tab[(2 * t) * 64] | ((uint64_t)tab[(2 * t + 1) * 64] << 32);
                                 const uint32_t b = tab[(18 + min(t, 7u)) * 64];
                                 base = t < 8u ? b : rb8;
                                 ++t;
                             }
                         },
                         body);
}

// The mask consumers (pressure, viscosity): the record's stream words go to
// this lane's LDS column of the wave's table (tab), then gather_bits.
template <int NB, bool SKIP_SELF, int AUXN, typename Body>
__device__ __forceinline__ void gather_record(const SceneDesc &S, const float4 *__restrict__ xps,
                                              const float4 *__restrict__ aux,
                                              const uint32_t *__restrict__ bounds,
                                              const uint32_t *__restrict__ buf, uint32_t n, uint32_t k,
                                              const float4 &pi, uint64_t lens, uint32_t *tab, Body &&body)
{
    const uint4 *bits = reinterpret_cast<const uint4 *>(buf);
    const uint4 a = bits[k], c = bits[(size_t)n + k];
    tab[0] = a.x;
    tab[64] = a.y;
    tab[128] = a.z;
    tab[192] = a.w;
    tab[256] = c.x;
    tab[320] = c.y;
    tab[384] = c.z;
    tab[448] = c.w;
    uint32_t total = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) total += (uint32_t)(lens >> (7 * r)) & 127u;
    if (total > 256u) {
        const uint4 e = bits[2 * (size_t)n + k];
        tab[512] = e.x;
        tab[576] = e.y;
        tab[640] = e.z;
        tab[704] = e.w;
    }
    if (total > 384u) {
        const uint4 e = bits[3 * (size_t)n + k];
        tab[768] = e.x;
        tab[832] = e.y;
        tab[896] = e.z;
        tab[960] = e.w;
    }
    gather_bits<NB, SKIP_SELF, AUXN>(S, xps, aux, bounds, k, pi, lens, tab, tab, body);
}

template <bool GRID, bool SKIP_SELF, bool AUX, bool MASKS = false, bool DEFER = false,
          bool MASK_ONLY = false, typename Body>
__device__ __forceinline__ uint32_t walk(const SceneDesc &S, const float4 *__restrict__ xps,
                                     const float4 *__restrict__ aux,
                                     const uint32_t *__restrict__ start,
                                     const uint32_t *__restrict__ endt, uint32_t k,
                                     const float4 &pi, float4 *stage, bool clean, Body &&body,
                                     MaskOut *mo = nullptr)
{
    if constexpr (GRID)
        return walk_grid<SKIP_SELF, AUX, MASKS, DEFER, MASK_ONLY>(S, xps, aux, start, endt, k, pi, stage, clean,
                                                                 body, mo);
    else walk_hash<SKIP_SELF, AUX>(S, xps, aux, start, endt, k, pi, body);
    return 0u;
}

// Gather kernels (density, pressure, viscosity), in result-identical
// variants picked per workload by the autotuner (sph_engine.cpp):
//
// WORK: work-item i handles particle work[i] (ids sorted by the Morton code
//   of their cell, rebuilt every few updates) at sorted slot rank[id], so
//   neighbouring work-items own neighbouring cells and read largely the same
//   buckets, which stay in L2; workgroups are renumbered so each XCD takes one
//   contiguous run of the work order.  Without WORK, work-item i handles
//   sorted slot i (cells in hash order).
// LIST: two phases per particle -- the walk only tests candidates and
//   records the passing ones in a per-lane list in LDS (kListMax entries);
//   the arithmetic then runs in a dense loop over the list, so lanes do not
//   idle through the heavy body while other lanes test candidates.  The list
//   keeps walk order (sums keep the reference's order); a particle with more
//   than kListMax neighbours finishes with a second walk that skips the
//   listed ones.  Without LIST the body runs inside the walk.
// NBR (row mode): the density pass records each particle's neighbours and
//   pressure and viscosity take them from that record instead of walking:
//   64-entry neighbour rows under the reference key (gather_row), a 48-byte
//   record of each particle's row masks under the grid key (gather_record).
// GRID: the cell-grid bucket key (walk_grid), sorted slots in grid order.
// Which work-item handles a particle, and whether the body runs in the walk,
// from the list or from the row-mode record, never changes a result.
constexpr uint32_t kListMax = 56;

// GRID: work-item i handles sorted slot i (cells in grid order, already
//   spatially coherent), with the XCD-aware workgroup numbering.
// The grid covers the engine's capacity (a multiple of 8 workgroups); the
// numbering spreads the n live slots' workgroups over all 8 XCDs (block b
// runs on XCD b % 8), and the workgroups past them get whole blocks of
// indices >= n (block numbers from `live` on, below 2^32 / blockDim.x since
// the grid's capacity is < 2^31).  One scene: XCD x takes the contiguous
// live blocks [x * per, (x + 1) * per), so each XCD's L2 holds one slab of
// the grid.  Several scenes (MULTI): chunks of kXcdChunk consecutive blocks
// are dealt to the XCDs in turn (the blocks past the last whole chunks
// contiguously), so every XCD gets a sample of every part of the layout --
// scenes differ in cost per particle (a compressed pool against a falling
// block), and one XCD per slab would leave the others idle while it ends.
constexpr uint32_t kXcdChunk = 128;
template <bool WORK, bool GRID, bool MULTI = false>
__device__ __forceinline__ uint32_t gather_index(uint32_t n, uint32_t b)
{
    if (!WORK && !GRID) return b * blockDim.x + threadIdx.x;
    const uint32_t live = (n + blockDim.x - 1) / blockDim.x;
    const uint32_t per = (live + 7u) >> 3, x = b & 7u, j = b >> 3;
    uint32_t lb;
    if (j >= per) {
        lb = live + j;
    } else if (!MULTI) {
        lb = x * per + j;
    } else {
        const uint32_t C = max(1u, min(kXcdChunk, per >> 3));
        const uint32_t full = per / C * C;
        lb = j < full ? ((j / C) * 8u + x) * C + j % C : 8u * full + x * (per - full) + (j - full);
    }
    return lb * blockDim.x + threadIdx.x;
}

// Runs body(m, ox, oy, oz, d2, aux) over the neighbours of slot k with d < h
// (self excluded when SKIP_SELF), in walk order.  With LIST the candidate
// tests and the bodies are split into the two phases described above; `lst`
// is this work-item's column of the block's LDS list (stride blockDim.x).
template <bool GRID, bool LIST, bool SKIP_SELF, bool AUX, bool MASKS = false, typename Body>
__device__ __forceinline__ void gather(const SceneDesc &S, const float4 *__restrict__ xps,
                                       const float4 *__restrict__ aux,
                                       const uint32_t *__restrict__ start,
                                       const uint32_t *__restrict__ endt, uint32_t *lst,
                                       float4 *stage, bool clean, uint32_t k, const float4 &pi,
                                       Body &&body, MaskOut *mo = nullptr)
{
    if (!LIST) {
        walk<GRID, SKIP_SELF, AUX, MASKS>(S, xps, aux, start, endt, k, pi, stage, clean, body, mo);
        return;
    }
    const uint32_t stride = blockDim.x;
    uint32_t c = 0;
    walk<GRID, SKIP_SELF, false, MASKS>(S, xps, nullptr, start, endt, k, pi, stage, clean,
                           [&](uint32_t m, float, float, float, float, const float4 &) {
                               if (c < kListMax) lst[c * stride] = m;
                               ++c;
                           }, mo);
    const uint32_t n1 = min(c, kListMax);
    for (uint32_t j = 0; j < n1; j += KB) {
        uint32_t m[KB];
        float4 r[KB], a[KB];
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            m[q] = k;
            if (j + q < n1) m[q] = lst[(j + q) * stride];
        }
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            r[q] = pi;
            a[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (j + q < n1) {
                r[q] = xps[m[q]];
                if (AUX) a[q] = aux[m[q]];
            }
        }
#pragma unroll
        for (int q = 0; q < KB; ++q)
            if (j + q < n1) {
                const float ox = r[q].x - pi.x, oy = r[q].y - pi.y, oz = r[q].z - pi.z;
                body(m[q], ox, oy, oz, ox * ox + oy * oy + oz * oz, a[q]);
            }
    }
    if (c > kListMax) {
        uint32_t skip = kListMax;
        walk<GRID, SKIP_SELF, AUX>(S, xps, aux, start, endt, k, pi, stage, clean,
                             [&](uint32_t m, float ox, float oy, float oz, float d2,
                                 const float4 &a) {
                                 if (skip) --skip;
                                 else body(m, ox, oy, oz, d2, a);
                             });
    }
}

// LDS words per wave of a gather kernel past its lists: the staging planes,
// with neighbour masks (density) the stream words and the mask table over
// them, or (mask consumers) the table alone.
__host__ __device__ constexpr uint32_t gather_wave_words(bool grid, bool nbr, bool consumer)
{
    return consumer ? kMaskTabWords
                    : !grid ? 0u
                            : !nbr ? 4 * kStageBuf
                                   : (4 * kStageBuf + kMaskWords * 64 > kMaskTabWords ? 4 * kStageBuf + kMaskWords * 64
                                                                                       : kMaskTabWords);
}

// Orders every lane's use of the wave's mask table before the walk of the
// particles whose record is unusable, which stages rows over the table: the
// wave's lanes have reconverged here, and the release/acquire pair at
// wavefront scope keeps the table's LDS reads ahead of the staging writes.
#define SPH_TABLE_DONE()                                                            \
    do {                                                                            \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");                      \
        __builtin_amdgcn_wave_barrier();                                            \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");                      \
    } while (0)

#define SPH_GATHER_PROLOGUE()                                                       \
    extern __shared__ uint32_t lds_list[];                                          \
    const uint32_t i = gather_index<WORK, GRID, MULTI>(ed->n_total, vblock_);      \
    if (i >= ed->n_total) return;                                                   \
    const uint32_t g = WORK ? work[i] : svals[i];                                   \
    const uint32_t k_slot_ = WORK ? rank[g] : i;                                    \
    const float4 pi = xps[k_slot_];                                                 \
    uint32_t *lst = lds_list + threadIdx.x;                                         \
    /* GRID: each wave's LDS follows the lists (gather_wave_words words).    */  \
    /* The walking passes (CONSUMER false) stage rows there; with neighbour */  \
    /* masks the density pass's stream words (kMaskWords x 64) follow the   */  \
    /* staging, and its mask table lies over both.  The mask consumers hold */  \
    /* only the table; the particles whose record is unusable walk after   */  \
    /* the wave's table users are done, staging rows over the table.       */  \
    uint32_t *wave_lds = lds_list + (LIST ? kListMax * blockDim.x : 0) +            \
                         (threadIdx.x >> 6) * gather_wave_words(GRID, NBR, CONSUMER); \
    float4 *stage = (GRID && !CONSUMER) ? reinterpret_cast<float4 *>(wave_lds) : nullptr; \
    float4 *fallback_stage = GRID ? reinterpret_cast<float4 *>(wave_lds) : nullptr;    \
    uint32_t *mwords = wave_lds + 4 * kStageBuf + (threadIdx.x & 63u);            \
    uint32_t *mtab = wave_lds + (threadIdx.x & 63u);                               \
    /* the rest of the kernel, for scene S (closed by SPH_GATHER_EPILOGUE,   */  \
    /* which runs it through run_scene)                                      */  \
    auto scene_body_ = [&](const SceneDesc &S) __attribute__((always_inline)) {    \
    /* mask consumers: the slot, opaque to the optimiser at each pass of     */  \
    /* run_scene's loop, so the record addresses derived from it are formed  */  \
    /* where used, not hoisted out of the loop and kept live (spilled)       */  \
    uint32_t k = k_slot_;                                                           \
    if (CONSUMER) asm volatile("" : "+v"(k));                                       \
    /* grid key: no particle of this scene has its cell outside the grid's box */ \
    const bool clean = GRID && sflags[S.index] == 0u;                              \
    (void)mtab;                                                                     \
    (void)fallback_stage;                                                           \
    (void)mwords;                                                                   \
    (void)clean;                                                                    \
    (void)g;                                                                        \
    (void)lst

#define SPH_GATHER_EPILOGUE()                                                       \
    }                                                                               \
    ;                                                                               \
    run_scene<MULTI>(sd, scene_of, k_slot_, scene_body_)

#define SPH_GATHER_FWD ed, sd, scene_of, work, rank, svals, xps, start, endt, nbr, ncnt, sflags

#define SPH_GATHER_PARAMS                                                                     \
    const EngineDesc *__restrict__ ed, const SceneDesc *__restrict__ sd,                      \
        const uint32_t *__restrict__ scene_of, const uint32_t *__restrict__ work,             \
        const uint32_t *__restrict__ rank, const uint32_t *__restrict__ svals,                \
        const float4 *__restrict__ xps, const uint32_t *__restrict__ start,                   \
        const uint32_t *__restrict__ endt, uint32_t *__restrict__ nbr,                        \
        uint32_t *__restrict__ ncnt, const uint32_t *__restrict__ sflags

// NBR (with WORK): the density pass also stores each particle's neighbours
// with d < h (self excluded, walk order) as row i = work index of the
// particle: nbr[i * kNbrMax + j], count in ncnt[i].  A wave's rows form one
// contiguous block, so the writes stay in L2 until whole.  Pressure and
// viscosity then loop over the row instead of walking the buckets; their
// neighbour reads are spatially local in the work order.  A particle with
// more than kNbrMax neighbours walks the buckets instead.
constexpr uint32_t kNbrMax = 64;

template <typename Body>
__device__ __forceinline__ void gather_row(const SceneDesc &S, const float4 *__restrict__ xps,
                                           const float4 *__restrict__ aux,
                                           const uint32_t *__restrict__ start,
                                           const uint32_t *__restrict__ endt,
                                           const uint32_t *__restrict__ row, uint32_t c,
                                           uint32_t k, const float4 &pi, Body &&body)
{
    if (c > kNbrMax) {
        walk<false, true, true>(S, xps, aux, start, endt, k, pi, nullptr, false, body);
        return;
    }
    const uint4 *row4 = reinterpret_cast<const uint4 *>(row);
    for (uint32_t j = 0; j < c; j += 4) {
        const uint4 mm = row4[j >> 2];
        const uint32_t m[4] = {mm.x, mm.y, mm.z, mm.w};
        float4 r[4], a[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            r[q] = pi;
            a[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (j + q < c) {
                r[q] = xps[m[q]];
                a[q] = aux[m[q]];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (j + q < c) {
                const float ox = r[q].x - pi.x, oy = r[q].y - pi.y, oz = r[q].z - pi.z;
                body(m[q], ox, oy, oz, ox * ox + oy * oy + oz * oz, a[q]);
            }
    }
}

// Density of the slots of (virtual) workgroup vblock_ of the launch.
template <bool MULTI, bool WORK, bool LIST, bool NBR, bool GRID>
__device__ __forceinline__ void density_vblock(uint32_t vblock_, SPH_GATHER_PARAMS, float4 *__restrict__ dens)
{
    constexpr bool CONSUMER = false;
    SPH_GATHER_PROLOGUE();
    const float h = S.h, w2c = S.w2c, w3c = S.w3c;
    float rho = 0.0f, rhon = 0.0f;
    constexpr bool ROWS = NBR && !GRID;   // hash key: neighbour rows
    constexpr bool MASKS = NBR && GRID;   // grid key: neighbour masks
    uint32_t *row = ROWS ? nbr + (size_t)i * kNbrMax : nullptr;
    uint32_t c = 0;
    auto body = [&](uint32_t m, float, float, float, float d2, const float4 &) {
        const float d = sqrtf(d2);
        const float v = h - d;
        rho += v * v * w2c;       // SmoothingPow2, kernels.h:31
        rhon += v * v * v * w3c;  // SmoothingPow3, kernels.h:43
        if (ROWS && m != k) {
            if (c < kNbrMax) row[c] = m;
            ++c;
        }
    };
    // MASKS: the walk records the neighbour masks (the stream in this lane's
    // LDS column, the record in the neighbour-row buffer) and the bodies then
    // run over all of them, decoded from LDS into the mask table, which lies
    // over the wave's staging planes (free once every lane's walk is done);
    // a particle whose record is unusable walks again with the body once the
    // wave's other lanes are done with their tables, staging its rows there.
    // Grid key without masks or the LDS list: the body runs over each row's
    // mask as the walk goes (DEFER).
    MaskOut mo{nbr, ed->n_total, mwords, 0};
    if (GRID && !LIST && MASKS) {
        const uint32_t ok = walk<GRID, false, false, true, false, true>(
            S, xps, nullptr, start, endt, k, pi, stage, clean,
            [](uint32_t, float, float, float, float, const float4 &) {}, &mo);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (ok) gather_bits<4, false, 0>(S, xps, nullptr, start, k, pi, mo.lens, mwords, mtab, body);
        SPH_TABLE_DONE();
        if (!ok) walk<GRID, false, false, false, true>(S, xps, nullptr, start, endt, k, pi, stage, clean, body);
    } else if (GRID && !LIST)
        walk<GRID, false, false, false, true>(S, xps, nullptr, start, endt, k, pi, stage, clean, body);
    else
        gather<GRID, LIST, false, false, MASKS>(S, xps, nullptr, start, endt, lst, stage, clean, k, pi, body,
                                                MASKS ? &mo : nullptr);
    if (ROWS) ncnt[i] = c;
    // pressures of cc:371-372 / 407-408, computed once per particle
    // the pressure pass divides by these densities: their reciprocals ride along
    dens[k] = make_float4(rho, rhon, reciprocal_or_zero(rho), reciprocal_or_zero(rhon));
    SPH_GATHER_EPILOGUE();
}

// The second __launch_bounds__ argument is the minimum waves per SIMD the
// register allocation must allow: the single-scene density pass fits 64
// VGPRs (8 waves); the multi-scene density pass keeps the 7 it reaches
// unconstrained.
template <bool MULTI, bool WORK, bool LIST, bool NBR, bool GRID>
__global__ __launch_bounds__(256, MULTI ? 6 : 8) void k_density(SPH_GATHER_PARAMS, float4 *__restrict__ dens)
{
    density_vblock<MULTI, WORK, LIST, NBR, GRID>(blockIdx.x, SPH_GATHER_FWD, dens);
}

// Density pass of the neighbour-row mode.  Each work-item stages its row in
// LDS (row-major with a 65-word pitch, so the 64 lanes writing entry c of
// their rows hit 32 distinct banks); after a barrier the workgroup writes its
// rows -- one contiguous block of global memory -- with coalesced 16-byte
// stores, skipping the chunks beyond each row's count.
constexpr uint32_t kRowPitch = kNbrMax + 1;

template <bool MULTI>
__global__ __launch_bounds__(256) void k_density_rows(SPH_GATHER_PARAMS, float4 *__restrict__ dens)
{
    extern __shared__ uint32_t lds_rows[];  // [blockDim.x][kRowPitch], then counts[blockDim.x]
    const uint32_t B = blockDim.x, t = threadIdx.x;
    const uint32_t i = gather_index<true, false>(ed->n_total, blockIdx.x);
    const uint32_t n = ed->n_total;
    uint32_t *my = lds_rows + t * kRowPitch;
    uint32_t *cnt = lds_rows + B * kRowPitch;
    uint32_t c = 0;
    if (i < n) {
        const uint32_t k = rank[work[i]];
        const SceneDesc &S = scene_of_slot<MULTI>(sd, scene_of, k);
        const float4 pi = xps[k];
        const float h = S.h, w2c = S.w2c, w3c = S.w3c;
        float rho = 0.0f, rhon = 0.0f;
        walk<false, false, false>(S, xps, nullptr, start, endt, k, pi, nullptr, false,
                           [&](uint32_t m, float, float, float, float d2, const float4 &) {
            const float d = sqrtf(d2);
            const float v = h - d;
            rho += v * v * w2c;       // SmoothingPow2, kernels.h:31
            rhon += v * v * v * w3c;  // SmoothingPow3, kernels.h:43
            if (m != k) {
                if (c < kNbrMax) my[c] = m;
                ++c;
            }
        });
        dens[k] = make_float4(rho, rhon, reciprocal_or_zero(rho), reciprocal_or_zero(rhon));
        ncnt[i] = c;
    }
    cnt[t] = c;
    __syncthreads();
    const uint32_t i0 = i - t;  // first work index of this workgroup
    uint4 *out = reinterpret_cast<uint4 *>(nbr + (size_t)i0 * kNbrMax);
    for (uint32_t idx = t; idx < B * (kNbrMax / 4); idx += B) {
        const uint32_t r = idx / (kNbrMax / 4), q = idx % (kNbrMax / 4);
        if (i0 + r >= n || 4 * q >= min(cnt[r], kNbrMax)) continue;
        const uint32_t *src = lds_rows + r * kRowPitch + 4 * q;
        out[idx] = make_uint4(src[0], src[1], src[2], src[3]);
    }
}

template <bool MULTI, bool WORK, bool LIST, bool NBR, bool GRID>
__device__ __forceinline__ void pressure_vblock(uint32_t vblock_, SPH_GATHER_PARAMS,
                                                const float4 *__restrict__ dens,
                                                const float4 *__restrict__ vs, float4 *__restrict__ vs2)
{
    constexpr bool CONSUMER = NBR && GRID;  // reads the neighbour masks
    SPH_GATHER_PROLOGUE();
    const float4 di = dens[k];
    const float h = S.h, dw2c = S.dw2c, dw3c = S.dw3c, rho0 = S.rho0, kp = S.k, kn = S.k_near;
    const float P = (di.x - rho0) * kp, Pn = di.y * kn;  // cc:371-372
    // the quotients by d need only d's range when the coordinate differences
    // to every neighbour are either zero or of magnitude >= 2^-84
    const int near_ok = (int)coords_clear_of_zero(pi.x, pi.y, pi.z);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    // Every division is correctly rounded.  dj = (rho_j, rho_near_j,
    // RN(1/rho_j), RN(1/rho_near_j)); the six divisions by rho_j / rho_near_j
    // use those reciprocals and the three by d use RN(1/d), through
    // quotient() (exact in its stated range), and fall back to IEEE division
    // outside that range
    auto body = [&](uint32_t, float ox, float oy, float oz, float d2, const float4 &dj) {
        const float pj = (dj.x - rho0) * kp, pnj = dj.y * kn;  // cc:407-408
        const float sp = (P + pj) * 0.5f;     // cc:410
        const float snp = (Pn + pnj) * 0.5f;  // cc:411
        const float d = sqrtf(d2);
        float dirx = 0.0f, diry = 1.0f, dirz = 0.0f;  // cc:414
        if (d > 0.0f) {
            if ((near_ok & (int)divisor_ok(d)) != 0) {
                const float yd = recip_rn(d);
                dirx = quotient(ox, d, yd);
                diry = quotient(oy, d, yd);
                dirz = quotient(oz, d, yd);
            } else {
                dirx = ox / d;
                diry = oy / d;
                dirz = oz / d;
            }
        }
        const float v = h - d;
        const float a = -v * dw2c;      // SmoothingDerivativePow2, kernels.h:55
        const float b = -v * v * dw3c;  // SmoothingDerivativePow3, kernels.h:67
        const float tx = dirx * a * sp, ty = diry * a * sp, tz = dirz * a * sp;
        const float ux = dirx * b * snp, uy = diry * b * snp, uz = dirz * b * snp;
        const bool fast = ((int)(dj.z != 0.0f) & (int)(dj.w != 0.0f) &
                           (int)quotients_ok(tx, ty, tz, ux, uy, uz)) != 0;
        if (fast) {
            fx += quotient(tx, dj.x, dj.z);  // cc:416
            fy += quotient(ty, dj.x, dj.z);
            fz += quotient(tz, dj.x, dj.z);
            fx += quotient(ux, dj.y, dj.w);  // cc:417
            fy += quotient(uy, dj.y, dj.w);
            fz += quotient(uz, dj.y, dj.w);
        } else {
            fx += tx / dj.x;
            fy += ty / dj.x;
            fz += tz / dj.x;
            fx += ux / dj.y;
            fy += uy / dj.y;
            fz += uz / dj.y;
        }
    };
    if (NBR && GRID) {
        const uint2 mt2 = reinterpret_cast<const uint2 *>(nbr)[kMaskMetaAt * (size_t)ed->n_total + k];
        const uint64_t meta = (uint64_t)mt2.x | ((uint64_t)mt2.y << 32);
        const bool usable = !(meta & kMaskUnusable);
        if (usable) gather_record<KB, true, 4>(S, xps, dens, start, nbr, ed->n_total, k, pi, meta, mtab, body);
        SPH_TABLE_DONE();
        if (!usable) gather<GRID, false, true, true>(S, xps, dens, start, endt, lst, fallback_stage, clean, k, pi, body);
    } else if (NBR) {
        gather_row(S, xps, dens, start, endt, nbr + (size_t)i * kNbrMax, ncnt[i], k, pi, body);
    } else {
        gather<GRID, LIST, true, true>(S, xps, dens, start, endt, lst, stage, clean, k, pi, body);
    }
    float4 v = vs[k];
    const float dt = S.dt, rho = di.x;
    v.x = v.x + (fx / rho) * dt;  // cc:421
    v.y = v.y + (fy / rho) * dt;
    v.z = v.z + (fz / rho) * dt;
    vs2[k] = v;
    SPH_GATHER_EPILOGUE();
}

// The single-scene mask consumer is held to 80 VGPRs (6 waves per SIMD,
// what the LDS mask table allows; 82 unconstrained, 5 waves): its few spills
// lie off the take loop (-0.4% to -1.2% in A/B runs).  The multi-scene one
// keeps its 90 (at 80 its spills cost 5%).
template <bool MULTI, bool WORK, bool LIST, bool NBR, bool GRID>
__global__ __launch_bounds__(256, (NBR && GRID && !MULTI) ? 6 : 1) void k_pressure(
    SPH_GATHER_PARAMS, const float4 *__restrict__ dens, const float4 *__restrict__ vs, float4 *__restrict__ vs2)
{
    pressure_vblock<MULTI, WORK, LIST, NBR, GRID>(blockIdx.x, SPH_GATHER_FWD, dens, vs, vs2);
}

// Writes the new state at the sorted slot (pos, vel, ids: the next resident
// order) and records each particle's bucket for the next update's sort, with
// this update's dt (record_bucket: keys, ticks, count, wkeys); rpos and
// colors (may be null) are the render buffers, in particle order.
#define SPH_VISCOSITY_PARAMS                                                                  \
    const float4 *__restrict__ vs2, const float4 *__restrict__ ps, float4 *__restrict__ pos,         \
        float4 *__restrict__ vel, uint32_t *__restrict__ ids, float4 *__restrict__ rpos,            \
        float4 *__restrict__ colors, uint32_t *__restrict__ keys, uint32_t *__restrict__ ticks,     \
        uint32_t *__restrict__ count, uint32_t *__restrict__ wkeys
#define SPH_VISCOSITY_FWD vs2, ps, pos, vel, ids, rpos, colors, keys, ticks, count, wkeys

template <bool MULTI, bool WORK, bool LIST, bool NBR, bool GRID>
__device__ __forceinline__ void viscosity_vblock(uint32_t vblock_, SPH_GATHER_PARAMS, SPH_VISCOSITY_PARAMS)
{
    constexpr bool CONSUMER = NBR && GRID;  // reads the neighbour masks
    SPH_GATHER_PROLOGUE();
    const float4 vi = vs2[k];
    const float4 x0 = ps[k];  // issued early: used after the gather
    const float hh = S.hh, c6 = S.poly6c;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    auto body = [&](uint32_t, float, float, float, float d2, const float4 &vj) {
        const float d = sqrtf(d2);
        const float v = hh - d * d;          // kernels.h:78
        const float w = v * v * v * c6;      // kernels.h:79
        fx += (vj.x - vi.x) * w;             // cc:460
        fy += (vj.y - vi.y) * w;
        fz += (vj.z - vi.z) * w;
    };
    if (NBR && GRID) {
        const uint2 mt2 = reinterpret_cast<const uint2 *>(nbr)[kMaskMetaAt * (size_t)ed->n_total + k];
        const uint64_t meta = (uint64_t)mt2.x | ((uint64_t)mt2.y << 32);
        const bool usable = !(meta & kMaskUnusable);
        if (usable) gather_record<4, true, 3>(S, xps, vs2, start, nbr, ed->n_total, k, pi, meta, mtab, body);
        SPH_TABLE_DONE();
        if (!usable) gather<GRID, false, true, true>(S, xps, vs2, start, endt, lst, fallback_stage, clean, k, pi, body);
    } else if (NBR) {
        gather_row(S, xps, vs2, start, endt, nbr + (size_t)i * kNbrMax, ncnt[i], k, pi, body);
    } else {
        gather<GRID, LIST, true, true>(S, xps, vs2, start, endt, lst, stage, clean, k, pi, body);
    }
    const float dt = S.dt, mu = S.mu;
    float4 v = vi;
    v.x = vi.x + fx * mu * dt;  // cc:463
    v.y = vi.y + fy * mu * dt;
    v.z = vi.z + fz * mu * dt;
    // stage F (cc:84-107), per particle
    float4 x = x0;
    x.x += v.x * dt;
    x.y += v.y * dt;
    x.z += v.z * dt;
    if (S.half[0] - fabsf(x.x) <= 0.0f) { x.x = S.half[0] * sign_f(x.x); v.x *= S.damp_neg; }
    if (S.half[1] - fabsf(x.y) <= 0.0f) { x.y = S.half[1] * sign_f(x.y); v.y *= S.damp_neg; }
    if (S.half[2] - fabsf(x.z) <= 0.0f) { x.z = S.half[2] * sign_f(x.z); v.z *= S.damp_neg; }
    x.w = 0.34f;  // OutPositions[i] = vec4(positions[i], 0.34f), cc:107
    pos[k] = x;
    vel[k] = v;
    ids[k] = g;
    if (rpos) rpos[g] = x;
    if (colors) {
        // render hand-off: the speed colour of the new velocity (fluidSimCPU.cc:100-125)
        float c[4];
        speed_color(v.x, v.y, v.z, ed->stops, c);
        colors[g] = make_float4(c[0], c[1], c[2], c[3]);
    }
    // stage A + B1 of the next update, from the new state
    record_bucket(*ed, S, predict_particle(S, x, v), true, k, keys, ticks, count, wkeys);
    SPH_GATHER_EPILOGUE();
}

// (5 waves per SIMD: 96 VGPRs, which the single-scene form fits in 76 and
// the multi-scene form in exactly 96, without spilling)
template <bool MULTI, bool WORK, bool LIST, bool NBR, bool GRID>
__global__ __launch_bounds__(256, 5) void k_viscosity_integrate(SPH_GATHER_PARAMS,
                                                                           SPH_VISCOSITY_PARAMS)
{
    viscosity_vblock<MULTI, WORK, LIST, NBR, GRID>(blockIdx.x, SPH_GATHER_FWD, SPH_VISCOSITY_FWD);
}

// ---------------------------------------------------------------------------
// Fused steady update: one persistent launch per Update
// ---------------------------------------------------------------------------
// The launch holds s.fused_grid workgroups of 256 -- no more than the device
// keeps resident at once (occupancy x compute units, a multiple of 8) -- and
// runs the six stages of a steady update (buckets recorded by the previous
// update's viscosity pass) in order, with a grid barrier between stages: the
// count scan by tickets (a workgroup waits only on chunks taken before its
// own), scatter and reorder over the particles, then the three gathers over
// the same virtual workgroups the separate launches use (virtual workgroup
// v runs on workgroup v mod grid, so on the XCD the numbering assumes).
// Each stage is the same device code as its own kernel, so results are
// identical.  Barrier waits are bounded (kFusedBarrierSpin polls 64 x 64
// clocks apart, about 0.1 s): a wait that reaches the bound raises scan_err
// (the host then reports the update invalid) and the launch still ends.

constexpr uint32_t kFusedBlock = 256;
constexpr uint32_t kFusedBarrierSpin = 1u << 16;  // polls

// Every workgroup arrives, in two levels so that no counter takes more than
// 64 arrivals: workgroup b on sub-counter b mod 64 (each on a cache line of
// its own), the last member of each sub-counter on the top counter; the last
// arrival there resets it and advances the generation, and every other
// workgroup waits for the generation to change (polling every 64 x 64
// clocks).  Each counter is reset by its last arrival before the generation
// moves, so all are zero at the next barrier and after the launch.
// Agent-scope release and acquire publish each workgroup's stores to the
// other XCDs' L2s.  bar: [0] top count, [1] generation, [64 + 16 q]
// sub-count q (kFusedBarWords words).
__device__ __forceinline__ void grid_barrier(uint32_t *bar, uint32_t nwg, uint32_t *err)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t gen = __hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t q = blockIdx.x & 63u;
        const uint32_t members = (nwg >> 6) + (q < (nwg & 63u) ? 1u : 0u);
        const uint32_t subs = nwg < 64u ? nwg : 64u;
        uint32_t *sub = bar + 64 + 16 * q;
        bool last = false;
        if (__hip_atomic_fetch_add(sub, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u == members) {
            __hip_atomic_store(sub, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u == subs;
        }
        if (last) {
            __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (++spins > kFusedBarrierSpin) {
                    *err = 1u;
                    break;
                }
                __builtin_amdgcn_s_sleep(64);
            }
        }
    }
    __syncthreads();
}
static_assert(64 + 16 * 63 < kFusedBarWords, "grid barrier words");

template <bool MULTI>
__global__ __launch_bounds__(kFusedBlock) void k_fused_update(const DevState s)
{
    __shared__ uint32_t chunk_s;
    const uint32_t nwg = gridDim.x, b = blockIdx.x, tid = threadIdx.x;
    // count scan: chunks by ticket (the counter starts at zero; the reorder
    // below clears it for the next update)
    const uint32_t na = scan_chunks_live(s.ed, s.count_blocks);
    for (;;) {
        __syncthreads();  // the previous chunk's shared values are read
        if (tid == 0) chunk_s = atomicAdd(reinterpret_cast<uint32_t *>(s.chain + s.count_blocks), 1u);
        __syncthreads();
        const uint32_t c = chunk_s;
        if (c >= na) break;
        scan_chunk(c, na, s.ed, s.count, s.chain, s.bounds, s.sflags, s.scan_err);
    }
    grid_barrier(s.fbar, nwg, s.scan_err);
    scatter_range(b * kFusedBlock + tid, nwg * kFusedBlock, s.ed, s.keys[0], s.keys[1], s.ids, s.bounds,
                  s.order);
    grid_barrier(s.fbar, nwg, s.scan_err);
    reorder_blocks<MULTI>(b, nwg, s.ed, s.sd, s.scene_of, s.keys[0], s.bounds, s.order, s.sids, s.vel, s.pos,
                          s.xps, s.vs, s.ps, nullptr, s.count, s.chain, s.count_blocks + 1);
    grid_barrier(s.fbar, nwg, s.scan_err);
    // the gathers over the virtual workgroups of the capacity's grid
    const uint32_t vblocks = (((s.capacity + kFusedBlock - 1) / kFusedBlock) + 7u) & ~7u;
    const uint32_t *const endt = s.bounds + 1;
    for (uint32_t v = b; v < vblocks; v += nwg)
        density_vblock<MULTI, false, false, true, true>(v, s.ed, s.sd, s.scene_of, nullptr, nullptr, s.sids,
                                                        s.xps, s.bounds, endt, s.nbr, s.ncnt, s.sflags, s.dens);
    grid_barrier(s.fbar, nwg, s.scan_err);
    for (uint32_t v = b; v < vblocks; v += nwg)
        pressure_vblock<MULTI, false, false, true, true>(v, s.ed, s.sd, s.scene_of, nullptr, nullptr, s.sids,
                                                         s.xps, s.bounds, endt, s.nbr, s.ncnt, s.sflags, s.dens,
                                                         s.vs, s.vs2);
    grid_barrier(s.fbar, nwg, s.scan_err);
    for (uint32_t v = b; v < vblocks; v += nwg)
        viscosity_vblock<MULTI, false, false, true, true>(v, s.ed, s.sd, s.scene_of, nullptr, nullptr, s.sids,
                                                          s.xps, s.bounds, endt, s.nbr, s.ncnt, s.sflags, s.vs2,
                                                          s.ps, s.pos, s.vel, s.ids, s.rpos, s.colors, s.keys[0],
                                                          s.keys[1], s.count, nullptr);
}

// dynamic LDS of the fused update's workgroup: the density pass's per-wave
// staging, stream words and mask table (the consumers' table is no larger)
constexpr size_t kFusedLds = sizeof(uint32_t) * kMaskTabWords * (kFusedBlock / 64);
static_assert(gather_wave_words(true, true, false) <= kMaskTabWords &&
                  gather_wave_words(true, true, true) <= kMaskTabWords,
              "the fused update's LDS holds every gather stage's");

// ---------------------------------------------------------------------------
// Download / parity helpers (not on the step path)
// ---------------------------------------------------------------------------
__global__ void k_unpermute_dens(const EngineDesc *__restrict__ ed,
                                 const uint32_t *__restrict__ svals,
                                 const float4 *__restrict__ dens, float2 *__restrict__ out)
{
    const uint32_t n = ed->n_total;
    for (uint32_t k = global_id(); k < n; k += grid_stride()) {
        const float4 d = dens[k];
        out[svals[k]] = make_float2(d.x, d.y);
    }
}

// Same walk as the reference's density loop, but without the d < h cut: the
// reported set is every record passing the key, hash and d2 <= r2 tests.
template <typename Body>
__device__ __forceinline__ void walk_prefilter(const SceneDesc &S, const float4 *__restrict__ xps,
                                               const uint32_t *__restrict__ start,
                                               const uint32_t *__restrict__ endt,
                                               const float4 &pi, Body &&body)
{
    const int cx = cell_coord(pi.x, S.h), cy = cell_coord(pi.y, S.h), cz = cell_coord(pi.z, S.h);
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const uint32_t hash = hash_cell(cx + dx, cy + dy, cz + dz);
                const uint32_t key =
                    S.key_off + (S.grid ? grid_key(cx + dx, cy + dy, cz + dz, S.cell_lo, S.gdim, S.gmod)
                                        : fastmod(hash, S.mod_m, S.table));
                const uint32_t b = start[key], e = endt[key];
                const float hf = (float)hash;
                for (uint32_t m = b; m < e; ++m) {
                    const float4 r = xps[m];
                    if (r.w != hf) continue;
                    const float ox = r.x - pi.x, oy = r.y - pi.y, oz = r.z - pi.z;
                    const float d2 = ox * ox + oy * oy + oz * oz;
                    if (d2 > S.r2) continue;
                    body(m);
                }
            }
}

template <bool MULTI>
__global__ void k_neighbour_count(const EngineDesc *__restrict__ ed,
                                  const SceneDesc *__restrict__ sd,
                                  const uint32_t *__restrict__ scene_of,
                                  const float4 *__restrict__ xps,
                                  const uint32_t *__restrict__ start,
                                  const uint32_t *__restrict__ endt,
                                  const uint32_t *__restrict__ svals, uint32_t *__restrict__ counts)
{
    const uint32_t n = ed->n_total;
    for (uint32_t k = global_id(); k < n; k += grid_stride()) {
        const SceneDesc &S = scene_of_slot<MULTI>(sd, scene_of, k);
        uint32_t c = 0;
        walk_prefilter(S, xps, start, endt, xps[k], [&](uint32_t) { ++c; });
        counts[svals[k]] = c;
    }
}

template <bool MULTI>
__global__ void k_neighbour_fill(const EngineDesc *__restrict__ ed,
                                 const SceneDesc *__restrict__ sd,
                                 const uint32_t *__restrict__ scene_of,
                                 const float4 *__restrict__ xps,
                                 const uint32_t *__restrict__ start,
                                 const uint32_t *__restrict__ endt,
                                 const uint32_t *__restrict__ svals,
                                 const uint64_t *__restrict__ offs, uint32_t *__restrict__ list)
{
    const uint32_t n = ed->n_total;
    for (uint32_t k = global_id(); k < n; k += grid_stride()) {
        const SceneDesc &S = scene_of_slot<MULTI>(sd, scene_of, k);
        const uint32_t g = svals[k];
        uint64_t o = offs[g];
        const uint32_t p0 = S.p_off;
        walk_prefilter(S, xps, start, endt, xps[k], [&](uint32_t m) { list[o++] = svals[m] - p0; });
    }
}

// Initial state generated on the device: the same per-particle function as
// the host generator (sph_generate.h), so the bytes are identical.
__global__ void k_generate(const GenPlan g, float4 *__restrict__ pos, float4 *__restrict__ vel)
{
    for (uint32_t q = global_id(); q < g.n; q += grid_stride()) {
        float p[4], v[4];
        generate_particle(g, q, p, v);
        pos[q] = make_float4(p[0], p[1], p[2], p[3]);
        vel[q] = make_float4(v[0], v[1], v[2], v[3]);
    }
}

// Floats with bit patterns in [lo, hi] whose recip_rn differs from IEEE 1/x:
// their count and the smallest such pattern.
__global__ void k_recip_check(uint32_t lo, uint32_t hi, unsigned long long *__restrict__ bad,
                              uint32_t *__restrict__ first_bad)
{
    uint32_t c = 0;
    for (uint64_t u = (uint64_t)lo + global_id(); u <= hi; u += grid_stride()) {
        const float x = __uint_as_float((uint32_t)u);
        if (__float_as_uint(recip_rn(x)) != __float_as_uint(1.0f / x)) {
            ++c;
            atomicMin(first_bad, (uint32_t)u);
        }
    }
    if (c) atomicAdd(bad, (unsigned long long)c);
}

// Plain streaming copy: the reference point for HBM byte counters.
__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
    for (size_t i = global_id(); i < n; i += grid_stride()) dst[i] = src[i];
}

__global__ void k_count_nonfinite(const EngineDesc *__restrict__ ed,
                                  const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                  uint32_t *__restrict__ counter)
{
    const uint32_t n = ed->n_total;
    uint32_t c = 0;
    for (uint32_t g = global_id(); g < n; g += grid_stride()) {
        const float4 x = pos[g], v = vel[g];
        c += !isfinite(x.x) + !isfinite(x.y) + !isfinite(x.z);
        c += !isfinite(v.x) + !isfinite(v.y) + !isfinite(v.z);
    }
    if (c) atomicAdd(counter, c);
}

// Blocks for a grid-stride kernel over `n` items: enough to fill the chip
// (256 CUs x 8 blocks), never more than the items need.
inline uint32_t grid_for(uint32_t n, int block)
{
    const uint32_t need = (n + block - 1) / block;
    const uint32_t cap = 256u * 32u;
    return need < 1 ? 1 : (need > cap ? cap : need);
}

}  // namespace

#define SPH_LAUNCH_CHECK() return hipGetLastError()

hipError_t launch_predict(const DevState &s, const LaunchCfg &c, hipStream_t st)
{
    const uint32_t g = grid_for(s.capacity, c.block_stream);
    uint32_t *const wk = s.grid ? nullptr : s.wkeys[0];  // no work order under the grid key
    k_count_clear<<<grid_for(s.count_size, c.block_stream), c.block_stream, 0, st>>>(s.count,
                                                                                    s.count_size);
    if (s.multi)
        k_predict<true><<<g, c.block_stream, 0, st>>>(s.ed, s.sd, s.scene_of, s.pos, s.vel,
                                                       s.keys[0], s.keys[1], s.count, wk);
    else
        k_predict<false><<<g, c.block_stream, 0, st>>>(s.ed, s.sd, s.scene_of, s.pos, s.vel,
                                                        s.keys[0], s.keys[1], s.count, wk);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_radix_sort(const DevState &s, uint32_t *const keys[2], uint32_t *const vals[2],
                             const uint32_t *first_vals, uint32_t bits, hipStream_t st,
                             int *result_buf)
{
    int cur = 0;
    const uint32_t nblk = s.radix_blocks;
    for (uint32_t shift = 0; shift < bits; shift += kRadixBits) {
        k_radix_hist<<<nblk, kSortBlock, 0, st>>>(s.ed, keys[cur], shift, s.hist);
        k_radix_scan_digit<<<kRadix, kSortBlock, 0, st>>>(s.hist, nblk, s.hist + kRadix * nblk);
        // first pass: values from first_vals, or the input slot indices
        k_radix_scatter<<<nblk, kSortBlock, 0, st>>>(s.ed, keys[cur], shift ? vals[cur] : first_vals,
                                                      keys[cur ^ 1],
                                                      vals[cur ^ 1], shift, s.hist,
                                                      s.hist + kRadix * nblk);
        cur ^= 1;
    }
    *result_buf = cur;
    SPH_LAUNCH_CHECK();
}

hipError_t launch_bucket_sort(const DevState &s, const LaunchCfg &c, hipStream_t st)
{
    const uint32_t nb = s.count_blocks, g = grid_for(s.capacity, c.block_stream);
    k_count_scan<<<nb, kSortBlock, 0, st>>>(s.ed, s.count, s.chain, s.bounds, s.sflags,
                                             s.scan_err);
    k_count_scatter<<<g, c.block_stream, 0, st>>>(s.ed, s.keys[0], s.keys[1], s.ids, s.bounds,
                                                   s.order);
    uint32_t *const rk = s.grid ? nullptr : s.rank;
    const uint32_t gr = grid_for(s.capacity, kReorderBlock);
    if (s.multi)
        k_count_reorder<true><<<gr, kReorderBlock, 0, st>>>(s.ed, s.sd, s.scene_of, s.keys[0],
                                                             s.bounds, s.order, s.sids, s.vel,
                                                             s.pos, s.xps, s.vs, s.ps, rk,
                                                             s.count, s.chain, nb + 1);
    else
        k_count_reorder<false><<<gr, kReorderBlock, 0, st>>>(s.ed, s.sd, s.scene_of, s.keys[0],
                                                              s.bounds, s.order, s.sids, s.vel,
                                                              s.pos, s.xps, s.vs, s.ps, rk,
                                                              s.count, s.chain, nb + 1);
    SPH_LAUNCH_CHECK();
}

// One work-item per slot; the grid is a multiple of 8 workgroups (XCD
// renumbering of the WORK variants).
inline uint32_t gather_grid(uint32_t cap, int block)
{
    const uint32_t blocks = (cap + block - 1) / block;
    return (blocks + 7u) & ~7u;
}

// Launches kernel template KER<MULTI, WORK, LIST, NBR, GRID> chosen at run time.
#define SPH_GATHER_LAUNCH(KER, var, blk, st, ...)                                             \
    do {                                                                                      \
        const GatherVariant v_ = (var);                                                       \
        const int b_ = (blk);                                                                 \
        const uint32_t g_ = gather_grid(s.capacity, b_);                                      \
        const size_t lds_ = v_.list ? sizeof(uint32_t) * kListMax * (size_t)b_ : 0;           \
        const bool m_ = s.multi;                                                              \
        if (s.grid && c.nbr && v_.rows) {                                                     \
            const size_t ldsr_ = sizeof(uint32_t) * kMaskTabWords * (size_t)(b_ / 64);       \
            if (m_) KER<true, false, false, true, true><<<g_, b_, ldsr_, st>>>(__VA_ARGS__);  \
            else KER<false, false, false, true, true><<<g_, b_, ldsr_, st>>>(__VA_ARGS__);    \
        } else if (s.grid) {                                                                  \
            const size_t ldsg_ = lds_ + sizeof(uint32_t) * gather_wave_words(true, false, false) * (size_t)(b_ / 64); \
            if (m_ && v_.list) KER<true, false, true, false, true><<<g_, b_, ldsg_, st>>>(__VA_ARGS__);  \
            else if (m_) KER<true, false, false, false, true><<<g_, b_, ldsg_, st>>>(__VA_ARGS__);       \
            else if (v_.list) KER<false, false, true, false, true><<<g_, b_, ldsg_, st>>>(__VA_ARGS__);  \
            else KER<false, false, false, false, true><<<g_, b_, ldsg_, st>>>(__VA_ARGS__);              \
        } else if (c.nbr && v_.rows) {                                                        \
            if (m_) KER<true, true, false, true, false><<<g_, b_, 0, st>>>(__VA_ARGS__);      \
            else KER<false, true, false, true, false><<<g_, b_, 0, st>>>(__VA_ARGS__);        \
        } else if (m_ && v_.work && v_.list)                                                  \
            KER<true, true, true, false, false><<<g_, b_, lds_, st>>>(__VA_ARGS__);           \
        else if (m_ && v_.work) KER<true, true, false, false, false><<<g_, b_, lds_, st>>>(__VA_ARGS__);   \
        else if (m_ && v_.list) KER<true, false, true, false, false><<<g_, b_, lds_, st>>>(__VA_ARGS__);   \
        else if (m_) KER<true, false, false, false, false><<<g_, b_, lds_, st>>>(__VA_ARGS__);             \
        else if (v_.work && v_.list) KER<false, true, true, false, false><<<g_, b_, lds_, st>>>(__VA_ARGS__); \
        else if (v_.work) KER<false, true, false, false, false><<<g_, b_, lds_, st>>>(__VA_ARGS__);        \
        else if (v_.list) KER<false, false, true, false, false><<<g_, b_, lds_, st>>>(__VA_ARGS__);        \
        else KER<false, false, false, false, false><<<g_, b_, lds_, st>>>(__VA_ARGS__);                    \
    } while (0)

#define SPH_GATHER_ARGS                                                                \
    s.ed, s.sd, s.scene_of, s.work, s.rank, s.sids, s.xps, s.bounds, s.bounds + 1, s.nbr, s.ncnt, \
        s.sflags

hipError_t launch_density(const DevState &s, const LaunchCfg &c, hipStream_t st)
{
    if (c.nbr && s.grid) {
        // the density pass records the neighbour masks
        const int b = c.density.block;
        const uint32_t g = gather_grid(s.capacity, b);
        const size_t lds = (c.density.list ? sizeof(uint32_t) * kListMax * (size_t)b : 0) +
                           sizeof(uint32_t) * gather_wave_words(true, true, false) * (size_t)(b / 64);
        if (s.multi && c.density.list)
            k_density<true, false, true, true, true><<<g, b, lds, st>>>(SPH_GATHER_ARGS, s.dens);
        else if (s.multi)
            k_density<true, false, false, true, true><<<g, b, lds, st>>>(SPH_GATHER_ARGS, s.dens);
        else if (c.density.list)
            k_density<false, false, true, true, true><<<g, b, lds, st>>>(SPH_GATHER_ARGS, s.dens);
        else
            k_density<false, false, false, true, true><<<g, b, lds, st>>>(SPH_GATHER_ARGS, s.dens);
        SPH_LAUNCH_CHECK();
    }
    if (c.nbr && !s.grid) {
        const int b = c.density.block;
        const size_t lds = sizeof(uint32_t) * ((size_t)b * kRowPitch + b);
        if (s.multi)
            k_density_rows<true><<<gather_grid(s.capacity, b), b, lds, st>>>(SPH_GATHER_ARGS, s.dens);
        else
            k_density_rows<false><<<gather_grid(s.capacity, b), b, lds, st>>>(SPH_GATHER_ARGS, s.dens);
        SPH_LAUNCH_CHECK();
    }
    SPH_GATHER_LAUNCH(k_density, c.density, c.density.block, st, SPH_GATHER_ARGS, s.dens);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_pressure(const DevState &s, const LaunchCfg &c, hipStream_t st)
{
    SPH_GATHER_LAUNCH(k_pressure, c.pressure, c.pressure.block, st, SPH_GATHER_ARGS, s.dens, s.vs,
                      s.vs2);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_viscosity_integrate(const DevState &s, const LaunchCfg &c, hipStream_t st)
{
    SPH_GATHER_LAUNCH(k_viscosity_integrate, c.viscosity, c.viscosity.block, st, SPH_GATHER_ARGS,
                      s.vs2, s.ps, s.pos, s.vel, s.ids, s.rpos, s.colors, s.keys[0], s.keys[1],
                      s.count, s.grid ? nullptr : s.wkeys[0]);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_fused_update(const DevState &s, hipStream_t st)
{
    if (!s.grid || s.fused_grid == 0 || s.fused_grid % 8u != 0) return hipErrorInvalidConfiguration;
    if (s.multi) k_fused_update<true><<<s.fused_grid, kFusedBlock, kFusedLds, st>>>(s);
    else k_fused_update<false><<<s.fused_grid, kFusedBlock, kFusedLds, st>>>(s);
    SPH_LAUNCH_CHECK();
}

uint32_t fused_update_grid(bool multi, int device)
{
    int per_cu = 0, cus = 0;
    hipError_t e = multi ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused_update<true>, kFusedBlock,
                                                                         kFusedLds)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused_update<false>, kFusedBlock,
                                                                         kFusedLds);
    if (e != hipSuccess || per_cu <= 0) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        return 0;
    return (uint32_t)(per_cu * cus) & ~7u;
}

hipError_t launch_unpermute_dens(const DevState &s, float2 *out, hipStream_t st)
{
    k_unpermute_dens<<<grid_for(s.capacity, 256), 256, 0, st>>>(s.ed, s.sids, s.dens, out);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_unpermute_state(const DevState &s, uint32_t off, uint32_t n, float4 *dst_pos,
                                  float4 *dst_vel, hipStream_t st)
{
    k_unpermute_state<<<grid_for(n, 256), 256, 0, st>>>(s.pos + off, s.vel + off, s.ids + off, n,
                                                        off, dst_pos, dst_vel);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_iota(uint32_t *a, uint32_t n, uint32_t first, hipStream_t st)
{
    k_iota<<<grid_for(n, 256), 256, 0, st>>>(a, n, first);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_scene_map(const SceneDesc *sd, uint32_t *scene_of, uint32_t n_scenes, uint32_t max_n,
                            hipStream_t st)
{
    if (n_scenes == 0) return hipSuccess;
    const uint32_t bx = grid_for(max_n, 256) < 1024u ? grid_for(max_n, 256) : 1024u;
    k_scene_map<<<dim3(bx, n_scenes), 256, 0, st>>>(sd, scene_of);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_clear_words(uint32_t *a, uint32_t n, hipStream_t st)
{
    k_count_clear<<<grid_for(n, 256), 256, 0, st>>>(a, n);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_neighbour_count(const DevState &s, uint32_t *counts, hipStream_t st)
{
    const uint32_t g = grid_for(s.capacity, 256);
    if (s.multi)
        k_neighbour_count<true><<<g, 256, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.bounds, s.bounds + 1,
                                                   s.sids, counts);
    else
        k_neighbour_count<false><<<g, 256, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.bounds, s.bounds + 1,
                                                    s.sids, counts);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_neighbour_fill(const DevState &s, const uint64_t *offs,
                                 uint32_t *list, hipStream_t st)
{
    const uint32_t g = grid_for(s.capacity, 256);
    if (s.multi)
        k_neighbour_fill<true><<<g, 256, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.bounds, s.bounds + 1,
                                                  s.sids, offs, list);
    else
        k_neighbour_fill<false><<<g, 256, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.bounds, s.bounds + 1,
                                                   s.sids, offs, list);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_generate(const GenPlan &g, float4 *pos, float4 *vel, hipStream_t st)
{
    k_generate<<<grid_for(g.n, 256), 256, 0, st>>>(g, pos, vel);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_recip_check(uint32_t lo, uint32_t hi, unsigned long long *bad,
                              uint32_t *first_bad, hipStream_t st)
{
    k_recip_check<<<4096, 256, 0, st>>>(lo, hi, bad, first_bad);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_copy(const float4 *src, float4 *dst, size_t n, hipStream_t st)
{
    k_copy<<<2048, 256, 0, st>>>(src, dst, n);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_count_nonfinite(const DevState &s, uint32_t *counter, hipStream_t st)
{
    k_count_nonfinite<<<grid_for(s.capacity, 256), 256, 0, st>>>(s.ed, s.pos, s.vel, counter);
    SPH_LAUNCH_CHECK();
}

}  // namespace sph
