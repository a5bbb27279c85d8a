// sph_kernels.hip -- hand-written gfx950 kernels of the SPH step.
//
// Reference: Allkams/Fluid-Simulation-3D, engine/physics/physicsWorld.cc
// (FluidSimulation::Update, cc:39-111) and engine/physics/kernels.h.
//
// Pipeline of one Update(dt) over all scenes of an engine:
//   k_predict          stage A + B1  (cc:42-48, 473-482), id order
//   k_fill             bucket table <- INT_MAX (cc:481)
//   k_radix_hist/scan/scatter  stage B2 (cc:484): stable LSD radix sort of
//                      (key, slot); stability gives (key, particle index) order
//   k_reorder          stage B3 (cc:486-496) + gather of xp, v into sorted order
//   k_density          stage C (cc:304-311, 325-365)
//   k_pressure         stage D (cc:62-69, 367-422)
//   k_viscosity_integrate  stage E (cc:71-78, 424-464, Jacobi) + F (cc:80-108)
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
// buckets, which stay resident in the per-XCD L2.
#include <hip/hip_runtime.h>

#include "sph_device.h"
#include "sph_launch.h"

namespace sph {

namespace {

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortBlock = 256;  // 4 waves
constexpr int kSortWaves = kSortBlock / 64;
constexpr int kScanBlock = 1024;

__device__ __forceinline__ uint32_t grid_stride() { return gridDim.x * blockDim.x; }
__device__ __forceinline__ uint32_t global_id() { return blockIdx.x * blockDim.x + threadIdx.x; }

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
template <bool MULTI>
__global__ void k_predict(const EngineDesc *__restrict__ ed, const SceneDesc *__restrict__ sd,
                          const uint32_t *__restrict__ scene_of, const float4 *__restrict__ pos,
                          float4 *__restrict__ vel, float4 *__restrict__ xp,
                          uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    const uint32_t n = ed->n_total;
    for (uint32_t g = global_id(); g < n; g += grid_stride()) {
        const SceneDesc &S = scene_of_slot<MULTI>(sd, scene_of, g);
        const float dt = S.dt;
        float4 v = vel[g];
        const float4 x = pos[g];
        // cc:45: velocity += CalculateExternalFoce(...) * deltatime
        v.x = v.x + S.g[0] * dt;
        v.y = v.y + S.g[1] * dt;
        v.z = v.z + S.g[2] * dt;
        // cc:47: predictedPositions = positions + velocity * (1/120)
        const float px = x.x + v.x * S.lookahead;
        const float py = x.y + v.y * S.lookahead;
        const float pz = x.z + v.z * S.lookahead;
        const uint32_t hash =
            hash_cell(cell_coord(px, S.h), cell_coord(py, S.h), cell_coord(pz, S.h));
        vel[g] = v;
        xp[g] = make_float4(px, py, pz, __uint_as_float(hash));
        keys[g] = S.key_off + fastmod(hash, S.mod_m, S.table);
        vals[g] = g;
    }
}

__global__ void k_fill(const EngineDesc *__restrict__ ed, uint32_t *__restrict__ a, uint32_t v)
{
    const uint32_t n = ed->table_total;
    for (uint32_t i = global_id(); i < n; i += grid_stride()) a[i] = v;
}

// ---------------------------------------------------------------------------
// Stage B2: stable LSD radix sort.  A fixed grid of `nblk` blocks; block b owns
// the contiguous chunk [b*chunk, min((b+1)*chunk, n)) and walks it tile by
// tile in order, so equal digits keep their input order (stability).
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

// Exclusive scan of `m` counters in place, one block of 1024 work-items,
// 8 consecutive counters per work-item per round.
constexpr int kScanItems = 8;
__global__ __launch_bounds__(kScanBlock) void k_scan_exclusive(uint32_t *__restrict__ a, uint32_t m)
{
    __shared__ uint32_t wsum[kScanBlock / 64];
    __shared__ uint32_t carry_s;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < m; base += kScanBlock * kScanItems) {
        const uint32_t i0 = base + tid * kScanItems;
        uint32_t v[kScanItems];
        uint32_t tsum = 0;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            v[j] = i0 + j < m ? a[i0 + j] : 0u;
            tsum += v[j];
        }
        // inclusive wave scan of the per-item sums
        uint32_t x = tsum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        if (w == 0) {
            uint32_t s = lane < kScanBlock / 64 ? wsum[lane] : 0;
#pragma unroll
            for (int off = 1; off < kScanBlock / 64; off <<= 1) {
                const uint32_t y = __shfl_up(s, off, 64);
                if (lane >= (uint32_t)off) s += y;
            }
            if (lane < kScanBlock / 64) wsum[lane] = s;  // inclusive over waves
        }
        __syncthreads();
        const uint32_t carry = carry_s;
        uint32_t run = carry + (w ? wsum[w - 1] : 0) + x - tsum;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            if (i0 + j < m) a[i0 + j] = run;
            run += v[j];
        }
        __syncthreads();
        if (tid == kScanBlock - 1) carry_s = carry + wsum[kScanBlock / 64 - 1];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(
    const EngineDesc *__restrict__ ed, const uint32_t *__restrict__ kin,
    const uint32_t *__restrict__ vin, uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
    uint32_t shift, const uint32_t *__restrict__ hist)
{
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t wcnt[kSortWaves][kRadix];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (uint32_t d = tid; d < kRadix; d += kSortBlock) {
        base[d] = hist[d * gridDim.x + blockIdx.x];
        for (int q = 0; q < kSortWaves; ++q) wcnt[q][d] = 0;
    }
    __syncthreads();
    uint32_t beg, end;
    chunk_of(ed->n_total, beg, end);
    const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (uint32_t t0 = beg; t0 < end; t0 += kSortBlock) {
        const uint32_t e = t0 + tid;
        const bool valid = e < end;
        const uint32_t key = valid ? kin[e] : 0u;
        const uint32_t val = valid ? vin[e] : 0u;
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
// Stage B3 (cc:486-496) + gather into sorted order
// ---------------------------------------------------------------------------
__global__ void k_reorder(const EngineDesc *__restrict__ ed, const uint32_t *__restrict__ skeys,
                          const uint32_t *__restrict__ svals, const float4 *__restrict__ xp,
                          const float4 *__restrict__ vel, float4 *__restrict__ xps,
                          float4 *__restrict__ vs, uint32_t *__restrict__ start,
                          uint32_t *__restrict__ endt)
{
    const uint32_t n = ed->n_total;
    for (uint32_t k = global_id(); k < n; k += grid_stride()) {
        const uint32_t key = skeys[k];
        const uint32_t g = svals[k];
        const float4 r = xp[g];
        // the reference's lookup record keeps the hash as a float (cc:480)
        xps[k] = make_float4(r.x, r.y, r.z, (float)__float_as_uint(r.w));
        vs[k] = vel[g];
        const uint32_t prev = k ? skeys[k - 1] : 0xffffffffu;
        const uint32_t next = k + 1 < n ? skeys[k + 1] : 0xffffffffu;
        if (key != prev) start[key] = k;
        if (key != next) endt[key] = k + 1;
    }
}

// ---------------------------------------------------------------------------
// Neighbour walk (cc:331-363, 378-419, 432-462): 27 cells in the reference's
// offset order (physicsWorld.h:131-143: dx outer, dz inner); within a bucket
// in sorted order; records of other cells sharing the bucket are skipped by
// the float-rounded hash test.  `body(m, ox, oy, oz, d2)` runs for every
// record with d2 <= r2 (the sqrRadius prefilter) and d < h; records with
// d >= h contribute exactly zero to every sum of the reference.
// ---------------------------------------------------------------------------
template <bool SKIP_SELF, typename Body>
__device__ __forceinline__ void walk(const SceneDesc &S, const float4 *__restrict__ xps,
                                     const uint32_t *__restrict__ start,
                                     const uint32_t *__restrict__ endt, uint32_t k,
                                     const float4 &pi, Body &&body)
{
    const int cx = cell_coord(pi.x, S.h), cy = cell_coord(pi.y, S.h), cz = cell_coord(pi.z, S.h);
#pragma unroll 1
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll 1
            for (int dz = -1; dz <= 1; ++dz) {
                const uint32_t hash = hash_cell(cx + dx, cy + dy, cz + dz);
                const uint32_t key = S.key_off + fastmod(hash, S.mod_m, S.table);
                const uint32_t b = start[key];
                if (b == SPH_START_EMPTY) continue;
                const uint32_t e = endt[key];
                const float hf = (float)hash;
                for (uint32_t m = b; m < e; ++m) {
                    const float4 r = xps[m];
                    if (r.w != hf) continue;
                    if (SKIP_SELF && m == k) continue;
                    const float ox = r.x - pi.x, oy = r.y - pi.y, oz = r.z - pi.z;
                    const float d2 = ox * ox + oy * oy + oz * oz;  // glm::dot order
                    if (d2 > S.r2) continue;
                    if (!(d2 < S.h2_lt)) continue;
                    body(m, ox, oy, oz, d2);
                }
            }
}

template <bool MULTI>
__global__ void k_density(const EngineDesc *__restrict__ ed, const SceneDesc *__restrict__ sd,
                          const uint32_t *__restrict__ scene_of, const float4 *__restrict__ xps,
                          const uint32_t *__restrict__ start, const uint32_t *__restrict__ endt,
                          float4 *__restrict__ dens)
{
    const uint32_t n = ed->n_total;
    for (uint32_t k = global_id(); k < n; k += grid_stride()) {
        const SceneDesc &S = scene_of_slot<MULTI>(sd, scene_of, k);
        const float4 pi = xps[k];
        const float h = S.h, w2c = S.w2c, w3c = S.w3c;
        float rho = 0.0f, rhon = 0.0f;
        walk<false>(S, xps, start, endt, k, pi, [&](uint32_t, float, float, float, float d2) {
            const float d = sqrtf(d2);
            const float v = h - d;
            rho += v * v * w2c;       // SmoothingPow2, kernels.h:31
            rhon += v * v * v * w3c;  // SmoothingPow3, kernels.h:43
        });
        // pressures of cc:371-372 / 407-408, computed once per particle
        dens[k] = make_float4(rho, rhon, (rho - S.rho0) * S.k, rhon * S.k_near);
    }
}

template <bool MULTI>
__global__ void k_pressure(const EngineDesc *__restrict__ ed, const SceneDesc *__restrict__ sd,
                           const uint32_t *__restrict__ scene_of, const float4 *__restrict__ xps,
                           const uint32_t *__restrict__ start, const uint32_t *__restrict__ endt,
                           const float4 *__restrict__ dens, const float4 *__restrict__ vs,
                           float4 *__restrict__ vs2)
{
    const uint32_t n = ed->n_total;
    for (uint32_t k = global_id(); k < n; k += grid_stride()) {
        const SceneDesc &S = scene_of_slot<MULTI>(sd, scene_of, k);
        const float4 pi = xps[k];
        const float4 di = dens[k];
        const float h = S.h, dw2c = S.dw2c, dw3c = S.dw3c;
        const float P = di.z, Pn = di.w;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        walk<true>(S, xps, start, endt, k, pi, [&](uint32_t m, float ox, float oy, float oz, float d2) {
            const float4 dj = dens[m];
            const float sp = (P + dj.z) * 0.5f;   // cc:410
            const float snp = (Pn + dj.w) * 0.5f; // cc:411
            const float d = sqrtf(d2);
            float dirx = 0.0f, diry = 1.0f, dirz = 0.0f;  // cc:414
            if (d > 0.0f) {
                dirx = ox / d;
                diry = oy / d;
                dirz = oz / d;
            }
            const float v = h - d;
            const float a = -v * dw2c;      // SmoothingDerivativePow2, kernels.h:55
            fx += dirx * a * sp / dj.x;     // cc:416
            fy += diry * a * sp / dj.x;
            fz += dirz * a * sp / dj.x;
            const float b = -v * v * dw3c;  // SmoothingDerivativePow3, kernels.h:67
            fx += dirx * b * snp / dj.y;    // cc:417
            fy += diry * b * snp / dj.y;
            fz += dirz * b * snp / dj.y;
        });
        float4 v = vs[k];
        const float dt = S.dt, rho = di.x;
        v.x = v.x + (fx / rho) * dt;  // cc:421
        v.y = v.y + (fy / rho) * dt;
        v.z = v.z + (fz / rho) * dt;
        vs2[k] = v;
    }
}

template <bool MULTI>
__global__ void k_viscosity_integrate(const EngineDesc *__restrict__ ed,
                                      const SceneDesc *__restrict__ sd,
                                      const uint32_t *__restrict__ scene_of,
                                      const float4 *__restrict__ xps,
                                      const uint32_t *__restrict__ start,
                                      const uint32_t *__restrict__ endt,
                                      const float4 *__restrict__ vs2,
                                      const uint32_t *__restrict__ svals,
                                      float4 *__restrict__ pos, float4 *__restrict__ vel)
{
    const uint32_t n = ed->n_total;
    for (uint32_t k = global_id(); k < n; k += grid_stride()) {
        const SceneDesc &S = scene_of_slot<MULTI>(sd, scene_of, k);
        const float4 pi = xps[k];
        const float4 vi = vs2[k];
        const float hh = S.hh, c6 = S.poly6c;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        walk<true>(S, xps, start, endt, k, pi, [&](uint32_t m, float, float, float, float d2) {
            const float d = sqrtf(d2);
            const float v = hh - d * d;          // kernels.h:78
            const float w = v * v * v * c6;      // kernels.h:79
            const float4 vj = vs2[m];
            fx += (vj.x - vi.x) * w;             // cc:460
            fy += (vj.y - vi.y) * w;
            fz += (vj.z - vi.z) * w;
        });
        const float dt = S.dt, mu = S.mu;
        float4 v = vi;
        v.x = vi.x + fx * mu * dt;  // cc:463
        v.y = vi.y + fy * mu * dt;
        v.z = vi.z + fz * mu * dt;
        // stage F (cc:84-107), per particle
        const uint32_t g = svals[k];
        float4 x = pos[g];
        x.x += v.x * dt;
        x.y += v.y * dt;
        x.z += v.z * dt;
        if (S.half[0] - fabsf(x.x) <= 0.0f) { x.x = S.half[0] * sign_f(x.x); v.x *= S.damp_neg; }
        if (S.half[1] - fabsf(x.y) <= 0.0f) { x.y = S.half[1] * sign_f(x.y); v.y *= S.damp_neg; }
        if (S.half[2] - fabsf(x.z) <= 0.0f) { x.z = S.half[2] * sign_f(x.z); v.z *= S.damp_neg; }
        x.w = 0.34f;  // OutPositions[i] = vec4(positions[i], 0.34f), cc:107
        pos[g] = x;
        vel[g] = v;
    }
}

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
                const uint32_t key = S.key_off + fastmod(hash, S.mod_m, S.table);
                const uint32_t b = start[key];
                if (b == SPH_START_EMPTY) continue;
                const uint32_t e = endt[key];
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
    if (s.multi)
        k_predict<true><<<g, c.block_stream, 0, st>>>(s.ed, s.sd, s.scene_of, s.pos, s.vel, s.xp,
                                                       s.keys[0], s.vals[0]);
    else
        k_predict<false><<<g, c.block_stream, 0, st>>>(s.ed, s.sd, s.scene_of, s.pos, s.vel, s.xp,
                                                        s.keys[0], s.vals[0]);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_clear_table(const DevState &s, const LaunchCfg &c, hipStream_t st)
{
    k_fill<<<grid_for(s.table_capacity, c.block_stream), c.block_stream, 0, st>>>(
        s.ed, s.start, SPH_START_EMPTY);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_radix_sort(const DevState &s, uint32_t bits, hipStream_t st, int *result_buf)
{
    int cur = 0;
    const uint32_t nblk = s.radix_blocks;
    for (uint32_t shift = 0; shift < bits; shift += kRadixBits) {
        k_radix_hist<<<nblk, kSortBlock, 0, st>>>(s.ed, s.keys[cur], shift, s.hist);
        k_scan_exclusive<<<1, kScanBlock, 0, st>>>(s.hist, nblk * kRadix);
        k_radix_scatter<<<nblk, kSortBlock, 0, st>>>(s.ed, s.keys[cur], s.vals[cur],
                                                      s.keys[cur ^ 1], s.vals[cur ^ 1], shift,
                                                      s.hist);
        cur ^= 1;
    }
    *result_buf = cur;
    SPH_LAUNCH_CHECK();
}

hipError_t launch_reorder(const DevState &s, const LaunchCfg &c, int buf, hipStream_t st)
{
    k_reorder<<<grid_for(s.capacity, c.block_stream), c.block_stream, 0, st>>>(
        s.ed, s.keys[buf], s.vals[buf], s.xp, s.vel, s.xps, s.vs, s.start, s.end);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_density(const DevState &s, const LaunchCfg &c, hipStream_t st)
{
    const uint32_t g = grid_for(s.capacity, c.block_gather);
    if (s.multi)
        k_density<true><<<g, c.block_gather, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.start,
                                                       s.end, s.dens);
    else
        k_density<false><<<g, c.block_gather, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.start,
                                                        s.end, s.dens);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_pressure(const DevState &s, const LaunchCfg &c, hipStream_t st)
{
    const uint32_t g = grid_for(s.capacity, c.block_gather);
    if (s.multi)
        k_pressure<true><<<g, c.block_gather, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.start,
                                                        s.end, s.dens, s.vs, s.vs2);
    else
        k_pressure<false><<<g, c.block_gather, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.start,
                                                         s.end, s.dens, s.vs, s.vs2);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_viscosity_integrate(const DevState &s, const LaunchCfg &c, int buf,
                                      hipStream_t st)
{
    const uint32_t g = grid_for(s.capacity, c.block_gather);
    if (s.multi)
        k_viscosity_integrate<true><<<g, c.block_gather, 0, st>>>(
            s.ed, s.sd, s.scene_of, s.xps, s.start, s.end, s.vs2, s.vals[buf], s.pos, s.vel);
    else
        k_viscosity_integrate<false><<<g, c.block_gather, 0, st>>>(
            s.ed, s.sd, s.scene_of, s.xps, s.start, s.end, s.vs2, s.vals[buf], s.pos, s.vel);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_unpermute_dens(const DevState &s, int buf, float2 *out, hipStream_t st)
{
    k_unpermute_dens<<<grid_for(s.capacity, 256), 256, 0, st>>>(s.ed, s.vals[buf], s.dens, out);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_neighbour_count(const DevState &s, int buf, uint32_t *counts, hipStream_t st)
{
    const uint32_t g = grid_for(s.capacity, 256);
    if (s.multi)
        k_neighbour_count<true><<<g, 256, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.start, s.end,
                                                   s.vals[buf], counts);
    else
        k_neighbour_count<false><<<g, 256, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.start, s.end,
                                                    s.vals[buf], counts);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_neighbour_fill(const DevState &s, int buf, const uint64_t *offs,
                                 uint32_t *list, hipStream_t st)
{
    const uint32_t g = grid_for(s.capacity, 256);
    if (s.multi)
        k_neighbour_fill<true><<<g, 256, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.start, s.end,
                                                  s.vals[buf], offs, list);
    else
        k_neighbour_fill<false><<<g, 256, 0, st>>>(s.ed, s.sd, s.scene_of, s.xps, s.start, s.end,
                                                   s.vals[buf], offs, list);
    SPH_LAUNCH_CHECK();
}

hipError_t launch_count_nonfinite(const DevState &s, uint32_t *counter, hipStream_t st)
{
    k_count_nonfinite<<<grid_for(s.capacity, 256), 256, 0, st>>>(s.ed, s.pos, s.vel, counter);
    SPH_LAUNCH_CHECK();
}

}  // namespace sph
