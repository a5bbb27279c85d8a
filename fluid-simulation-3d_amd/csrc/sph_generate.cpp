// sph_generate.cpp -- synthetic initial states (host only).
//
// Every value depends only on (seed, particle index, component), through a
// counter-based generator (splitmix64), so the bytes do not depend on threads
// or call order and the GPU path and the CPU checkers see identical inputs.
#include <math.h>
#include <string.h>

#include "../../include/sph_amd.h"

namespace {

uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Uniform float in [0, 1) with 24 random bits, for (seed, particle, component).
float uniform01(uint64_t seed, uint64_t i, int c)
{
    const uint64_t r = splitmix64(splitmix64(seed) + i * 8u + (uint64_t)c);
    return (float)(r >> 40) * (1.0f / 16777216.0f);
}

float jit(uint64_t seed, uint64_t i, int c, float jitter)
{
    return jitter == 0.0f ? 0.0f : (2.0f * uniform01(seed, i, c) - 1.0f) * jitter;
}

void put(float *a, uint32_t i, float x, float y, float z, float w)
{
    a[4 * (size_t)i + 0] = x;
    a[4 * (size_t)i + 1] = y;
    a[4 * (size_t)i + 2] = z;
    a[4 * (size_t)i + 3] = w;
}

// The reference's InitializeData lattice: GridArrangement(ceil(cbrt n), 0.215)
// (physicsWorld.cc:139-142, 518-557): y, x, z loops, z fastest, centred.
void grid_arrangement(uint32_t n, float *pos4)
{
    const int per_axis = (int)ceilf(powf((float)n, (1.0f / 3.0f)));
    const float gap = 0.215f;
    const float total_w = per_axis * gap;
    for (uint32_t q = 0; q < n; ++q) put(pos4, q, 0.0f, 0.0f, 0.0f, 0.25f);  // cc:131-132
    uint32_t i = 0;
    for (int ly = 0; ly < per_axis; ly++)
        for (int lx = 0; lx < per_axis; lx++)
            for (int lz = 0; lz < per_axis; lz++) {
                if (i >= n) return;
                const float xo = lx * gap, yo = ly * gap, zo = lz * gap;
                const float wx = (0 - ((total_w - gap) / 2.0f));
                const float wy = (0 + (total_w - gap) / 2.0f);
                const float wz = (0 - (total_w - gap) / 2.0f);
                put(pos4, i, wx + xo, wy - yo, wz + zo, 0.34f);
                i++;
            }
}

}  // namespace

extern "C" int sph_generate(int kind, uint64_t seed, uint32_t n, const float box[3],
                            float spacing, float jitter, float speed, float *pos4, float *vel4)
{
    if (n == 0 || !pos4 || !vel4) return SPH_ERR_INVALID;
    for (uint32_t i = 0; i < n; ++i) put(vel4, i, 0.0f, 0.0f, 0.0f, 0.0f);
    if (kind == SPH_INIT_GRID_ARRANGEMENT) {
        grid_arrangement(n, pos4);
        return SPH_OK;
    }
    if (!box || !(box[0] > 0) || !(box[1] > 0) || !(box[2] > 0)) return SPH_ERR_INVALID;
    const float lo[3] = {-0.5f * box[0], -0.5f * box[1], -0.5f * box[2]};
    const float s = spacing;
    switch (kind) {
    case SPH_INIT_DAM_BREAK: {
        // block nx x 2nx x 2nx (x, y, z), z fastest, at the -x,-y,-z corner
        if (!(s > 0)) return SPH_ERR_INVALID;
        uint32_t nx = 1;
        while ((uint64_t)nx * (2u * nx) * (2u * nx) < n) ++nx;
        const uint32_t ny = 2 * nx, nz = 2 * nx;
        for (uint32_t q = 0; q < n; ++q) {
            const uint32_t ix = q / (ny * nz), iy = (q / nz) % ny, iz = q % nz;
            put(pos4, q, lo[0] + ((float)ix + 0.5f) * s + jit(seed, q, 0, jitter),
                lo[1] + ((float)iy + 0.5f) * s + jit(seed, q, 1, jitter),
                lo[2] + ((float)iz + 0.5f) * s + jit(seed, q, 2, jitter), 0.34f);
        }
        return SPH_OK;
    }
    case SPH_INIT_UNIFORM_LOWER:
        for (uint32_t q = 0; q < n; ++q)
            put(pos4, q, lo[0] + uniform01(seed, q, 0) * box[0],
                lo[1] + uniform01(seed, q, 1) * (0.5f * box[1]),
                lo[2] + uniform01(seed, q, 2) * box[2], 0.34f);
        return SPH_OK;
    case SPH_INIT_TWO_BLOCKS: {
        // two mirrored lattice blocks on the floor, at the -x and +x walls,
        // moving towards each other at `speed`
        if (!(s > 0)) return SPH_ERR_INVALID;
        uint32_t nz = (uint32_t)(0.96f * box[2] / s), ny = (uint32_t)(0.64f * box[1] / s);
        if (nz < 1) nz = 1;
        if (ny < 1) ny = 1;
        const uint32_t half = (n + 1) / 2;
        for (uint32_t q = 0; q < n; ++q) {
            const bool right = q >= half;
            const uint32_t r = right ? q - half : q;
            const uint32_t ix = r / (ny * nz), iy = (r / nz) % ny, iz = r % nz;
            const float xo = ((float)ix + 0.5f) * s;
            const float x = right ? -lo[0] - xo : lo[0] + xo;
            const float y = lo[1] + ((float)iy + 0.5f) * s;
            const float z = -0.5f * (float)nz * s + ((float)iz + 0.5f) * s;
            put(pos4, q, x + jit(seed, q, 0, jitter), y + jit(seed, q, 1, jitter),
                z + jit(seed, q, 2, jitter), 0.34f);
            put(vel4, q, right ? -speed : speed, 0.0f, 0.0f, 0.0f);
        }
        return SPH_OK;
    }
    case SPH_INIT_CENTRED_BLOCK: {
        if (!(s > 0)) return SPH_ERR_INVALID;
        uint32_t m = 1;
        while ((uint64_t)m * m * m < n) ++m;
        const float c0 = -0.5f * (float)(m - 1) * s;
        for (uint32_t q = 0; q < n; ++q) {
            const uint32_t ix = q / (m * m), iy = (q / m) % m, iz = q % m;
            put(pos4, q, c0 + (float)ix * s + jit(seed, q, 0, jitter),
                c0 + (float)iy * s + jit(seed, q, 1, jitter),
                c0 + (float)iz * s + jit(seed, q, 2, jitter), 0.34f);
        }
        return SPH_OK;
    }
    default:
        return SPH_ERR_INVALID;
    }
}
