// sph_generate.h -- synthetic initial states, one particle at a time.
//
// Every value depends only on (seed, particle index, component), through a
// counter-based generator (splitmix64).  The per-particle function compiles
// for host and device from this one source, so the states generated on the
// GPU (sph_initialize, the server's generated requests) and on the host
// (sph_generate, used by the oracle's tests) are the same bytes.  Per-scene sizes
// (lattice dimensions) are planned once on the host.
#pragma once
#include <stdint.h>

#include "sph_device.h"

namespace sph {

struct GenPlan {
    int32_t kind;     // SPH_INIT_*
    uint32_t n;
    uint64_t seed;
    float box[3];
    float lo[3];      // -box/2
    float spacing, jitter, speed;
    uint32_t d0, d1, d2;  // lattice dimensions of the kind (see plan_generate)
    float c0;             // centred block / lattice origin
    float gap, wx, wy, wz;  // GridArrangement (physicsWorld.cc:518-557)
};

SPH_HD uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Uniform float in [0, 1) with 24 random bits, for (seed, particle, component).
SPH_HD float uniform01(uint64_t seed, uint64_t i, int c)
{
    const uint64_t r = splitmix64(splitmix64(seed) + i * 8u + (uint64_t)c);
    return (float)(r >> 40) * (1.0f / 16777216.0f);
}

SPH_HD float jit(uint64_t seed, uint64_t i, int c, float jitter)
{
    return jitter == 0.0f ? 0.0f : (2.0f * uniform01(seed, i, c) - 1.0f) * jitter;
}

// Position and velocity of particle q (float4, w = 0.34 for positions: the
// OutPositions layout of physicsWorld.cc:107; 0 for velocities).
SPH_HD void generate_particle(const GenPlan &g, uint32_t q, float pos[4], float vel[4])
{
    const float s = g.spacing;
    float x = 0.0f, y = 0.0f, z = 0.0f, vx = 0.0f;
    switch (g.kind) {
    case 0: {  // SPH_INIT_GRID_ARRANGEMENT: y, x, z loops, z fastest (cc:525-555)
        const uint32_t p = g.d0;
        const uint32_t lz = q % p, lx = (q / p) % p, ly = q / (p * p);
        const float xo = (float)(int)lx * g.gap, yo = (float)(int)ly * g.gap, zo = (float)(int)lz * g.gap;
        x = g.wx + xo;
        y = g.wy - yo;
        z = g.wz + zo;
        break;
    }
    case 1: {  // SPH_INIT_DAM_BREAK: block d0 x d1 x d2 (x, y, z), z fastest, at the -x,-y,-z corner
        const uint32_t ix = q / (g.d1 * g.d2), iy = (q / g.d2) % g.d1, iz = q % g.d2;
        x = g.lo[0] + ((float)ix + 0.5f) * s + jit(g.seed, q, 0, g.jitter);
        y = g.lo[1] + ((float)iy + 0.5f) * s + jit(g.seed, q, 1, g.jitter);
        z = g.lo[2] + ((float)iz + 0.5f) * s + jit(g.seed, q, 2, g.jitter);
        break;
    }
    case 2:  // SPH_INIT_UNIFORM_LOWER
        x = g.lo[0] + uniform01(g.seed, q, 0) * g.box[0];
        y = g.lo[1] + uniform01(g.seed, q, 1) * (0.5f * g.box[1]);
        z = g.lo[2] + uniform01(g.seed, q, 2) * g.box[2];
        break;
    case 3: {  // SPH_INIT_TWO_BLOCKS: mirrored blocks at the -x and +x walls, moving together
        const uint32_t ny = g.d1, nz = g.d2;
        const uint32_t half = (g.n + 1) / 2;
        const bool right = q >= half;
        const uint32_t r = right ? q - half : q;
        const uint32_t ix = r / (ny * nz), iy = (r / nz) % ny, iz = r % nz;
        const float xo = ((float)ix + 0.5f) * s;
        const float bx = right ? -g.lo[0] - xo : g.lo[0] + xo;
        const float by = g.lo[1] + ((float)iy + 0.5f) * s;
        const float bz = -0.5f * (float)nz * s + ((float)iz + 0.5f) * s;
        x = bx + jit(g.seed, q, 0, g.jitter);
        y = by + jit(g.seed, q, 1, g.jitter);
        z = bz + jit(g.seed, q, 2, g.jitter);
        vx = right ? -g.speed : g.speed;
        break;
    }
    default: {  // 4, SPH_INIT_CENTRED_BLOCK: cube of side d0, centred
        const uint32_t m = g.d0;
        const uint32_t ix = q / (m * m), iy = (q / m) % m, iz = q % m;
        x = g.c0 + (float)ix * s + jit(g.seed, q, 0, g.jitter);
        y = g.c0 + (float)iy * s + jit(g.seed, q, 1, g.jitter);
        z = g.c0 + (float)iz * s + jit(g.seed, q, 2, g.jitter);
        break;
    }
    }
    pos[0] = x;
    pos[1] = y;
    pos[2] = z;
    pos[3] = 0.34f;
    vel[0] = vx;
    vel[1] = 0.0f;
    vel[2] = 0.0f;
    vel[3] = 0.0f;
}

// Plans a scene (host): validates the arguments and fixes the per-scene
// sizes.  Returns 0 or SPH_ERR_INVALID.
int plan_generate(int kind, uint64_t seed, uint32_t n, const float box[3], float spacing,
                  float jitter, float speed, GenPlan *out);

}  // namespace sph
