// sph_generate.cpp -- synthetic initial states on the host (sph_generate) and
// the per-scene planning shared with the device generator (sph_generate.h).
#include "sph_generate.h"

#include <math.h>
#include <string.h>

#include "../../include/sph_amd.h"

namespace sph {

int plan_generate(int kind, uint64_t seed, uint32_t n, const float box[3], float spacing,
                  float jitter, float speed, GenPlan *g)
{
    if (n == 0 || !g) return SPH_ERR_INVALID;
    memset(g, 0, sizeof(*g));
    g->kind = kind;
    g->n = n;
    g->seed = seed;
    g->spacing = spacing;
    g->jitter = jitter;
    g->speed = speed;
    if (kind == SPH_INIT_GRID_ARRANGEMENT) {
        // InitializeData: GridArrangement(ceil(cbrt n), 0.215) (physicsWorld.cc:139-142, 518-557)
        const int per_axis = (int)ceilf(powf((float)n, (1.0f / 3.0f)));
        const float gap = 0.215f;
        const float total_w = per_axis * gap;
        g->d0 = (uint32_t)per_axis;
        g->gap = gap;
        g->wx = (0 - ((total_w - gap) / 2.0f));
        g->wy = (0 + (total_w - gap) / 2.0f);
        g->wz = (0 - (total_w - gap) / 2.0f);
        return SPH_OK;
    }
    if (!box || !(box[0] > 0) || !(box[1] > 0) || !(box[2] > 0)) return SPH_ERR_INVALID;
    for (int c = 0; c < 3; ++c) {
        g->box[c] = box[c];
        g->lo[c] = -0.5f * box[c];
    }
    const float s = spacing;
    switch (kind) {
    case SPH_INIT_DAM_BREAK: {
        if (!(s > 0)) return SPH_ERR_INVALID;
        uint32_t nx = 1;
        while ((uint64_t)nx * (2u * nx) * (2u * nx) < n) ++nx;
        g->d0 = nx;
        g->d1 = 2 * nx;
        g->d2 = 2 * nx;
        return SPH_OK;
    }
    case SPH_INIT_UNIFORM_LOWER:
        return SPH_OK;
    case SPH_INIT_TWO_BLOCKS: {
        if (!(s > 0)) return SPH_ERR_INVALID;
        uint32_t nz = (uint32_t)(0.96f * box[2] / s), ny = (uint32_t)(0.64f * box[1] / s);
        g->d1 = ny < 1 ? 1 : ny;
        g->d2 = nz < 1 ? 1 : nz;
        return SPH_OK;
    }
    case SPH_INIT_CENTRED_BLOCK: {
        if (!(s > 0)) return SPH_ERR_INVALID;
        uint32_t m = 1;
        while ((uint64_t)m * m * m < n) ++m;
        g->d0 = m;
        g->c0 = -0.5f * (float)(m - 1) * s;
        return SPH_OK;
    }
    default:
        return SPH_ERR_INVALID;
    }
}

}  // namespace sph

extern "C" int sph_generate(int kind, uint64_t seed, uint32_t n, const float box[3],
                            float spacing, float jitter, float speed, float *pos4, float *vel4)
{
    if (!pos4 || !vel4) return SPH_ERR_INVALID;
    sph::GenPlan g;
    const int rc = sph::plan_generate(kind, seed, n, box, spacing, jitter, speed, &g);
    if (rc) return rc;
    for (uint32_t q = 0; q < n; ++q)
        sph::generate_particle(g, q, pos4 + 4 * (size_t)q, vel4 + 4 * (size_t)q);
    return SPH_OK;
}
