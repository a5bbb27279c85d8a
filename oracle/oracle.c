/*
 * oracle.c -- CPU restatement of the reference SPH step (see oracle.h).
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this (as liboracle.so, through ctypes).  It only
 * checks results; it is never the thing measured as the product or shipped.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fno-fast-math -fopenmp).
 * FP32 functions keep the reference's operation order; no FMA contraction.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void orc_set_threads(int nthreads) { g_threads = nthreads < 1 ? 1 : nthreads; }
int orc_get_threads(void) { return g_threads; }

/* glm::pi<float>() (exts/glm/ext/scalar_constants.inl:13-17) */
#define PI_F ((float)3.14159265358979323846264338327950288)

/* offsets[27], physicsWorld.h:131-143: dx outer, dy, dz inner, each -1..1 */
static const float OFFS[27][3] = {
    {-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, 1},
    {-1, 1, -1},  {-1, 1, 0},  {-1, 1, 1},  {0, -1, -1}, {0, -1, 0}, {0, -1, 1},
    {0, 0, -1},   {0, 0, 0},   {0, 0, 1},   {0, 1, -1},  {0, 1, 0},  {0, 1, 1},
    {1, -1, -1},  {1, -1, 0},  {1, -1, 1},  {1, 0, -1},  {1, 0, 0},  {1, 0, 1},
    {1, 1, -1},   {1, 1, 0},   {1, 1, 1}};

void orc_default_params(orc_params *p)
{
    p->h = 0.35f;             /* physicsWorld.h:97 */
    p->rho0 = 99.7f;          /* physicsWorld.h:98 */
    p->k = 300.0f;            /* physicsWorld.h:99 */
    p->k_near = 20.0f;        /* physicsWorld.h:100 */
    p->mu = 0.5f;             /* physicsWorld.h:101 */
    p->gravity_on = 0;        /* physicsWorld.h:105 */
    p->gravity_scale = 10.0f; /* physicsWorld.h:106 */
    p->bounds[0] = p->bounds[1] = p->bounds[2] = 20.0f; /* physicsWorld.h:145 */
    p->lookahead = 1.0f / 120.0f; /* physicsWorld.cc:47 */
    p->damp = 0.95f;              /* physicsWorld.cc:87 */
    p->r2 = 0.35f * 0.35f;        /* physicsWorld.h:96 */
}

/* ------------------------------------------------------------------ */
/* kernels.h:25-82, evaluated per call, as the reference does          */
/* ------------------------------------------------------------------ */
float orc_W2(float dist, float radius) /* kernels.h:25-34 */
{
    if (dist < radius) {
        float volume = 15 / (2 * PI_F * powf(radius, 5));
        float v = radius - dist;
        return v * v * volume;
    }
    return 0;
}

float orc_W3(float dist, float radius) /* kernels.h:37-46 */
{
    if (dist < radius) {
        float volume = 15.0f / (PI_F * powf(radius, 6));
        float v = radius - dist;
        return v * v * v * volume;
    }
    return 0;
}

float orc_dW2(float dist, float radius) /* kernels.h:49-58 */
{
    if (dist <= radius) {
        float scale = 15.0f / (powf(radius, 5) * PI_F);
        float v = (radius - dist);
        return -v * scale;
    }
    return 0;
}

float orc_dW3(float dist, float radius) /* kernels.h:61-70 */
{
    if (dist <= radius) {
        float scale = 45 / (powf(radius, 6) * PI_F);
        float v = (radius - dist);
        return -v * v * scale;
    }
    return 0;
}

float orc_poly6(float dist, float radius) /* kernels.h:73-82 */
{
    if (dist < radius) {
        float scale = 315 / (64 * PI_F * powf(fabsf(radius), 9));
        float v = radius * radius - dist * dist;
        return v * v * v * scale;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* spatial hash, physicsWorld.cc:499-516                               */
/* ------------------------------------------------------------------ */
void orc_cell_coord(const float xp[3], float h, float cell[3]) /* cc:499-503 */
{
    /* glm: floor(pos / r) per component, then (int), stored back as float */
    for (int c = 0; c < 3; ++c) cell[c] = (float)(int)floorf(xp[c] / h);
}

/* cc:505-511.  (uint32_t)float of a negative value is formally undefined in
 * C++; the reference's x86-64 builds convert through a 64-bit signed integer
 * (cvttss2si) and keep the low 32 bits, i.e. two's-complement wrap. */
static inline uint32_t f2u_wrap(float f) { return (uint32_t)(int64_t)f; }

uint32_t orc_hash_cell(const float cell[3])
{
    uint32_t a = f2u_wrap(cell[0]) * 15823u;
    uint32_t b = f2u_wrap(cell[1]) * 9737333u;
    uint32_t c = f2u_wrap(cell[2]) * 440817757u;
    return a + b + c;
}

uint32_t orc_key(uint32_t hash, uint32_t table_size) { return hash % table_size; } /* cc:513-516 */

/* Test hook (no reference counterpart): the cell-grid bucket key. */
static int g_grid_key = 0;
static uint32_t g_gdim[3];
static int32_t g_glo[3];

/* Reference sort (timing only): stage B2 as std::sort of 12-byte float
 * records (oracle_sort.cpp) instead of the stable counting sort. */
static int g_reference_sort = 0;
void orc_std_sort_records(float *rec3, uint32_t n); /* oracle_sort.cpp */
void orc_set_reference_sort(int on) { g_reference_sort = on != 0; }

void orc_set_grid_key(const uint32_t *gdim, const int32_t *lo)
{
    g_grid_key = gdim != NULL && lo != NULL;
    for (int c = 0; c < 3; ++c) {
        g_gdim[c] = g_grid_key ? gdim[c] : 0;
        g_glo[c] = g_grid_key ? lo[c] : 0;
    }
}

/* Bucket of a cell: the reference's hash % T, or (hook set) the cell-grid
 * key (wx * Gy + wy) * Gz + wz with w = (u32(cell) - u32(lo)) mod G. */
static inline uint32_t cell_key(const float cell[3], uint32_t hash, uint32_t T)
{
    if (!g_grid_key) return orc_key(hash, T);
    uint32_t w[3];
    for (int c = 0; c < 3; ++c) w[c] = (f2u_wrap(cell[c]) - (uint32_t)g_glo[c]) % g_gdim[c];
    return (w[0] * g_gdim[1] + w[1]) * g_gdim[2] + w[2];
}

/* ------------------------------------------------------------------ */
/* stages                                                              */
/* ------------------------------------------------------------------ */
void orc_stage_predict(uint32_t n, const orc_params *p, float dt, const float *pos4,
                       float *vel4, float *xp4)
{
    /* cc:313-323: g = gravity ? (0,-gravityScale,0) : 0 */
    const float g[3] = {0.0f, p->gravity_on ? -p->gravity_scale : 0.0f, 0.0f};
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        for (int c = 0; c < 3; ++c) {
            float gv = g[c] * dt;                       /* cc:45: force * deltatime */
            vel4[4 * i + c] = vel4[4 * i + c] + gv;     /* velocity[i] += ... */
            xp4[4 * i + c] = pos4[4 * i + c] + vel4[4 * i + c] * p->lookahead; /* cc:47 */
        }
        xp4[4 * i + 3] = 0.0f;
    }
}

int orc_stage_spatial(uint32_t n, uint32_t T, float h, const float *xp4, uint32_t *hash_id,
                      uint32_t *key_id, uint32_t *sorted_key, uint32_t *sorted_idx,
                      uint32_t *start)
{
    if (n == 0 || T == 0 || n > (1u << 24) || T > (1u << 24) || !start) return -1;
    if (g_grid_key && (uint64_t)g_gdim[0] * g_gdim[1] * g_gdim[2] > T) return -1;
    if (g_grid_key && (g_gdim[0] == 0 || g_gdim[1] == 0 || g_gdim[2] == 0)) return -1;
    uint32_t *hid = hash_id ? hash_id : (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *kid = key_id ? key_id : (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *sk = sorted_key ? sorted_key : (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *si = sorted_idx ? sorted_idx : (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *cnt = (uint32_t *)calloc((size_t)T + 1, sizeof(uint32_t));
    int rc = 0;
    if (!hid || !kid || !sk || !si || !cnt) { rc = -1; goto out; }

    /* B1, cc:473-482 */
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        float cell[3];
        orc_cell_coord(&xp4[4 * i], h, cell);
        uint32_t hash = orc_hash_cell(cell);
        hid[i] = hash;
        kid[i] = cell_key(cell, hash, T);
    }
    for (uint32_t k = 0; k < T; ++k) start[k] = ORC_START_EMPTY;

    if (g_reference_sort) {
        /* B2, cc:480 + 484 as the reference runs it: {i, hash, key} as float
         * records, std::sort by key, serially */
        float *rec = (float *)malloc(sizeof(float) * 3 * (size_t)n);
        if (!rec) { rc = -1; goto out; }
#pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            rec[3 * i + 0] = (float)(uint32_t)i;
            rec[3 * i + 1] = (float)hid[i];
            rec[3 * i + 2] = (float)kid[i];
        }
        orc_std_sort_records(rec, n);
        for (uint32_t k = 0; k < n; ++k) {
            si[k] = (uint32_t)rec[3 * k + 0];
            sk[k] = (uint32_t)rec[3 * k + 2];
        }
        free(rec);
    } else {
        /* B2, cc:484: sort by key -- restated as a stable counting sort */
        for (uint32_t i = 0; i < n; ++i) cnt[kid[i] + 1]++;
        for (uint32_t k = 0; k < T; ++k) cnt[k + 1] += cnt[k];
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t pos = cnt[kid[i]]++;
            sk[pos] = kid[i];
            si[pos] = i;
        }
    }

    /* B3, cc:486-496 */
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t key = sk[i];
        uint32_t keyPrev = i == 0 ? UINT32_MAX : sk[i - 1];
        if (key != keyPrev) start[key] = i;
    }
out:
    if (!hash_id) free(hid);
    if (!key_id) free(kid);
    if (!sorted_key) free(sk);
    if (!sorted_idx) free(si);
    free(cnt);
    return rc;
}

/* The bucket walk of cc:331-363 / 378-419 / 432-462.  The reference's
 * spatialLookup record is {idx, hash, key} stored as float (cc:480), and the
 * walk compares index.z != key and index.y != hash as float vs uint32, i.e.
 * after converting the integer to float.  WALK(...) calls BODY(j, o[3], d2)
 * for every record that passes the key/hash/prefilter tests. */
#define WALK_BEGIN(PI_XP, SELF_SKIP, IDX_BREAK)                                           \
    {                                                                                     \
        float cell_[3];                                                                   \
        orc_cell_coord((PI_XP), p->h, cell_);                                            \
        for (int oi_ = 0; oi_ < 27; ++oi_) {                                              \
            float c_[3] = {cell_[0] + OFFS[oi_][0], cell_[1] + OFFS[oi_][1],              \
                           cell_[2] + OFFS[oi_][2]};                                      \
            uint32_t hash_ = orc_hash_cell(c_);                                           \
            uint32_t key_ = cell_key(c_, hash_, T);                                       \
            uint32_t cur_ = start[key_];                                                  \
            while (cur_ < n) {                                                            \
                uint32_t ridx_ = sorted_idx[cur_];                                        \
                float rkey_ = (float)sorted_key[cur_];                                    \
                float rhash_ = (float)hash_id[ridx_];                                     \
                float ridxf_ = (float)ridx_;                                              \
                cur_++;                                                                   \
                if (rkey_ != (float)key_) break;                                          \
                if (rhash_ != (float)hash_) continue;                                     \
                if ((IDX_BREAK) && ridxf_ >= (float)n) break;                             \
                uint32_t j = (uint32_t)ridxf_;                                            \
                if ((SELF_SKIP) && j == (uint32_t)i) continue;                            \
                float o[3] = {xp4[4 * j + 0] - (PI_XP)[0], xp4[4 * j + 1] - (PI_XP)[1],  \
                              xp4[4 * j + 2] - (PI_XP)[2]};                               \
                float d2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];                       \
                if (d2 > p->r2) continue;

#define WALK_END \
    }            \
    }            \
    }

void orc_stage_density(uint32_t n, uint32_t T, const orc_params *p, const float *xp4,
                       const uint32_t *sorted_key, const uint32_t *sorted_idx,
                       const uint32_t *hash_id, const uint32_t *start, float *dens2)
{
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float *pos = &xp4[4 * i];
        float density = 0, near_density = 0;
        WALK_BEGIN(pos, 0, 1)
        float dist = sqrtf(d2);
        density += orc_W2(dist, p->h);
        near_density += orc_W3(dist, p->h);
        WALK_END
        dens2[2 * i + 0] = density;
        dens2[2 * i + 1] = near_density;
    }
}

void orc_stage_pressure(uint32_t n, uint32_t T, const orc_params *p, float dt,
                        const float *xp4, const uint32_t *sorted_key,
                        const uint32_t *sorted_idx, const uint32_t *hash_id,
                        const uint32_t *start, const float *dens2, float *vel4)
{
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float density = dens2[2 * i + 0];
        const float near_density = dens2[2 * i + 1];
        const float pressure = (density - p->rho0) * p->k;       /* cc:371 */
        const float near_pressure = near_density * p->k_near;    /* cc:372 */
        float F[3] = {0, 0, 0};
        const float *pos = &xp4[4 * i];
        WALK_BEGIN(pos, 1, 0)
        float nd = dens2[2 * j + 0], nnd = dens2[2 * j + 1];
        float npress = (nd - p->rho0) * p->k;                    /* cc:407 */
        float nnpress = nnd * p->k_near;                         /* cc:408 */
        float sp = (pressure + npress) * 0.5f;                   /* cc:410 */
        float snp = (near_pressure + nnpress) * 0.5f;            /* cc:411 */
        float dist = sqrtf(d2);                                  /* cc:413 */
        float dir[3];
        if (dist > 0) { dir[0] = o[0] / dist; dir[1] = o[1] / dist; dir[2] = o[2] / dist; }
        else { dir[0] = 0; dir[1] = 1; dir[2] = 0; }             /* cc:414 */
        float a = orc_dW2(dist, p->h);
        for (int c = 0; c < 3; ++c) F[c] += dir[c] * a * sp / nd;   /* cc:416 */
        float b = orc_dW3(dist, p->h);
        for (int c = 0; c < 3; ++c) F[c] += dir[c] * b * snp / nnd; /* cc:417 */
        WALK_END
        for (int c = 0; c < 3; ++c) vel4[4 * i + c] += (F[c] / density) * dt; /* cc:421 */
    }
}

static void viscosity_one(int64_t i, uint32_t n, uint32_t T, const orc_params *p, float dt,
                          const float *xp4, const uint32_t *sorted_key,
                          const uint32_t *sorted_idx, const uint32_t *hash_id,
                          const uint32_t *start, const float *vread, float *vwrite)
{
    const float *pos = &xp4[4 * i];
    float F[3] = {0, 0, 0};
    const float velo[3] = {vread[4 * i + 0], vread[4 * i + 1], vread[4 * i + 2]};
    WALK_BEGIN(pos, 1, 0)
    float dist = sqrtf(d2);
    float influence = orc_poly6(dist, p->h);
    for (int c = 0; c < 3; ++c) F[c] += (vread[4 * j + c] - velo[c]) * influence; /* cc:460 */
    WALK_END
    for (int c = 0; c < 3; ++c) vwrite[4 * i + c] = velo[c] + F[c] * p->mu * dt; /* cc:463 */
}

void orc_stage_viscosity(uint32_t n, uint32_t T, const orc_params *p, float dt,
                         const float *xp4, const uint32_t *sorted_key,
                         const uint32_t *sorted_idx, const uint32_t *hash_id,
                         const uint32_t *start, float *vel4, int gauss_seidel)
{
    if (gauss_seidel) {
        /* in place, index order: velocity[j] for j < i is already updated */
        for (int64_t i = 0; i < (int64_t)n; ++i)
            viscosity_one(i, n, T, p, dt, xp4, sorted_key, sorted_idx, hash_id, start, vel4,
                          vel4);
        return;
    }
    float *snap = (float *)malloc(sizeof(float) * 4 * (size_t)n);
    if (!snap) return;
    memcpy(snap, vel4, sizeof(float) * 4 * (size_t)n);
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        viscosity_one(i, n, T, p, dt, xp4, sorted_key, sorted_idx, hash_id, start, snap, vel4);
    free(snap);
}

static inline float sign_f(float x) { return (float)(0 < x) - (float)(x < 0); } /* glm::sign */

void orc_stage_integrate(uint32_t n, const orc_params *p, float dt, float *pos4, float *vel4)
{
    const float damp_neg = -1 * p->damp; /* cc:94: `*= -1 * dampFactor` */
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        for (int c = 0; c < 3; ++c) {
            pos4[4 * i + c] += vel4[4 * i + c] * dt;               /* cc:84 */
            const float half = p->bounds[c] * 0.5f;                /* cc:88 */
            const float edge = half - fabsf(pos4[4 * i + c]);      /* cc:89 */
            if (edge <= 0) {                                       /* cc:91-105 */
                pos4[4 * i + c] = half * sign_f(pos4[4 * i + c]);
                vel4[4 * i + c] *= damp_neg;
            }
        }
        pos4[4 * i + 3] = 0.34f;                                   /* cc:107 */
    }
}

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

int orc_update(uint32_t n, uint32_t T, const orc_params *p, float dt, float *pos4,
               float *vel4, float *dens2, int gauss_seidel, double *stage_ms)
{
    if (n == 0 || T == 0 || n > (1u << 24) || T > (1u << 24)) return -1;
    float *xp4 = (float *)malloc(sizeof(float) * 4 * (size_t)n);
    uint32_t *hid = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *sk = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *si = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *st = (uint32_t *)malloc(sizeof(uint32_t) * T);
    float *d2 = dens2 ? dens2 : (float *)malloc(sizeof(float) * 2 * (size_t)n);
    int rc = 0;
    if (!xp4 || !hid || !sk || !si || !st || !d2) { rc = -1; goto out; }
    double t0 = now_ms();
    orc_stage_predict(n, p, dt, pos4, vel4, xp4);
    double t1 = now_ms();
    rc = orc_stage_spatial(n, T, p->h, xp4, hid, NULL, sk, si, st);
    if (rc) goto out;
    double t2 = now_ms();
    orc_stage_density(n, T, p, xp4, sk, si, hid, st, d2);
    double t3 = now_ms();
    orc_stage_pressure(n, T, p, dt, xp4, sk, si, hid, st, d2, vel4);
    double t4 = now_ms();
    orc_stage_viscosity(n, T, p, dt, xp4, sk, si, hid, st, vel4, gauss_seidel);
    double t5 = now_ms();
    orc_stage_integrate(n, p, dt, pos4, vel4);
    double t6 = now_ms();
    if (stage_ms) {
        stage_ms[0] = t1 - t0; stage_ms[1] = t2 - t1; stage_ms[2] = t3 - t2;
        stage_ms[3] = t4 - t3; stage_ms[4] = t5 - t4; stage_ms[5] = t6 - t5;
    }
out:
    free(xp4); free(hid); free(sk); free(si); free(st);
    if (!dens2) free(d2);
    return rc;
}

size_t orc_neighbours(uint32_t n, uint32_t T, const orc_params *p, const float *xp4,
                      const uint32_t *sorted_key, const uint32_t *sorted_idx,
                      const uint32_t *hash_id, const uint32_t *start, uint64_t *offs,
                      uint32_t *list)
{
    /* counting pass (serial: the output is a CSR in particle order) */
    size_t total = 0;
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float *pos = &xp4[4 * i];
        if (offs) offs[i] = total;
        WALK_BEGIN(pos, 0, 1)
        (void)d2;
        if (list) list[total] = j;
        total++;
        WALK_END
    }
    if (offs) offs[n] = total;
    return total;
}

/* ------------------------------------------------------------------ */
/* FP64 restatement on the FP32 neighbour structure                    */
/* ------------------------------------------------------------------ */
static const double PI_D = 3.14159265358979323846264338327950288;

void orc64_density(uint32_t n, uint32_t T, const orc_params *p, const float *xp4,
                   const uint32_t *sorted_key, const uint32_t *sorted_idx,
                   const uint32_t *hash_id, const uint32_t *start, double *dens2)
{
    const double h = p->h;
    const double vol2 = 15.0 / (2.0 * PI_D * pow(h, 5));
    const double vol3 = 15.0 / (PI_D * pow(h, 6));
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float *pos = &xp4[4 * i];
        double rho = 0, rhon = 0;
        WALK_BEGIN(pos, 0, 1)
        (void)o; (void)d2;
        double ox = (double)xp4[4 * j] - pos[0], oy = (double)xp4[4 * j + 1] - pos[1],
               oz = (double)xp4[4 * j + 2] - pos[2];
        double d = sqrt(ox * ox + oy * oy + oz * oz);
        if (d < h) {
            double v = h - d;
            rho += v * v * vol2;
            rhon += v * v * v * vol3;
        }
        WALK_END
        dens2[2 * i] = rho;
        dens2[2 * i + 1] = rhon;
    }
}

void orc64_pressure(uint32_t n, uint32_t T, const orc_params *p, double dt, const float *xp4,
                    const uint32_t *sorted_key, const uint32_t *sorted_idx,
                    const uint32_t *hash_id, const uint32_t *start, const double *dens2,
                    double *dv3, double *mag3)
{
    const double h = p->h, rho0 = p->rho0, k = p->k, kn = p->k_near;
    const double s2 = 15.0 / (pow(h, 5) * PI_D);
    const double s3 = 45.0 / (pow(h, 6) * PI_D);
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float *pos = &xp4[4 * i];
        const double rho = dens2[2 * i], rhon = dens2[2 * i + 1];
        const double P = (rho - rho0) * k, Pn = rhon * kn;
        double F[3] = {0, 0, 0}, M[3] = {0, 0, 0};
        WALK_BEGIN(pos, 1, 0)
        (void)o; (void)d2;
        double ov[3] = {(double)xp4[4 * j] - pos[0], (double)xp4[4 * j + 1] - pos[1],
                        (double)xp4[4 * j + 2] - pos[2]};
        double d = sqrt(ov[0] * ov[0] + ov[1] * ov[1] + ov[2] * ov[2]);
        double dir[3] = {0, 1, 0};
        if (d > 0) { dir[0] = ov[0] / d; dir[1] = ov[1] / d; dir[2] = ov[2] / d; }
        double rj = dens2[2 * j], rnj = dens2[2 * j + 1];
        double sp = 0.5 * (P + (rj - rho0) * k), snp = 0.5 * (Pn + rnj * kn);
        double a = d <= h ? -(h - d) * s2 : 0.0;
        double b = d <= h ? -(h - d) * (h - d) * s3 : 0.0;
        /* scale: the terms with |P_i| + |P_j| in place of |P_i + P_j|, so a
         * shared pressure that cancels keeps the size of its two inputs */
        const double up = 0.5 * (fabs(P) + fabs((rj - rho0) * k)), unp = 0.5 * (fabs(Pn) + fabs(rnj * kn));
        for (int c = 0; c < 3; ++c) {
            double t1 = dir[c] * a * sp / rj, t2 = dir[c] * b * snp / rnj;
            F[c] += t1 + t2;
            M[c] += fabs(dir[c] * a) * up / fabs(rj) + fabs(dir[c] * b) * unp / fabs(rnj);
        }
        WALK_END
        for (int c = 0; c < 3; ++c) {
            dv3[3 * i + c] = F[c] / rho * dt;
            if (mag3) mag3[3 * i + c] = M[c] / fabs(rho) * dt;
        }
    }
}

void orc64_viscosity(uint32_t n, uint32_t T, const orc_params *p, double dt, const float *xp4,
                     const uint32_t *sorted_key, const uint32_t *sorted_idx,
                     const uint32_t *hash_id, const uint32_t *start, const double *vin3,
                     double *dv3, double *mag3)
{
    const double h = p->h, mu = p->mu;
    const double s = 315.0 / (64.0 * PI_D * pow(fabs(h), 9));
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float *pos = &xp4[4 * i];
        double F[3] = {0, 0, 0}, M[3] = {0, 0, 0};
        WALK_BEGIN(pos, 1, 0)
        (void)o; (void)d2;
        double ox = (double)xp4[4 * j] - pos[0], oy = (double)xp4[4 * j + 1] - pos[1],
               oz = (double)xp4[4 * j + 2] - pos[2];
        double dd = ox * ox + oy * oy + oz * oz;
        double w = 0.0;
        if (sqrt(dd) < h) {
            double v = h * h - dd;
            w = v * v * v * s;
        }
        for (int c = 0; c < 3; ++c) {
            double t = (vin3[3 * j + c] - vin3[3 * i + c]) * w;
            F[c] += t;
            M[c] += fabs(t);
        }
        WALK_END
        for (int c = 0; c < 3; ++c) {
            dv3[3 * i + c] = F[c] * mu * dt;
            if (mag3) mag3[3 * i + c] = M[c] * fabs(mu) * dt;
        }
    }
}

/* ------------------------------------------------------------------ */
/* GridArrangement, physicsWorld.cc:139-142, 518-557                   */
/* ------------------------------------------------------------------ */
void orc_grid_arrangement(uint32_t n, float *pos4)
{
    int per_axis = (int)ceilf(powf((float)n, (1.0f / 3.0f))); /* cc:139 */
    float gap = 0.215f;                                         /* cc:140 */
    uint32_t i = 0;
    float total_w = per_axis * gap;                             /* cc:522 */
    for (uint32_t q = 0; q < n; ++q) {
        pos4[4 * q + 0] = pos4[4 * q + 1] = pos4[4 * q + 2] = 0.0f;
        pos4[4 * q + 3] = 0.25f;                                /* cc:132 */
    }
    for (int ly = 0; ly < per_axis; ly++)
        for (int lx = 0; lx < per_axis; lx++)
            for (int lz = 0; lz < per_axis; lz++) {
                if (i >= n) return;                             /* cc:532-535 */
                float xo = lx * gap, yo = ly * gap, zo = lz * gap;
                float wx = (0 - ((total_w - gap) / 2.0f));
                float wy = (0 + (total_w - gap) / 2.0f);
                float wz = (0 - (total_w - gap) / 2.0f);
                pos4[4 * i + 0] = wx + xo;
                pos4[4 * i + 1] = wy - yo;
                pos4[4 * i + 2] = wz + zo;
                pos4[4 * i + 3] = 0.34f;                        /* cc:551 */
                i++;
            }
}

/* ------------------------------------------------------------------ */
/* render hand-off, fluidSimCPU.cc:100-125 + physicsWorld.cc:178-182  */
/* ------------------------------------------------------------------ */
static inline float glm_min(float x, float y) { return (y < x) ? y : x; } /* func_common.inl:17-21 */
static inline float glm_max(float x, float y) { return (x < y) ? y : x; } /* func_common.inl:25-30 */

void orc_speed_colors(uint32_t n, const float *vel4, const float stops[16], float *rgba4)
{
    const float *Color1 = stops, *Color2 = stops + 4, *Color3 = stops + 8, *Color4 = stops + 12;
    for (uint32_t i = 0; i < n; ++i) {
        const float *v = &vel4[4 * (size_t)i];
        /* getSpeedNormalzied: glm::clamp(glm::length(velocity), 0.0f, 1.5f) / 1.5f */
        float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        float normalized = glm_min(glm_max(len, 0.0f), 1.5f) / 1.5f;
        float breakpoint1 = 0.33f; /* fluidSimCPU.cc:107-108 */
        float breakpoint2 = 0.66f;
        float *out = &rgba4[4 * (size_t)i];
        for (int c = 0; c < 4; ++c) {
            if (normalized <= breakpoint1) {
                out[c] = (1.0f - normalized / breakpoint1) * Color1[c] +
                         (normalized / breakpoint1) * Color2[c];
            } else if (normalized <= breakpoint2) {
                out[c] = (1.0f - (normalized - breakpoint1) / (breakpoint2 - breakpoint1)) * Color2[c] +
                         ((normalized - breakpoint1) / (breakpoint2 - breakpoint1)) * Color3[c];
            } else {
                out[c] = (1.0f - (normalized - breakpoint2) / (1.0f - breakpoint2)) * Color3[c] +
                         ((normalized - breakpoint2) / (1.0f - breakpoint2)) * Color4[c];
            }
        }
    }
}
