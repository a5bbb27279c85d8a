// sph_device.h -- integer and floating-point building blocks shared by the
// gfx950 kernels and by host code (the same source compiles for both, so the
// host-side unit tests exercise the exact device arithmetic).
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SPH_HD __host__ __device__ __forceinline__
#else
#define SPH_HD inline
#endif

namespace sph {

// HashCell (physicsWorld.cc:505-511) on integer cell coordinates.  The
// reference converts the float cell coordinate to uint32_t; on its x86-64
// builds a negative value wraps (two's complement), which the int32 -> uint32
// conversion here reproduces for |cell| < 2^31.  (AMDGPU's direct
// float -> u32 conversion would saturate negatives to 0 instead.)
SPH_HD uint32_t hash_cell(int32_t cx, int32_t cy, int32_t cz)
{
    return (uint32_t)cx * 15823u + (uint32_t)cy * 9737333u + (uint32_t)cz * 440817757u;
}

// hash % table (GetKeyFromHash, physicsWorld.cc:513-516) without a divide:
// Lemire, Kaser & Kurz, "Faster remainder by direct computation" (2019).
// m = floor((2^64 - 1) / table) + 1 (m = 0 for table == 1, still exact).
// Exact for every 32-bit numerator and divisor.
SPH_HD uint64_t fastmod_constant(uint32_t table) { return ~0ull / table + 1ull; }

SPH_HD uint32_t fastmod(uint32_t a, uint64_t m, uint32_t table)
{
    const uint64_t low = m * (uint64_t)a;  // mod 2^64
    const uint32_t lo = (uint32_t)low, hi = (uint32_t)(low >> 32);
    // floor(low * table / 2^64) = floor((hi*table + floor(lo*table / 2^32)) / 2^32)
    const uint64_t lo_hi = ((uint64_t)lo * table) >> 32;
    return (uint32_t)(((uint64_t)hi * table + lo_hi) >> 32);
}

// Cell-grid bucket key, per axis: w = (u32(c) - u32(lo)) mod G, with
// (u32(c) - u32(lo)) wrapping mod 2^32.  For c in [lo, lo + G) this is c - lo.
SPH_HD uint32_t grid_wrap(uint32_t u, uint32_t g, uint64_t m) { return u < g ? u : fastmod(u, m, g); }

SPH_HD uint32_t grid_key(int32_t cx, int32_t cy, int32_t cz, const int32_t lo[3],
                         const uint32_t g[3], const uint64_t m[3])
{
    const uint32_t wx = grid_wrap((uint32_t)cx - (uint32_t)lo[0], g[0], m[0]);
    const uint32_t wy = grid_wrap((uint32_t)cy - (uint32_t)lo[1], g[1], m[1]);
    const uint32_t wz = grid_wrap((uint32_t)cz - (uint32_t)lo[2], g[2], m[2]);
    return (wx * g[1] + wy) * g[2] + wz;
}

// Render colour of a particle from its velocity: FluidSimCPU::updateColors
// (fluidSimCPU.cc:100-125) over getSpeedNormalzied (physicsWorld.cc:178-182):
// clamp(length(v), 0, 1.5) / 1.5 with glm's length = sqrt(dot(v, v)),
// clamp = min(max(x, lo), hi), max(x, y) = x < y ? y : x, min(x, y) = y < x ?
// y : x; then the 4-stop gradient with breakpoints 0.33f and 0.66f held in
// float variables (so b2 - b1 and 1 - b2 are float operations), scalar * vec4
// per component.  stops = Color1..Color4, RGBA each.
SPH_HD void speed_color(float vx, float vy, float vz, const float *stops, float out[4])
{
    const float len = sqrtf((vx * vx + vy * vy) + vz * vz);
    const float mx = len < 0.0f ? 0.0f : len;  // glm::max(len, 0)
    const float nrm = (1.5f < mx ? 1.5f : mx) / 1.5f;
    const float b1 = 0.33f, b2 = 0.66f;
    const float *ca, *cb;
    float t;
    if (nrm <= b1) {
        t = nrm / b1;
        ca = stops;
        cb = stops + 4;
    } else if (nrm <= b2) {
        t = (nrm - b1) / (b2 - b1);
        ca = stops + 4;
        cb = stops + 8;
    } else {
        t = (nrm - b2) / (1.0f - b2);
        ca = stops + 8;
        cb = stops + 12;
    }
    const float u = 1.0f - t;
    for (int c = 0; c < 4; ++c) out[c] = u * ca[c] + t * cb[c];
}

// Correctly rounded quotient a / b from y = RN(1/b), without a divide
// (Markstein's theorem: with y within half an ulp of 1/b and q0 = RN(a*y)
// within one ulp of a/b, r = a - b*q0 is exact and RN(q0 + r*y) = RN(a/b),
// provided nothing under- or overflows).  quotient_ok() states the range in
// which that holds for the kernels' use: b in [2^-26, 2^26] (y stored only
// then), and a = 0 or |a| in [2^-91, 2^100] -- so r is exact (|a| >= 2^-103),
// y and q0 are normal and q0 is finite.  (For a = -0 the result is +0 where
// IEEE gives -0; a zero term does not change a sum that starts at +0.)
SPH_HD bool quotient_ok(float a)
{
    const float m = fabsf(a);
    return a == 0.0f || (m >= 0x1p-91f && m <= 0x1p100f);
}

// Sign-cleared bit pattern: non-negative floats order as these integers.
SPH_HD uint32_t mag_bits(float x)
{
    uint32_t u;
    __builtin_memcpy(&u, &x, sizeof u);
    return u & 0x7fffffffu;
}

constexpr uint32_t kQuotLo = 0x12000000u;  // 2^-91
constexpr uint32_t kQuotHi = 0x71800000u;  // 2^100

// quotient_ok() of every argument, without branches: with u the magnitude
// bits, a value fails when it is nonzero below the range (u - 1 < lo - 1,
// unsigned, so zero wraps above) or above it (u > hi; infinities and NaNs
// too).  One min and one max per value.
template <typename... F>
SPH_HD bool quotients_ok(F... a)
{
    uint32_t lo = 0xffffffffu, hi = 0u;
    ((lo = (mag_bits(a) - 1u < lo ? mag_bits(a) - 1u : lo),
      hi = (mag_bits(a) > hi ? mag_bits(a) : hi)),
     ...);
    return lo >= kQuotLo - 1u && hi <= kQuotHi;
}

// d in [2^-26, 2^26], the divisor range of quotient() (the bit pattern
// itself: negative values, zeros and NaNs fall outside).
SPH_HD bool divisor_ok(float d)
{
    uint32_t u;
    __builtin_memcpy(&u, &d, sizeof u);
    return u - 0x32800000u <= 0x4c800000u - 0x32800000u;
}

SPH_HD float quotient(float a, float b, float y)
{
    const float q0 = a * y;
    const float r = fmaf(-b, q0, a);
    return fmaf(r, y, q0);
}

// RN(1/x) for x in [2^-26, 2^26] (divisor_ok).  On the GPU: the hardware
// reciprocal estimate y0 (within one ulp) refined by one Newton-Raphson step
// with fused multiply-adds, e = 1 - x*y0 (exact) and y0 + e*y0 (Markstein's
// correctly rounded reciprocal, IBM J. Res. Dev. 34(1), 1990); it equals IEEE
// 1/x for every float of the range (checked exhaustively on the GPU by
// sph_debug_reciprocal_mismatches).  On the host: IEEE 1/x.
SPH_HD float recip_rn(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float y0 = __builtin_amdgcn_rcpf(x);
    const float e = fmaf(-x, y0, 1.0f);
    return fmaf(e, y0, y0);
#else
    return 1.0f / x;
#endif
}

// y = RN(1/x) for x in [2^-26, 2^26], else 0 (no fast quotient by x).
SPH_HD float reciprocal_or_zero(float x) { return divisor_ok(x) ? recip_rn(x) : 0.0f; }

// True when |x|, |y|, |z| >= 2^-60.  Then every nonzero difference between
// this point's coordinates and another float's has magnitude >= 2^-84:
// floats of magnitude >= 2^-61 are multiples of 2^-84, and a float below
// 2^-61 lies at least 2^-61 away.  Such differences are valid quotient()
// numerators whenever the divisor d bounds them (|difference| <= d(1+2^-22)
// for the components of a distance d <= 2^26).
SPH_HD bool coords_clear_of_zero(float x, float y, float z)
{
    const uint32_t a = mag_bits(x), b = mag_bits(y), c = mag_bits(z);
    const uint32_t m = a < b ? (a < c ? a : c) : (b < c ? b : c);
    return m >= 0x21800000u;  // 2^-60
}

// glm::sign for float (func_common.inl:144-150): (0 < x) - (x < 0).
SPH_HD float sign_f(float x) { return (float)(0.0f < x) - (float)(x < 0.0f); }

}  // namespace sph
