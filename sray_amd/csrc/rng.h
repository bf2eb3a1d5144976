/*
 * rng.h — the counter-based random numbers of the device path, usable from
 * host code (projection maps) and device code (camera jitter, light samples,
 * diffuse directions, photon emission and bounces).
 *
 * Every draw is a pure function of (seed, a, b, c, slot): a SplitMix64 chain
 * (Steele, Lea, Flood 2014) over the key words, so a draw does not depend on
 * thread scheduling or on how work is split over launches.  The reference
 * draws from one global rand() stream (src/sray.cc, src/light.cc); its
 * sequence is not reproduced, its distributions are.
 */
#ifndef SRAY_RNG_H_
#define SRAY_RNG_H_

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define SRAY_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define SRAY_HD inline
#endif

namespace srd {

SRAY_HD uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
SRAY_HD uint32_t rng_u32(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    uint64_t h = splitmix64(seed);
    h = splitmix64(h ^ (uint64_t)a);
    h = splitmix64(h ^ (((uint64_t)b << 32) | (uint64_t)c));
    h = splitmix64(h ^ (uint64_t)d);
    return (uint32_t)(h >> 32);
}
/* uniform in [0,1) with 24 random bits */
SRAY_HD float rng_unit(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    return (float)(rng_u32(seed, a, b, c, d) >> 8) * (1.0f / 16777216.0f);
}
/* slot word: kind<<30 | i1<<20 | i2<<10 | attempt<<3 | dim (kind 0 camera,
 * 1 light sample, 2 diffuse GI direction, 3 photon emission/bounce) */
SRAY_HD uint32_t rslot(uint32_t kind, uint32_t i1, uint32_t i2, uint32_t att, uint32_t dim)
{
    return (kind << 30) | (i1 << 20) | (i2 << 10) | (att << 3) | dim;
}
constexpr uint32_t K_CAMERA = 0u, K_LIGHT = 1u, K_DIFFUSE = 2u, K_PHOTON = 3u;

/* photon draws: a = emission index, b = map kind << 24 | light, c = bounce;
 * the purpose goes into i1, the attempt number into i2 (high) and att (low 7 bits) */
constexpr uint32_t P_TIME = 0u, P_LPOINT = 1u, P_EMIT = 2u, P_RR = 3u, P_DIFF = 4u, P_PROJ = 5u;
constexpr uint32_t PHOTON_ATTEMPTS = 1u << 17; /* emission rejection attempts towards full cells */
constexpr int PHOTON_MAX_BOUNCES = 1024;       /* guard on the bounce loops */
constexpr uint32_t PHOTON_SPHERE_ATTEMPTS = 128u;

SRAY_HD float prand(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t purpose, uint32_t att, uint32_t dim)
{
    return rng_unit(seed, a, b, c, rslot(K_PHOTON, purpose, att >> 7, att & 127u, dim));
}

/* one attempt of Marsaglia's (1972) uniform point on the unit sphere from two
 * uniforms; false if the pair is rejected */
SRAY_HD bool marsaglia(float u1, float u2, float out[3])
{
    float x1 = 2.0f * u1 - 1.0f, x2 = 2.0f * u2 - 1.0f;
    float s = x1 * x1 + x2 * x2;
    if (!(s < 1.0f && s > 0.0f)) return false;
    float q = 2.0f * sqrtf(1.0f - s);
    out[0] = x1 * q;
    out[1] = x2 * q;
    out[2] = 1.0f - 2.0f * s;
    return true;
}

} // namespace srd

#endif
