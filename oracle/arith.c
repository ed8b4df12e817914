/*
 * arith.c -- the per-element arithmetic shared by every oracle instance:
 * noise hash, Poisson-kick decision, Gaussian draw, portable exp, int-ring
 * quantisation.  Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdint.h>
#include "oracle.h"

/* sim/src/state.rs:283-295 -- splitmix64 over i*phi + t*c1 + seed*c2,
 * top 32 bits of the finalised value. */
uint32_t or_noise_hash(uint32_t i, uint64_t t, uint64_t seed)
{
    uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull + t * 0xBF58476D1CE4E5B9ull +
                 seed * 0x94D049BB133111EBull;
    h ^= h >> 30;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27;
    h *= 0x94D049BB133111EBull;
    h ^= h >> 31;
    return (uint32_t)(h >> 32);
}

/* sim/src/state.rs:298-301 -- (f32)hash * 2.3283064e-10 < p */
int or_noise_fires(uint32_t i, uint64_t t, uint64_t seed, float p)
{
    return (float)or_noise_hash(i, t, seed) * 2.3283064e-10f < p;
}

static uint64_t mix64(uint64_t h)
{
    h ^= h >> 30;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27;
    h *= 0x94D049BB133111EBull;
    h ^= h >> 31;
    return h;
}

/* Gaussian background draw (not in the reference; defined here and in the
 * product identically).  Irwin-Hall(12) over twelve 8-bit fields: the four
 * bytes of the LOW half of the kick hash h0 = splitmix64(key) (its high half
 * decides the kick) and the eight bytes of h1 = splitmix64(key +
 * 0xD1B54A32D192ED03), where key is the kick key of (i, t, seed).
 * g = (sum(bytes) - 1530) / 256: mean exactly 0, variance (12 * (256^2 - 1)
 * / 12) / 65536 = 1 - 2^-16, range +-5.98, values on a 2^-8 lattice.  Only
 * integer work, one exact int->float conversion and an exact power-of-two
 * scale, so every IEEE implementation returns the same bits; two hashes per
 * neuron-step in all (one shared with the kick). */
static uint32_t byte_sum32(uint32_t x)
{
    return (x & 0xFF) + ((x >> 8) & 0xFF) + ((x >> 16) & 0xFF) + (x >> 24);
}

float or_gauss_draw(uint32_t i, uint64_t t, uint64_t seed)
{
    uint64_t key = (uint64_t)i * 0x9E3779B97F4A7C15ull + t * 0xBF58476D1CE4E5B9ull +
                   seed * 0x94D049BB133111EBull;
    uint64_t h0 = mix64(key);
    uint64_t h1 = mix64(key + 0xD1B54A32D192ED03ull);
    uint32_t sum = byte_sum32((uint32_t)h0) + byte_sum32((uint32_t)h1) +
                   byte_sum32((uint32_t)(h1 >> 32));
    return (float)((int32_t)sum - 1530) * (1.0f / 256.0f);
}

/* Portable single-precision exp for AdEx (not libm's): Cody & Waite (1980)
 * argument reduction x = k ln2 + r with the two-part ln2 of fdlibm's
 * e_expf.c (ln2_hi has 12 trailing zero bits so k*ln2_hi is exact), then
 * the degree-7 Taylor polynomial of e^r in Horner form, then 2^k applied
 * through the exponent field in two halves (so subnormal results are
 * formed by one final rounding multiply).  Plain f32 add/mul only -- with
 * contraction off it gives the same bits on any IEEE machine. */
float or_portable_expf(float x)
{
    if (x != x)
        return x;
    if (x > 88.72283f)
        return INFINITY;
    if (x < -103.97208f)
        return 0.0f;
    const float log2e = 0x1.715476p+0f;   /* 1.44269504 */
    const float ln2_hi = 0x1.62e400p-1f;  /* 6.9314575195e-01 */
    const float ln2_lo = 0x1.7f7d1cp-20f; /* 1.4286067653e-06 */
    float kf = rintf(x * log2e);
    int k = (int)kf;
    float r = (x - kf * ln2_hi) - kf * ln2_lo;
    float p = 0x1.a01a02p-13f;            /* 1/5040 */
    p = p * r + 0x1.6c16c2p-10f;          /* 1/720 */
    p = p * r + 0x1.111112p-7f;           /* 1/120 */
    p = p * r + 0x1.555556p-5f;           /* 1/24 */
    p = p * r + 0x1.555556p-3f;           /* 1/6 */
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int k1 = k / 2;
    int k2 = k - k1;
    union { uint32_t u; float f; } s1, s2;
    s1.u = (uint32_t)(k1 + 127) << 23;
    s2.u = (uint32_t)(k2 + 127) << 23;
    return (p * s1.f) * s2.f;
}

/* sim/src/kernels.rs:61 -- (int)lrintf(w * scale), round half to even */
int32_t or_ring_put(float w, float scale)
{
    return (int32_t)lrintf(w * scale);
}
