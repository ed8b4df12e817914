// device.h -- per-neuron arithmetic of the step, written for gfx950.
//
// Every function here is plain IEEE f32 add/mul/div with contraction off
// (the library is compiled with -ffp-contract=off), so each result equals
// the CPU restatement of the reference step bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nh {

// splitmix64 finaliser.
__device__ __forceinline__ uint64_t mix64(uint64_t h) {
    h ^= h >> 30;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27;
    h *= 0x94D049BB133111EBull;
    h ^= h >> 31;
    return h;
}

__device__ __forceinline__ uint64_t noise_key(uint32_t i, uint64_t t, uint64_t seed) {
    return (uint64_t)i * 0x9E3779B97F4A7C15ull + t * 0xBF58476D1CE4E5B9ull +
           seed * 0x94D049BB133111EBull;
}

// Poisson-kick decision: top 32 bits of h = splitmix64 over the (i, t, seed)
// key, scaled by 2^-32, compared with rate*dt/1000
// (sim/src/state.rs:283-301, sim/src/kernels.rs:69-77,128-130).
__device__ __forceinline__ bool kick_fires(uint64_t h, float p) {
    // the high word converted as a 32-bit integer (one v_cvt_f32_u32,
    // round to nearest even); left to itself the compiler widens this to a
    // 64-bit integer conversion sequence with the same result
    float f;
    asm("v_cvt_f32_u32 %0, %1" : "=v"(f) : "v"((uint32_t)(h >> 32)));
    return f * 0x1p-32f < p;
}

// Gaussian current draw: Irwin-Hall(12) over twelve 8-bit fields -- the four
// bytes of the low half of the kick hash h0 (the high half decides the kick)
// and the eight bytes of h1 = splitmix64(key + 0xD1B54A32D192ED03); integer
// byte sums (v_sad_u8 against zero), exact conversion, exact 2^-8 scale --
// mean 0, variance 1 - 2^-16.
__device__ __forceinline__ float gauss_draw(uint64_t h0, uint64_t key) {
    const uint64_t h1 = mix64(key + 0xD1B54A32D192ED03ull);
    uint32_t sum = __builtin_amdgcn_sad_u8((uint32_t)h0, 0u, 0u);
    sum = __builtin_amdgcn_sad_u8((uint32_t)h1, 0u, sum);
    sum = __builtin_amdgcn_sad_u8((uint32_t)(h1 >> 32), 0u, sum);
    return (float)((int32_t)sum - 1530) * 0x1p-8f;
}

// Portable exp: Cody & Waite reduction with fdlibm's two-part ln2, degree-7
// Taylor polynomial in Horner form, 2^k applied in two exponent-field
// halves.  Same operation sequence as the CPU restatement.
__device__ __forceinline__ float portable_expf(float x) {
    if (x != x) return x;
    if (x > 0x1.62e42ep+6f) return __builtin_inff();
    if (x < -0x1.9fe368p+6f) return 0.0f;
    const float kf = __builtin_rintf(x * 0x1.715476p+0f);
    const int k = (int)kf;
    const float r = (x - kf * 0x1.62e400p-1f) - kf * 0x1.7f7d1cp-20f;
    float p = 0x1.a01a02p-13f;
    p = p * r + 0x1.6c16c2p-10f;
    p = p * r + 0x1.111112p-7f;
    p = p * r + 0x1.555556p-5f;
    p = p * r + 0x1.555556p-3f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const int k1 = k / 2;
    const int k2 = k - k1;
    const float s1 = __uint_as_float((uint32_t)(k1 + 127) << 23);
    const float s2 = __uint_as_float((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

// Izhikevich step (sim/src/model/izh.rs:309-317): reset for last step's
// crossing, integrate v, integrate u with the new v, spike test.
__device__ __forceinline__ bool izh_step(float a, float b, float c, float d, float2 &x,
                                         float I, float dt) {
    float v = x.x, u = x.y;
    if (v >= 30.0f) {
        v = c;
        u += d;
    }
    v += dt * (0.04f * v * v + 5.0f * v + 140.0f - u + I);
    u += dt * a * (b * v - u);
    x = make_float2(v, u);
    return v >= 30.0f;
}

// AdEx step (sim/src/model/adex.rs:163-176), row
// [C, gL, EL, VT, dT, a, tau_w, b, Vr, Vpeak]; exp is portable_expf.
__device__ __forceinline__ bool adex_step(const float *p, float2 &x, float I, float dt) {
    float v = x.x, w = x.y;
    if (v >= p[9]) {
        v = p[8];
        w += p[7];
    }
    float q = (v - p[3]) / p[4];
    q = q > 10.0f ? 10.0f : q;
    const float ex = portable_expf(q);
    v += dt * (-p[1] * (v - p[2]) + p[1] * p[4] * ex - w + I) / p[0];
    w += dt * (p[5] * (v - p[2]) - w) / p[6];
    const bool sp = v >= p[9];
    if (v > p[9]) v = p[9];
    x = make_float2(v, w);
    return sp;
}

// Leaky integrate-and-fire step (the reference's third model, a test-only
// plug-in: sim/tests/lif.rs:14-104), row [tau, e, r, vth, vreset], state v in
// x.x (x.y unused): reset for last step's crossing, integrate, spike test.
__device__ __forceinline__ bool lif_step(const float *p, float2 &x, float I, float dt) {
    float v = x.x;
    if (v >= p[3]) v = p[4];
    v += dt * (p[1] - v + p[2] * I) / p[0];
    x.x = v;
    return v >= p[3];
}

}  // namespace nh
