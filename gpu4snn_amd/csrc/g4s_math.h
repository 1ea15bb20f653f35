// g4s_math.h -- random streams and neuron arithmetic shared by the HIP
// kernels and the C++ host runtime (network builder).
//
// Every float expression below is written as the sequence of binary32
// operations the specification fixes; all translation units that include it
// are compiled with -ffp-contract=off, and only +, -, *, fmaf and integer
// operations are used, so host and gfx950 produce identical bits.
//
// Reference: Crypto-TII/GPU4SNN gpu4snn.cu.  The reference draws from cuRAND
// XORWOW (cu:273-283, cu:358) and glibc rand() (cu:873-893), both seeded by
// time(NULL); this file replaces them by a counter-based Philox4x32-10
// stream: key = 64-bit network seed, counter = (a, b, tag, 0).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define G4S_HD __host__ __device__ __forceinline__

namespace g4s {

enum : uint32_t { kTagParam = 1, kTagConn = 2, kTagWeight = 3, kTagThal = 4 };

struct Word4 { uint32_t x0, x1, x2, x3; };

// Philox4x32-10 (Salmon et al., SC'11): ten rounds of two 32x32->64 bit
// products with the Philox multipliers, key bumped by the Weyl constants.
G4S_HD Word4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Word4{c0, c1, c2, c3};
}

G4S_HD Word4 draw(uint64_t seed, uint32_t a, uint32_t b, uint32_t tag)
{
    return philox(a, b, tag, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
}

G4S_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
G4S_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// (2*(x>>9)+1) * 2^-24: exact, strictly inside (0,1).
G4S_HD float u01(uint32_t x) { return (float)(((x >> 9) << 1) | 1u) * 0x1p-24f; }

// 1/y for y in [1,4): bit-pattern seed and three Newton steps with fma.
G4S_HD float rcp_nr(float y)
{
    float r = u2f(0x7EF311C3u - f2u(y));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float e = __builtin_fmaf(-y, r, 1.0f);
        r = __builtin_fmaf(r, e, r);
    }
    return r;
}

// ln(x) for positive normal x: x = m*2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2*atanh(s), s = (m-1)/(m+1), series through s^11.
G4S_HD float ln_f(float x)
{
    const uint32_t bits = f2u(x);
    int32_t e = (int32_t)(bits >> 23) - 127;
    float m = u2f((bits & 0x007FFFFFu) | 0x3F800000u);
    if (m > 0x1.6a09e6p+0f) {
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f * rcp_nr(m + 1.0f);
    const float z = s * s;
    float p = __builtin_fmaf(z, 0x1.745d18p-4f, 0x1.c71c72p-4f);
    p = __builtin_fmaf(z, p, 0x1.24924ap-3f);
    p = __builtin_fmaf(z, p, 0x1.99999ap-3f);
    p = __builtin_fmaf(z, p, 0x1.555556p-2f);
    p = p * z;
    const float s2 = s + s;
    const float lm = __builtin_fmaf(s2, p, s2);
    const float ef = (float)e;
    return __builtin_fmaf(ef, 0x1.62e4p-1f, __builtin_fmaf(ef, 0x1.7f7d1cp-20f, lm));
}

// sqrt(x) for positive normal x: inverse-sqrt seed, three Newton steps,
// one residual correction.
G4S_HD float sqrt_f(float x)
{
    float y = u2f(0x5F3759DFu - (f2u(x) >> 1));
    const float h = 0.5f * x;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float hy = h * y;
        y = y * __builtin_fmaf(-hy, y, 1.5f);
    }
    const float s = x * y;
    const float r = __builtin_fmaf(-s, s, x);
    return __builtin_fmaf(r, 0.5f * y, s);
}

// cos(2*pi*((x>>8)+0.5)/2^24) by exact integer quadrant reduction and Taylor
// polynomials of sin/cos on (0, pi/4].
G4S_HD float cos_2pi(uint32_t x)
{
    const uint32_t k = x >> 8;
    const uint32_t q = k >> 22;
    uint32_t r = k & 0x3FFFFFu;
    uint32_t use_sin = q & 1u;
    if (r >= 0x200000u) {
        r = 0x3FFFFFu - r;
        use_sin ^= 1u;
    }
    const float a = (float)(2u * r + 1u) * 0x1.921fb6p-23f;
    const float a2 = a * a;
    float v;
    if (use_sin) {
        float p = __builtin_fmaf(a2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
        p = __builtin_fmaf(a2, p, 0x1.111112p-7f);
        p = __builtin_fmaf(a2, p, -0x1.555556p-3f);
        v = __builtin_fmaf(a * a2, p, a);
    } else {
        float p = __builtin_fmaf(a2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
        p = __builtin_fmaf(a2, p, -0x1.6c16c2p-10f);
        p = __builtin_fmaf(a2, p, 0x1.555556p-5f);
        p = __builtin_fmaf(a2, p, -0.5f);
        v = __builtin_fmaf(a2, p, 1.0f);
    }
    return (q == 1u || q == 2u) ? -v : v;
}

// Box-Muller N(0,1) from two words.
G4S_HD float normal(uint32_t x0, uint32_t x1)
{
    const float u1 = u01(x0);
    const float l = ln_f(u1);
    const float w = -2.0f * l;
    return sqrt_f(w) * cos_2pi(x1);
}

// Thalamic draw of neuron i at step t (replaces curand_normal, cu:358).
G4S_HD float thalamic(uint64_t seed, uint32_t t, uint32_t i)
{
    const Word4 o = draw(seed, t, i, kTagThal);
    return normal(o.x0, o.x1);
}

// Regime gains (cu:362-374) times the thalamic scale.
inline void regime_gains(int regime, float scale, float* ge, float* gi)
{
    float e = 5.0f, g = 2.0f;
    if (regime == 0) { e = 2.5f; g = 1.0f; }
    if (regime == 2) { e = 7.5f; g = 3.0f; }
    *ge = e * scale;
    *gi = g * scale;
}

struct Params { float a, b, c, d; };

// initNeuron (cu:286-320) with r*r for powf(r, 2).
G4S_HD Params neuron_params(uint64_t seed, uint32_t i, uint32_t n_exc)
{
    const float r = u01(draw(seed, i, 0u, kTagParam).x0);
    Params p;
    if (i < n_exc) {
        const float r2 = r * r;
        p.a = 0.02f;
        p.b = 0.2f;
        p.c = -65.0f + 15.0f * r2;
        p.d = 8.0f - 6.0f * r2;
    } else {
        p.a = 0.02f + 0.08f * r;
        p.b = 0.25f - 0.05f * r;
        p.c = -65.0f;
        p.d = 2.0f;
    }
    return p;
}

// Izhikevich step of stateupdate (cu:386-407): two 0.5 ms Euler half steps
// of v in the reference's left-to-right order, then u, then threshold and
// reset.  Returns true on a spike (v >= 30 or v not finite).
G4S_HD bool izh_step(float& v, float& u, float a, float b, float c, float d, float I)
{
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float q = 0.04f * v;
        q = q * v;
        const float l = 5.0f * v;
        float s = q + l;
        s = s + 140.0f;
        s = s - u;
        s = s + I;
        v = v + 0.5f * s;
    }
    const float bv = b * v;
    u = u + a * (bv - u);
    const bool spk = (v >= 30.0f) || !__builtin_isfinite(v);
    if (spk) {
        v = c;
        u = u + d;
    }
    return spk;
}

}  // namespace g4s
