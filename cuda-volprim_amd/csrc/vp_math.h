// vp_math.h — deterministic FP32 math for the volumetric path tracer.
//
// Every function here is built only from IEEE-754 operations that are
// correctly rounded on both gfx950 and x86-64 (add, mul, fma, div, sqrt,
// rint, floor, bit manipulation).  Compiled with -ffp-contract=off on both
// sides, a HIP kernel and a host build of the same expression produce the
// same bits.  That is what lets the GPU renderer be checked bit for bit
// against a CPU restatement of the estimator.
//
// Methods (all from the mathematics or a named published method):
//   exp   : Cody–Waite reduction by ln2 (two-constant, fma), degree-7 Taylor
//           polynomial on |r| <= ln2/2, scaling by 2^k split in two halves.
//   log   : x = m 2^e, m in [sqrt(1/2), sqrt(2)); log(m) = 2 atanh(s) with
//           s = f/(2+f), Taylor series of atanh to s^11.
//   sin/cos: Cody–Waite reduction by pi/2 (three constants), Taylor
//           polynomials on |r| <= pi/4.
//   atan2 : octant reduction, atan(t) = pi/6 + atan((sqrt3 t - 1)/(t + sqrt3))
//           for t > 2 - sqrt3, Taylor series to u^13.
//   erf   : x P(x^2) on |x| <= 1; erfc(x) = exp(-x^2) erfcx(x) on x > 1 with
//           erfcx from three centred polynomial fits (gen_math_coeffs.py).
//   erfinv: variable substitution of M. Giles, "Approximating the erfinv
//           function" (GPU Computing Gems, 2011), coefficients refitted
//           (gen_math_coeffs.py).
//   chord : the Taylor series of e^(x^2) erf(x) (all terms positive) in
//           d = 2 x^2, refitted as a minimax polynomial on [0, 9].
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define VP_HD __host__ __device__ __forceinline__
#else
#define VP_HD inline __attribute__((always_inline))
#endif

// VP_FAST_MATH=1 (device code only) replaces the correctly rounded
// reciprocal, square root and reciprocal square root and the polynomial
// exp/log with the hardware's approximate instructions (v_rcp_f32,
// v_sqrt_f32, v_rsq_f32, v_exp_f32, v_log_f32; about 1 ulp).  Results then
// agree with the host build statistically, not bit for bit.
#if defined(VP_FAST_MATH) && VP_FAST_MATH && defined(__HIP_DEVICE_COMPILE__)
#define VP_FAST_DEVICE 1
#else
#define VP_FAST_DEVICE 0
#endif

namespace vp {

// ---------------------------------------------------------------------------
// Primitive operations
// ---------------------------------------------------------------------------
VP_HD float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
VP_HD float sqrtf_(float x) {
#if VP_FAST_DEVICE
    return __builtin_amdgcn_sqrtf(x);
#else
    return __builtin_sqrtf(x);
#endif
}
VP_HD float fabsf_(float x) { return __builtin_fabsf(x); }
VP_HD float floorf_(float x) { return __builtin_floorf(x); }
VP_HD float rintf_(float x) { return __builtin_rintf(x); }
VP_HD float copysignf_(float x, float y) { return __builtin_copysignf(x, y); }
VP_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
VP_HD float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }
VP_HD bool isnanf_(float x) { return x != x; }
VP_HD bool isfinitef_(float x) { return (f2u(x) & 0x7f800000u) != 0x7f800000u; }

// Comparison-based min/max/clamp with a fixed NaN behaviour:
//   minf(a, b) = a < b ? a : b,  maxf(a, b) = a > b ? a : b,
//   clampf(x, lo, hi) = maxf(lo, minf(x, hi))   (so clampf(NaN, lo, hi) = hi).
VP_HD float minf(float a, float b) { return a < b ? a : b; }
VP_HD float maxf(float a, float b) { return a > b ? a : b; }
VP_HD float clampf(float x, float lo, float hi) { return maxf(lo, minf(x, hi)); }
#if VP_FAST_DEVICE
VP_HD float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
VP_HD float rsqrtf_(float x) { return __builtin_amdgcn_rsqf(x); }
#else
VP_HD float rcpf(float x) { return 1.0f / x; }
VP_HD float rsqrtf_(float x) { return 1.0f / sqrtf_(x); }
#endif

constexpr float kInf = __builtin_huge_valf();
constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kOneOverPi = 1.0f / kPi;
constexpr float kOneOverTwoPi = 1.0f / (2.0f * kPi);
constexpr float kOneOverFourPi = 1.0f / (4.0f * kPi);
constexpr float kOneOverTwoPiPow32 = 0.0634936359f;  // (2 pi)^(-3/2)
constexpr float kRootTwoPi = 2.5066282746f;          // sqrt(2 pi)
constexpr float kRootHalfPi = kRootTwoPi * 0.5f;     // sqrt(pi/2)
constexpr float kRootTwo = 1.41421356237309504880f;
constexpr float kOneOverRootTwo = 1.0f / kRootTwo;
constexpr float kDegToRad = kPi / 180.0f;

// ---------------------------------------------------------------------------
// exp
// ---------------------------------------------------------------------------
VP_HD float exp2i_(int k) { return u2f(static_cast<uint32_t>(k + 127) << 23); }

VP_HD float expf_(float x) {
#if VP_FAST_DEVICE
    return __builtin_amdgcn_exp2f(x * 1.44269502162933349609375f);
#endif
    if (isnanf_(x)) return x;
    if (x > 88.72283935546875f) return kInf;
    if (x < -103.97208404541015625f) return 0.0f;
    const float k = rintf_(x * 1.44269502162933349609375f);
    float r = fmaf_(-k, 0.693147182464599609375f, x);
    r = fmaf_(-k, -0x1.05c610p-29f, r);
    float p = 1.0f / 5040.0f;
    p = fmaf_(p, r, 1.0f / 720.0f);
    p = fmaf_(p, r, 1.0f / 120.0f);
    p = fmaf_(p, r, 1.0f / 24.0f);
    p = fmaf_(p, r, 1.0f / 6.0f);
    p = fmaf_(p, r, 0.5f);
    p = fmaf_(p, r, 1.0f);
    p = fmaf_(p, r, 1.0f);
    const int ki = static_cast<int>(k);
    const int k1 = ki >> 1;  // floor(ki / 2): both halves stay in the normal range
    const int k2 = ki - k1;
    return (p * exp2i_(k1)) * exp2i_(k2);
}

// ---------------------------------------------------------------------------
// log
// ---------------------------------------------------------------------------
VP_HD float logf_(float x) {
#if VP_FAST_DEVICE
    return __builtin_amdgcn_logf(x) * 0.693147182464599609375f;
#endif
    if (isnanf_(x) || x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -kInf;
    if (x == kInf) return kInf;
    uint32_t b = f2u(x);
    int e = 0;
    if (b < 0x00800000u) {  // subnormal: scale into the normal range
        x = x * 8388608.0f;
        b = f2u(x);
        e = -23;
    }
    e += static_cast<int>(b >> 23) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float R = 2.0f / 11.0f;
    R = fmaf_(R, z, 2.0f / 9.0f);
    R = fmaf_(R, z, 2.0f / 7.0f);
    R = fmaf_(R, z, 2.0f / 5.0f);
    R = fmaf_(R, z, 2.0f / 3.0f);
    R = R * z;
    const float hfsq = 0.5f * f * f;
    const float lg = f - (hfsq - s * (hfsq + R));
    const float fe = static_cast<float>(e);
    return fmaf_(fe, 0.693145751953125f, fmaf_(fe, 1.428606765330187045e-06f, lg));
}

// ---------------------------------------------------------------------------
// sin / cos (arguments up to a few thousand radians)
// ---------------------------------------------------------------------------
VP_HD void sincosf_(float x, float* s_out, float* c_out) {
    const float q = rintf_(x * 0.636619746685028076171875f);  // 2/pi
    float r = fmaf_(-q, 0x1.921fb6p+0f, x);
    r = fmaf_(-q, -0x1.777a5cp-25f, r);
    r = fmaf_(-q, -0x1.ee59dap-50f, r);
    const float z = r * r;
    float ps = -1.0f / 39916800.0f;
    ps = fmaf_(ps, z, 1.0f / 362880.0f);
    ps = fmaf_(ps, z, -1.0f / 5040.0f);
    ps = fmaf_(ps, z, 1.0f / 120.0f);
    ps = fmaf_(ps, z, -1.0f / 6.0f);
    const float sr = fmaf_(r * z, ps, r);
    float pc = 1.0f / 479001600.0f;
    pc = fmaf_(pc, z, -1.0f / 3628800.0f);
    pc = fmaf_(pc, z, 1.0f / 40320.0f);
    pc = fmaf_(pc, z, -1.0f / 720.0f);
    pc = fmaf_(pc, z, 1.0f / 24.0f);
    const float cr = fmaf_(z * z, pc, fmaf_(-0.5f, z, 1.0f));
    const int n = static_cast<int>(q) & 3;
    float s, c;
    if (n == 0) {
        s = sr;
        c = cr;
    } else if (n == 1) {
        s = cr;
        c = -sr;
    } else if (n == 2) {
        s = -sr;
        c = -cr;
    } else {
        s = -cr;
        c = sr;
    }
    *s_out = s;
    *c_out = c;
}
VP_HD float sinf_(float x) {
    float s, c;
    sincosf_(x, &s, &c);
    return s;
}
VP_HD float cosf_(float x) {
    float s, c;
    sincosf_(x, &s, &c);
    return c;
}

// ---------------------------------------------------------------------------
// atan2 / acos
// ---------------------------------------------------------------------------
VP_HD float atan2f_(float y, float x) {
    const float ax = fabsf_(x);
    const float ay = fabsf_(y);
    float a;
    if (ax == 0.0f && ay == 0.0f) {
        a = (f2u(x) >> 31) ? kPi : 0.0f;
    } else {
        const bool swap = ay > ax;
        const float t = swap ? ax / ay : ay / ax;  // in [0, 1]
        float base = 0.0f;
        float u = t;
        if (t > 0.267949193445499f) {  // 2 - sqrt(3)
            u = fmaf_(t, 1.73205080756887729f, -1.0f) / (t + 1.73205080756887729f);
            base = 0.52359877559829887f;  // pi/6
        }
        const float z = u * u;
        float p = 1.0f / 13.0f;
        p = fmaf_(p, z, -1.0f / 11.0f);
        p = fmaf_(p, z, 1.0f / 9.0f);
        p = fmaf_(p, z, -1.0f / 7.0f);
        p = fmaf_(p, z, 1.0f / 5.0f);
        p = fmaf_(p, z, -1.0f / 3.0f);
        a = base + fmaf_(u * z, p, u);
        if (swap) a = (0x1.921fb6p+0f - a) + -0x1.777a5cp-25f;
        if (f2u(x) >> 31) a = (0x1.921fb6p+1f - a) + -0x1.777a5cp-24f;
    }
    return copysignf_(a, y);
}

// acos(y) for y in [-1, 1] written as atan2(sqrt((1-y)(1+y)), y).
VP_HD float acosf_(float y) { return atan2f_(sqrtf_((1.0f - y) * (1.0f + y)), y); }

// ---------------------------------------------------------------------------
// erf / erfc / erfinv
// ---------------------------------------------------------------------------
// erf(x)/x on [0, 1] in z = x^2.
constexpr float kErfP[7] = {
    0x1.20dd74p+0f, -0x1.81273ep-2f, 0x1.ce2cf8p-4f, -0x1.b7f912p-6f,
    0x1.5405d0p-8f, -0x1.a3f7bap-11f, 0x1.496bf8p-14f,
};
// exp(x^2) erfc(x) on [1, 2.2], [2.2, 4.5], [4.5, 10.1] in u = (1/x - m) s.
constexpr float kErfcxM[3] = {0x1.745d18p-1f, 0x1.5a814ap-2f, 0x1.48f112p-3f};
constexpr float kErfcxS[3] = {0x1.d55556p+1f, 0x1.137a70p+3f, 0x1.03b6dcp+4f};
constexpr float kErfcx[3][9] = {
    {0x1.5f88f6p-2f, 0x1.854636p-4f, -0x1.80f1a2p-7f, 0x1.05c45ep-10f, 0x1.de9856p-16f,
     -0x1.454288p-15f, 0x1.7753a0p-17f, -0x1.200130p-19f, 0x1.e10e88p-23f},
    {0x1.73a57cp-3f, 0x1.cfea9ap-5f, -0x1.422894p-9f, -0x1.347414p-14f, 0x1.e2d914p-16f,
     -0x1.7f0e6ap-19f, 0x1.05e1c6p-28f, 0x1.adb966p-25f, -0x1.30bdbep-27f},
    {0x1.6e8ebep-4f, 0x1.125e64p-5f, -0x1.df5fd0p-12f, -0x1.7af86cp-15f, 0x1.b25348p-19f,
     0x1.f71024p-25f, -0x1.9725f8p-26f, 0x1.3beadcp-30f, 0x1.0d1178p-33f},
};
// erfinv(x)/x in w - 2.5 (w < 5) and sqrt(w) - 3 (w >= 5), w = -log((1-x)(1+x)).
constexpr float kErfinvA[10] = {
    0x1.805c5ep+0f, 0x1.f91e8ap-3f, -0x1.11c4ecp-8f, -0x1.486616p-10f, 0x1.c9e6c8p-13f,
    -0x1.4d15a4p-18f, -0x1.d0c1b4p-19f, 0x1.f48bd4p-22f, 0x1.8f2344p-26f, -0x1.2e0e4ep-27f,
};
constexpr float kErfinvB[10] = {
    0x1.6a9efcp+1f, 0x1.006dc6p+0f, 0x1.3580cep-7f, -0x1.f48470p-8f, 0x1.74b38cp-8f,
    -0x1.d784f0p-9f, 0x1.854f58p-10f, -0x1.dc4896p-15f, -0x1.449e5ap-12f, 0x1.e5506cp-14f,
};

// erf on |x| <= 1 (sign carried by x).
VP_HD float erf_small_(float x) {
    const float z = x * x;
    float p = kErfP[6];
    p = fmaf_(p, z, kErfP[5]);
    p = fmaf_(p, z, kErfP[4]);
    p = fmaf_(p, z, kErfP[3]);
    p = fmaf_(p, z, kErfP[2]);
    p = fmaf_(p, z, kErfP[1]);
    p = fmaf_(p, z, kErfP[0]);
    return x * p;
}

// erfc(a) for 1 <= a; 0 beyond 10.1 (below the smallest subnormal).
VP_HD float erfc_tail_(float a) {
    if (a > 10.1f) return 0.0f;
    const int j = a < 2.2f ? 0 : (a < 4.5f ? 1 : 2);
    const float u = (1.0f / a - kErfcxM[j]) * kErfcxS[j];
    const float* c = kErfcx[j];
    float p = c[8];
    p = fmaf_(p, u, c[7]);
    p = fmaf_(p, u, c[6]);
    p = fmaf_(p, u, c[5]);
    p = fmaf_(p, u, c[4]);
    p = fmaf_(p, u, c[3]);
    p = fmaf_(p, u, c[2]);
    p = fmaf_(p, u, c[1]);
    p = fmaf_(p, u, c[0]);
    // exp(-a^2) with a^2 split into hi + lo so the rounding of a*a does not
    // enter the exponent: exp(-hi - lo) = exp(-hi) (1 - lo) to first order.
    const float hi = a * a;
    const float lo = fmaf_(a, a, -hi);
    float e = expf_(-hi);
    e = fmaf_(-e, lo, e);
    return e * p;
}

VP_HD float erff_(float x) {
    const float a = fabsf_(x);
    float r;
    if (a <= 1.0f) return erf_small_(x);
    if (a >= 3.92f) r = 1.0f;
    else r = 1.0f - erfc_tail_(a);
    return copysignf_(r, x);
}

VP_HD float erfcf_(float x) {
    if (isnanf_(x)) return x;
    const float a = fabsf_(x);
    if (a <= 1.0f) return 1.0f - erf_small_(x);
    if (x > 0.0f) return erfc_tail_(a);
    return 2.0f - erfc_tail_(a);
}

// erf(b) - erf(a) for a <= b, evaluated through erfc when both arguments lie
// on the same side of zero so the difference does not cancel.
VP_HD float erf_diff_(float a, float b) {
    if (a >= 0.0f) return erfcf_(a) - erfcf_(b);
    if (b <= 0.0f) return erfcf_(-b) - erfcf_(-a);
    return erff_(b) - erff_(a);
}

// S(d) = sum_n d^n / (2n+1)!! = sqrt(pi/2) e^(d/2) erf(sqrt(d/2)) / sqrt(d) on
// [0, 9], a minimax fit of degree 10 in d (gen_math_coeffs.py; relative error
// below 4e-7 in float32 Horner).  With it the optical depth of a whole chord
// through a 3-sigma-truncated Gaussian needs no exp and no erf: the chord of
// squared local half-length d has depth 2 e^(-9/2) M dnf sqrt(d)/|w| S(d).
constexpr float kChordS[11] = {
    0x1.000002p+0f, 0x1.5554b8p-2f, 0x1.111e7ap-4f, 0x1.373554p-7f, 0x1.1c9208p-10f, 0x1.504234p-14f,
    0x1.b774c6p-17f, -0x1.9bb328p-21f, 0x1.af65b8p-23f, -0x1.9ee29ap-27f, 0x1.452d92p-31f,
};
VP_HD float chord_poly_(float d) {
    float p = kChordS[10];
    for (int i = 9; i >= 0; --i) p = fmaf_(p, d, kChordS[i]);
    return p;
}

VP_HD float erfinvf_(float x) {
    if (isnanf_(x) || x > 1.0f || x < -1.0f) return __builtin_nanf("");
    if (x == 1.0f) return kInf;
    if (x == -1.0f) return -kInf;
    const float w = -logf_((1.0f - x) * (1.0f + x));
    float p;
    if (w < 5.0f) {
        const float u = w - 2.5f;
        p = kErfinvA[9];
        for (int i = 8; i >= 0; --i) p = fmaf_(p, u, kErfinvA[i]);
    } else {
        const float u = sqrtf_(w) - 3.0f;
        p = kErfinvB[9];
        for (int i = 8; i >= 0; --i) p = fmaf_(p, u, kErfinvB[i]);
    }
    return p * x;
}

}  // namespace vp
