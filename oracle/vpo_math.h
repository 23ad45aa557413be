/* vpo_math.h -- oracle copy of the FP32 transcendental functions.
 *
 * TEST INFRASTRUCTURE.  This file belongs to the CPU oracle under oracle/;
 * only tests, the smoke check and the CPU-baseline leg of the benchmark use
 * it.  The GPU renderer never links it.
 *
 * Each function restates, in plain C, the same published method the product
 * uses (Cody-Waite + Taylor exp; atanh-series log; Cody-Waite + Taylor
 * sin/cos; octant-reduced Taylor atan2; x P(x^2) / exp(-x^2) erfcx(x) erf;
 * Giles-substitution erfinv) with the same float32 constants, so that a
 * correctly rounded evaluation on x86-64 reproduces the kernel's bits.
 * Accuracy of every function is pinned in tests/test_oracle_math.py against
 * mpmath at 40 digits.  Compile with -ffp-contract=off.
 */
#ifndef VPO_MATH_H
#define VPO_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint32_t vpo_f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float vpo_u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
static inline float vpo_min(float a, float b) { return a < b ? a : b; }
static inline float vpo_max(float a, float b) { return a > b ? a : b; }
static inline float vpo_clamp(float x, float lo, float hi) { return vpo_max(lo, vpo_min(x, hi)); }
static inline float vpo_rcp(float x) { return 1.0f / x; }
static inline float vpo_rsqrt(float x) { return 1.0f / sqrtf(x); }

#define VPO_INF ((float)INFINITY)
#define VPO_PI 3.14159265358979323846f

static inline float vpo_exp2i(int k) { return vpo_u2f((uint32_t)(k + 127) << 23); }

static inline float vpo_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return VPO_INF;
    if (x < -103.97208404541015625f) return 0.0f;
    float k = rintf(x * 1.44269502162933349609375f);
    float r = fmaf(-k, 0.693147182464599609375f, x);
    r = fmaf(-k, -0x1.05c610p-29f, r);
    float p = 1.0f / 5040.0f;
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int ki = (int)k;
    int k1 = ki >> 1;
    int k2 = ki - k1;
    return (p * vpo_exp2i(k1)) * vpo_exp2i(k2);
}

static inline float vpo_logf(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -VPO_INF;
    if (x == VPO_INF) return VPO_INF;
    uint32_t b = vpo_f2u(x);
    int e = 0;
    if (b < 0x00800000u) {
        x = x * 8388608.0f;
        b = vpo_f2u(x);
        e = -23;
    }
    e += (int)(b >> 23) - 127;
    float m = vpo_u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = 2.0f / 11.0f;
    R = fmaf(R, z, 2.0f / 9.0f);
    R = fmaf(R, z, 2.0f / 7.0f);
    R = fmaf(R, z, 2.0f / 5.0f);
    R = fmaf(R, z, 2.0f / 3.0f);
    R = R * z;
    float hfsq = 0.5f * f * f;
    float lg = f - (hfsq - s * (hfsq + R));
    float fe = (float)e;
    return fmaf(fe, 0.693145751953125f, fmaf(fe, 1.428606765330187045e-06f, lg));
}

static inline void vpo_sincosf(float x, float* so, float* co) {
    float q = rintf(x * 0.636619746685028076171875f);
    float r = fmaf(-q, 0x1.921fb6p+0f, x);
    r = fmaf(-q, -0x1.777a5cp-25f, r);
    r = fmaf(-q, -0x1.ee59dap-50f, r);
    float z = r * r;
    float ps = -1.0f / 39916800.0f;
    ps = fmaf(ps, z, 1.0f / 362880.0f);
    ps = fmaf(ps, z, -1.0f / 5040.0f);
    ps = fmaf(ps, z, 1.0f / 120.0f);
    ps = fmaf(ps, z, -1.0f / 6.0f);
    float sr = fmaf(r * z, ps, r);
    float pc = 1.0f / 479001600.0f;
    pc = fmaf(pc, z, -1.0f / 3628800.0f);
    pc = fmaf(pc, z, 1.0f / 40320.0f);
    pc = fmaf(pc, z, -1.0f / 720.0f);
    pc = fmaf(pc, z, 1.0f / 24.0f);
    float cr = fmaf(z * z, pc, fmaf(-0.5f, z, 1.0f));
    int n = (int)q & 3;
    switch (n) {
        case 0: *so = sr; *co = cr; break;
        case 1: *so = cr; *co = -sr; break;
        case 2: *so = -sr; *co = -cr; break;
        default: *so = -cr; *co = sr; break;
    }
}
static inline float vpo_sinf(float x) { float s, c; vpo_sincosf(x, &s, &c); return s; }
static inline float vpo_cosf(float x) { float s, c; vpo_sincosf(x, &s, &c); return c; }

static inline float vpo_atan2f(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y), a;
    if (ax == 0.0f && ay == 0.0f) {
        a = (vpo_f2u(x) >> 31) ? VPO_PI : 0.0f;
    } else {
        int swap = ay > ax;
        float t = swap ? ax / ay : ay / ax;
        float base = 0.0f, u = t;
        if (t > 0.267949193445499f) {
            u = fmaf(t, 1.73205080756887729f, -1.0f) / (t + 1.73205080756887729f);
            base = 0.52359877559829887f;
        }
        float z = u * u;
        float p = 1.0f / 13.0f;
        p = fmaf(p, z, -1.0f / 11.0f);
        p = fmaf(p, z, 1.0f / 9.0f);
        p = fmaf(p, z, -1.0f / 7.0f);
        p = fmaf(p, z, 1.0f / 5.0f);
        p = fmaf(p, z, -1.0f / 3.0f);
        a = base + fmaf(u * z, p, u);
        if (swap) a = (0x1.921fb6p+0f - a) + -0x1.777a5cp-25f;
        if (vpo_f2u(x) >> 31) a = (0x1.921fb6p+1f - a) + -0x1.777a5cp-24f;
    }
    return copysignf(a, y);
}
static inline float vpo_acosf(float y) { return vpo_atan2f(sqrtf((1.0f - y) * (1.0f + y)), y); }

static const float VPO_ERF_P[7] = {
    0x1.20dd74p+0f, -0x1.81273ep-2f, 0x1.ce2cf8p-4f, -0x1.b7f912p-6f,
    0x1.5405d0p-8f, -0x1.a3f7bap-11f, 0x1.496bf8p-14f,
};
static const float VPO_ERFCX_M[3] = {0x1.745d18p-1f, 0x1.5a814ap-2f, 0x1.48f112p-3f};
static const float VPO_ERFCX_S[3] = {0x1.d55556p+1f, 0x1.137a70p+3f, 0x1.03b6dcp+4f};
static const float VPO_ERFCX[3][9] = {
    {0x1.5f88f6p-2f, 0x1.854636p-4f, -0x1.80f1a2p-7f, 0x1.05c45ep-10f, 0x1.de9856p-16f,
     -0x1.454288p-15f, 0x1.7753a0p-17f, -0x1.200130p-19f, 0x1.e10e88p-23f},
    {0x1.73a57cp-3f, 0x1.cfea9ap-5f, -0x1.422894p-9f, -0x1.347414p-14f, 0x1.e2d914p-16f,
     -0x1.7f0e6ap-19f, 0x1.05e1c6p-28f, 0x1.adb966p-25f, -0x1.30bdbep-27f},
    {0x1.6e8ebep-4f, 0x1.125e64p-5f, -0x1.df5fd0p-12f, -0x1.7af86cp-15f, 0x1.b25348p-19f,
     0x1.f71024p-25f, -0x1.9725f8p-26f, 0x1.3beadcp-30f, 0x1.0d1178p-33f},
};
static const float VPO_ERFINV_A[10] = {
    0x1.805c5ep+0f, 0x1.f91e8ap-3f, -0x1.11c4ecp-8f, -0x1.486616p-10f, 0x1.c9e6c8p-13f,
    -0x1.4d15a4p-18f, -0x1.d0c1b4p-19f, 0x1.f48bd4p-22f, 0x1.8f2344p-26f, -0x1.2e0e4ep-27f,
};
static const float VPO_ERFINV_B[10] = {
    0x1.6a9efcp+1f, 0x1.006dc6p+0f, 0x1.3580cep-7f, -0x1.f48470p-8f, 0x1.74b38cp-8f,
    -0x1.d784f0p-9f, 0x1.854f58p-10f, -0x1.dc4896p-15f, -0x1.449e5ap-12f, 0x1.e5506cp-14f,
};

static inline float vpo_erf_small(float x) {
    float z = x * x, p = VPO_ERF_P[6];
    for (int i = 5; i >= 0; --i) p = fmaf(p, z, VPO_ERF_P[i]);
    return x * p;
}
static inline float vpo_erfc_tail(float a) {
    if (a > 10.1f) return 0.0f;
    int j = a < 2.2f ? 0 : (a < 4.5f ? 1 : 2);
    float u = (1.0f / a - VPO_ERFCX_M[j]) * VPO_ERFCX_S[j];
    float p = VPO_ERFCX[j][8];
    for (int i = 7; i >= 0; --i) p = fmaf(p, u, VPO_ERFCX[j][i]);
    float hi = a * a;
    float lo = fmaf(a, a, -hi);
    float e = vpo_expf(-hi);
    e = fmaf(-e, lo, e);
    return e * p;
}
static inline float vpo_erff(float x) {
    float a = fabsf(x), r;
    if (a <= 1.0f) return vpo_erf_small(x);
    if (a >= 3.92f) r = 1.0f;
    else r = 1.0f - vpo_erfc_tail(a);
    return copysignf(r, x);
}
static inline float vpo_erfcf(float x) {
    if (x != x) return x;
    float a = fabsf(x);
    if (a <= 1.0f) return 1.0f - vpo_erf_small(x);
    if (x > 0.0f) return vpo_erfc_tail(a);
    return 2.0f - vpo_erfc_tail(a);
}
static inline float vpo_erf_diff(float a, float b) {
    if (a >= 0.0f) return vpo_erfcf(a) - vpo_erfcf(b);
    if (b <= 0.0f) return vpo_erfcf(-b) - vpo_erfcf(-a);
    return vpo_erff(b) - vpo_erff(a);
}
/* S(d) = sum_n d^n/(2n+1)!! = sqrt(pi/2) e^(d/2) erf(sqrt(d/2))/sqrt(d) on
 * [0, 9]: degree-10 minimax fit (cuda-volprim_amd/csrc/gen_math_coeffs.py),
 * the same float32 coefficients and Horner order as the product. */
static const float VPO_CHORD_S[11] = {
    0x1.000002p+0f, 0x1.5554b8p-2f, 0x1.111e7ap-4f, 0x1.373554p-7f, 0x1.1c9208p-10f, 0x1.504234p-14f,
    0x1.b774c6p-17f, -0x1.9bb328p-21f, 0x1.af65b8p-23f, -0x1.9ee29ap-27f, 0x1.452d92p-31f,
};
static inline float vpo_chord_poly(float d) {
    float p = VPO_CHORD_S[10];
    for (int i = 9; i >= 0; --i) p = fmaf(p, d, VPO_CHORD_S[i]);
    return p;
}

static inline float vpo_erfinvf(float x) {
    if (x != x || x > 1.0f || x < -1.0f) return NAN;
    if (x == 1.0f) return VPO_INF;
    if (x == -1.0f) return -VPO_INF;
    float w = -vpo_logf((1.0f - x) * (1.0f + x)), p, u;
    if (w < 5.0f) {
        u = w - 2.5f;
        p = VPO_ERFINV_A[9];
        for (int i = 8; i >= 0; --i) p = fmaf(p, u, VPO_ERFINV_A[i]);
    } else {
        u = sqrtf(w) - 3.0f;
        p = VPO_ERFINV_B[9];
        for (int i = 8; i >= 0; --i) p = fmaf(p, u, VPO_ERFINV_B[i]);
    }
    return p * x;
}

#endif
