/* vpo.c -- CPU oracle for the Gaussian-mixture volumetric path tracer.
 *
 * TEST INFRASTRUCTURE, NOT THE PRODUCT.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py load this library, and only as the
 * checker (or the timed CPU baseline).  The GPU renderer never calls it.
 *
 * A plain-C restatement of kfernandez31/cuda-volprim's render path, brute
 * force over all primitives (no acceleration structure), multi-threaded over
 * pixel rows.  Every function cites the reference lines it follows.  The
 * deliberate, documented differences from the reference are the ones the
 * GPU build makes too:
 *   - the per-primitive free-flight uniform is a counter hash of
 *     (path, bounce, role, primitive) instead of the next PCG draw
 *     (the reference's own cap-free experiment, latex/chapters/06-optimization.tex:470-476),
 *     so the argmin does not depend on the order primitives are visited in;
 *   - active sets are kept sorted by primitive index (set semantics as
 *     device/include CompactSet, capacity cap_active);
 *   - optical depth along a shadow ray is summed in 24.40 fixed point and the
 *     ray returns 0 as soon as the total reaches 88;
 *   - erf(b) - erf(a) is evaluated through erfc on the tails;
 *   - the environment is importance-sampled through an alias table over the
 *     same texel distribution (joint/total) as the reference's 2D CDF; the
 *     reference CDF is restated too (vpo_env_ref_cdf) so tests can pin the
 *     distribution;
 *   - bilinear environment lookup with exact float weights;
 *   - the entry hit list is unbounded (the reference drops beyond 128);
 *   - a pixel's mean is the exact 32.32 fixed-point sum of its samples
 *     divided by their count (accum_add / accum_mean), where the reference
 *     folds a running Welford mean in sample order (raygen.cuh:214-238): same
 *     estimator, rounded once at the end instead of once per sample.
 * Each function is pinned by tests/test_oracle_pins.py.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off).
 */
#include "vpo_math.h"

#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* vectors                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x + (-b.x), a.y + (-b.y), a.z + (-b.z)); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vfma(float a, v3 b, v3 c) { return V(fmaf(a, b.x, c.x), fmaf(a, b.y, c.y), fmaf(a, b.z, c.z)); }
/* math.h:254-256 */
static inline float vdot(v3 a, v3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
/* math.h:306-310 */
static inline v3 vcross(v3 a, v3 b) {
    return V(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
static inline float vmax3(v3 v) { return vpo_max(vpo_max(v.x, v.y), v.z); }
static inline v3 vnormalize(v3 v) { return vscale(v, vpo_rsqrt(vdot(v, v))); }

/* ------------------------------------------------------------------------ */
/* primitive record (primitive.h:27-87): 16 floats                           */
/* c.xyz, qconj.wxyz, rcp_scale.xyz, dnf, icf, albedo.rgb, mass              */
/* ------------------------------------------------------------------------ */
typedef struct {
    float cx, cy, cz, qw, qx, qy, qz, rsx, rsy, rsz, dnf, icf, ar, ag, ab, mass;
} prim_t;

#define ONE_OVER_TWO_PI_POW_3_2 0.0634936359f
#define ROOT_TWO_PI 2.5066282746f
#define ROOT_HALF_PI (ROOT_TWO_PI * 0.5f)
#define ROOT_TWO 1.41421356237309504880f
#define ONE_OVER_ROOT_TWO (1.0f / ROOT_TWO)
#define ONE_OVER_FOUR_PI (1.0f / (4.0f * VPO_PI))
#define INF_F 1e20f
#define EXTENT2 9.0f

/* Primitive(center, q.conjugate(), scale, albedo, M) with q = UnitQuaternion::from
 * (quat.h:54-57, primitive.h:70-87,108-111). */
void vpo_make_prims(const float* in, long n, float* out) {
    for (long i = 0; i < n; ++i) {
        const float* r = in + 14 * i;
        prim_t* p = (prim_t*)(out + 16 * i);
        float w = r[3], x = r[4], y = r[5], z = r[6];
        float l2 = fmaf(w, w, fmaf(x, x, fmaf(y, y, z * z)));
        float s = 1.0f / sqrtf(l2);
        p->cx = r[0]; p->cy = r[1]; p->cz = r[2];
        p->qw = w * s; p->qx = -(x * s); p->qy = -(y * s); p->qz = -(z * s);
        p->rsx = 1.0f / r[7]; p->rsy = 1.0f / r[8]; p->rsz = 1.0f / r[9];
        p->dnf = ONE_OVER_TWO_PI_POW_3_2 * (p->rsx * p->rsy * p->rsz);
        p->icf = r[13] * p->dnf * ROOT_HALF_PI;
        p->ar = r[10]; p->ag = r[11]; p->ab = r[12];
        p->mass = r[13];
    }
}

/* quat.h:67-71 */
static inline v3 qrot(const prim_t* p, v3 v) {
    v3 u = V(p->qx, p->qy, p->qz);
    v3 uv = vcross(u, v);
    v3 uuv = vcross(u, uv);
    return vfma(2.0f, vfma(p->qw, uv, uuv), v);
}
/* primitive.h:98-104 */
static inline v3 local_pos(const prim_t* p, v3 x) { return vmul(qrot(p, vsub(x, V(p->cx, p->cy, p->cz))), V(p->rsx, p->rsy, p->rsz)); }
static inline v3 local_dir(const prim_t* p, v3 d) { return vmul(qrot(p, d), V(p->rsx, p->rsy, p->rsz)); }

/* primitive.h:127-135 */
static float prim_pdf(const prim_t* p, v3 x) {
    v3 l = local_pos(p, x);
    return vpo_expf(-0.5f * vdot(l, l)) * p->dnf;
}

/* primitive.h:140-186 */
static float inv_cdf(const prim_t* pr, v3 o, v3 d, float tau) {
    v3 w = local_dir(pr, d), p = local_pos(pr, o);
    float wl2 = vdot(w, w);
    float wil = vpo_rsqrt(wl2);
    float wp = vdot(w, p) * wil;
    float pp = vdot(p, p);
    float diff = fmaf(-wp, wp, pp);
    float K = wil * vpo_expf(-0.5f * diff) * pr->icf;
    float raw = vpo_erff(wp * ONE_OVER_ROOT_TWO) + tau * vpo_rcp(K);
    if (raw > 1.0f) return INF_F;
    float v = vpo_erfinvf(vpo_clamp(raw, -1.0f, 1.0f));
    return (ROOT_TWO * v - wp) * wil;
}

/* primitive.h:203-246 */
static float inv_cdf_span(const prim_t* pr, v3 o, v3 d, float toff, float tau) {
    v3 w = local_dir(pr, d), p = local_pos(pr, o);
    float wl2 = vdot(w, w);
    float wil = vpo_rsqrt(wl2);
    float wl = wl2 * wil;
    float wp = vdot(w, p) * wil;
    float pp = vdot(p, p);
    float diff = fmaf(-wp, wp, pp);
    float K = wil * vpo_expf(-0.5f * diff) * pr->icf;
    float wpo = fmaf(toff, wl, wp);
    float raw = vpo_erff(wpo * ONE_OVER_ROOT_TWO) + tau * vpo_rcp(K);
    if (raw > 1.0f) return INF_F;
    float v = vpo_erfinvf(vpo_clamp(raw, -1.0f, 1.0f));
    return (ROOT_TWO * v - wp) * wil;
}

/* primitive.h:249-318 (erf difference through erfc on the tails) */
static float optical_depth(const prim_t* pr, v3 o, v3 d, float t0, float t1) {
    if (t1 - t0 <= 1e-6f) return 0.0f;
    v3 w = local_dir(pr, d), p = local_pos(pr, o);
    float wil = vpo_rsqrt(vdot(w, w));
    float t0s = vpo_max(0.0f, t0);
    float t1s = vpo_max(t0s + 1e-6f, t1);
    v3 wn = vscale(w, wil);
    v3 ps = vadd(p, vscale(w, t0s));
    float tl = vpo_clamp((t1s - t0s) * vpo_rcp(wil), 0.0f, 6.0f);
    float B = vdot(wn, ps);
    float C = vdot(ps, ps);
    float perp = fmaf(-B, B, C);
    float e = vpo_expf(-0.5f * perp);
    float G = ROOT_TWO_PI * wil;
    float bs = B * ONE_OVER_ROOT_TWO;
    float term = vpo_erf_diff(bs, fmaf(tl, ONE_OVER_ROOT_TWO, bs)) * 0.5f;
    return pr->mass * G * e * term * pr->dnf;
}

/* intersection.h:115-126 */
static float exit_inside(const prim_t* pr, v3 o, v3 d) {
    v3 w = local_dir(pr, d), p = local_pos(pr, o);
    float a = vdot(w, w);
    float b = vdot(p, w);
    float c = vdot(p, p) - EXTENT2;
    float disc = fmaf(-a, c, b * b);
    return (-b + sqrtf(disc)) * vpo_rcp(a);
}

/* Crossing of the 3-sigma bound from outside over [0, 1e20): the built-in
 * sphere contract (renderer.cpp:142-157 instance transform T R S(3 s);
 * anyhit.cuh:68-86; intersection.h:47-55 front face only), solved with the
 * stable quadratic of intersection.h:66-101 at radius^2 = 9.  Entry c/q; exit
 * the far root q/a (compute_exit_from_entry, intersection.h:142-173, is the
 * same root); optical depth of the whole chord (optical_depth over
 * [entry, exit], primitive.h:249-318) in the closed form
 *   M dnf sqrt(2 pi)/|w| e^(-|q0|^2/2) erf(sqrt(delta/2))
 *     = 2 e^(-9/2) M dnf (sqrt(a delta)/a) S(delta)
 * with S(d) = sum d^n/(2n+1)!! (vpo_chord_poly).  Chords of length <= 1e-6
 * have depth 0.  t0 < 0: no entry. */
typedef struct { float t0, t1, tau; } crossing_t;
#define CHORD_SCALE 0x1.6c0504p-6f /* 2 e^(-9/2) */
static crossing_t crossing(const prim_t* pr, v3 o, v3 d) {
    crossing_t r = {-1.0f, -1.0f, 0.0f};
    v3 w = local_dir(pr, d), p = local_pos(pr, o);
    float a = vdot(w, w);
    if (a < 1e-12f) return r;
    float c = vdot(p, p) - EXTENT2;
    if (!(c > 0.0f)) return r;
    float b = -vdot(p, w);
    if (!(b > 0.0f)) return r;
    float ai = vpo_rcp(a);
    v3 q0 = vfma(b * ai, w, p);
    float delta = EXTENT2 - vdot(q0, q0);
    if (delta < 1e-12f) return r;
    float sq = sqrtf(a * delta);
    float q = b + sq;
    float t = c * vpo_rcp(q);
    if (!(t >= 0.0f && t < INF_F)) return r;
    r.t0 = t;
    r.t1 = q * ai;
    if (r.t1 - t > 1e-6f && r.t1 < INF_F) r.tau = ((CHORD_SCALE * (pr->mass * pr->dnf)) * (sq * ai)) * vpo_chord_poly(delta);
    return r;
}

/* Chord of a ray whose origin is inside the 3-sigma bound, origin to exit,
 * from one transform: the exit of exit_inside (intersection.h:115-126) and
 * optical_depth over [0, exit] (primitive.h:249-318), written with
 * e^(x^2) erf(x) = (2/sqrt(pi)) x S(2x^2) as
 *   M dnf [e^(-9/2) (sqrt(a delta)/a) S(delta) - e^(-|p|^2/2) (b/a) S(b^2/a)]
 * with b = p.w and delta = (b^2 - a c)/a.  Exits <= 1e-6 have depth 0. */
typedef struct { float t1, tau; } inside_t;
static inside_t inside_chord(const prim_t* pr, v3 o, v3 d) {
    v3 w = local_dir(pr, d), p = local_pos(pr, o);
    float a = vdot(w, w);
    float b = vdot(p, w);
    float pp = vdot(p, p);
    float c = pp - EXTENT2;
    float disc = fmaf(-a, c, b * b);
    float sq = sqrtf(disc);
    float ai = vpo_rcp(a);
    inside_t r;
    r.t1 = (-b + sq) * ai;
    r.tau = 0.0f;
    if (r.t1 > 1e-6f && r.t1 < INF_F) {
        float fr = ((CHORD_SCALE * 0.5f) * (sq * ai)) * vpo_chord_poly(vpo_min(disc * ai, EXTENT2));
        float nr = (vpo_expf(-0.5f * pp) * (b * ai)) * vpo_chord_poly(vpo_min((b * b) * ai, EXTENT2));
        r.tau = vpo_max((pr->mass * pr->dnf) * (fr - nr), 0.0f);
    }
    return r;
}

/* Free flight from an origin inside the primitive (sampling.cuh:356-370):
 * inv_cdf (primitive.h:140-186) kept in [0, exit] when the sampled depth is
 * at most the chord's, INF otherwise. */
static float origin_flight(const prim_t* pr, v3 o, v3 d, inside_t x, float tau) {
    if (!(tau <= x.tau)) return INF_F;
    return vpo_clamp(inv_cdf(pr, o, d, tau), 0.0f, x.t1);
}

/* Free flight into an entered primitive (sampling.cuh:372-412): it ends in
 * the chord iff the sampled depth is at most the chord's; the distance is
 * inv_cdf_span from the entry (primitive.h:203-246) kept in [entry, exit]
 * (the reference clamps a result below the entry up to it and rejects one
 * beyond the exit).  INF when the flight passes the chord. */
static float entered_flight(const prim_t* pr, v3 o, v3 d, crossing_t x, float tau) {
    if (!(tau <= x.tau)) return INF_F;
    return vpo_clamp(inv_cdf_span(pr, o, d, x.t0, tau), x.t0, x.t1);
}

/* intersection.h:56-61 */
static int point_inside(const prim_t* pr, v3 x) {
    v3 l = local_pos(pr, x);
    return vdot(l, l) <= EXTENT2;
}

/* test access to the primitive functions: fn selects, args per row */
void vpo_prim_eval(int fn, const float* prim16, const float* args, long m, float* out) {
    const prim_t* p = (const prim_t*)prim16;
    for (long i = 0; i < m; ++i) {
        const float* a = args + 9 * i; /* o(3) d(3) t0 t1 tau */
        v3 o = V(a[0], a[1], a[2]), d = V(a[3], a[4], a[5]);
        switch (fn) {
            case 0: out[i] = inv_cdf(p, o, d, a[8]); break;
            case 1: out[i] = inv_cdf_span(p, o, d, a[6], a[8]); break;
            case 2: out[i] = optical_depth(p, o, d, a[6], a[7]); break;
            case 3: out[i] = exit_inside(p, o, d); break;
            case 4: out[i] = crossing(p, o, d).t0; break;
            case 5: { crossing_t x = crossing(p, o, d); out[i] = x.t0 >= 0.0f ? x.t1 : -1.0f; } break;
            case 6: out[i] = (float)point_inside(p, o); break;
            case 7: out[i] = prim_pdf(p, o); break;
            case 8: out[i] = crossing(p, o, d).tau; break;
            case 9: out[i] = entered_flight(p, o, d, crossing(p, o, d), a[8]); break;
            case 10: out[i] = inside_chord(p, o, d).tau; break;
            case 11: out[i] = origin_flight(p, o, d, inside_chord(p, o, d), a[8]); break;
            default: out[i] = NAN;
        }
    }
}

void vpo_math_eval(int fn, const float* x, const float* y, long n, float* out) {
    for (long i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = vpo_expf(x[i]); break;
            case 1: out[i] = vpo_logf(x[i]); break;
            case 2: out[i] = vpo_sinf(x[i]); break;
            case 3: out[i] = vpo_cosf(x[i]); break;
            case 4: out[i] = vpo_atan2f(x[i], y[i]); break;
            case 5: out[i] = vpo_acosf(x[i]); break;
            case 6: out[i] = vpo_erff(x[i]); break;
            case 7: out[i] = vpo_erfcf(x[i]); break;
            case 8: out[i] = vpo_erfinvf(x[i]); break;
            case 9: out[i] = vpo_erf_diff(x[i], y[i]); break;
            case 10: out[i] = sqrtf(x[i]); break;
            case 11: out[i] = x[i] / y[i]; break;
            case 12: out[i] = vpo_chord_poly(x[i]); break;
            default: out[i] = NAN;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Morton codes (host/utils/math.h:103-160) and stable sort by code           */
/* ------------------------------------------------------------------------ */
static uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
uint32_t vpo_expand_bits(uint32_t v) { return expand_bits(v); }

/* renderer.cpp:90-104: bounds of the centres, inv_extent = 1/max(ext, 1e-30) */
void vpo_morton_codes(const float* centers, long n, uint32_t* codes) {
    float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float mx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (long i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            float c = centers[3 * i + k];
            mn[k] = vpo_min(mn[k], c);
            mx[k] = vpo_max(mx[k], c);
        }
    float inv[3];
    for (int k = 0; k < 3; ++k) inv[k] = 1.0f / vpo_max(mx[k] + (-mn[k]), 1e-30f);
    for (long i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            float nrm = (centers[3 * i + k] + (-mn[k])) * inv[k];
            q[k] = (uint32_t)(vpo_clamp(nrm, 0.0f, 1.0f) * 1023.0f);
        }
        codes[i] = (expand_bits(q[2]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[0]);
    }
}

static const uint32_t* g_sort_codes;
static int cmp_code_idx(const void* a, const void* b) {
    uint32_t ia = *(const uint32_t*)a, ib = *(const uint32_t*)b;
    uint32_t ca = g_sort_codes[ia], cb = g_sort_codes[ib];
    if (ca != cb) return ca < cb ? -1 : 1;
    return ia < ib ? -1 : (ia > ib ? 1 : 0);
}
/* order[k] = original index of the k-th primitive by (code, index) */
void vpo_sort_order(const uint32_t* codes, long n, uint32_t* order) {
    for (long i = 0; i < n; ++i) order[i] = (uint32_t)i;
    g_sort_codes = codes;
    qsort(order, (size_t)n, sizeof(uint32_t), cmp_code_idx);
}

/* ------------------------------------------------------------------------ */
/* RNG (random.cuh:16-53) and the per-primitive counter draw                  */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t state, inc; } pcg_t;
static inline uint32_t pcg_next(pcg_t* r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ull + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
static inline uint64_t hash64(uint64_t x) {
    x = x * 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x = x * 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}
static inline pcg_t pcg_init(uint64_t seed, uint64_t seq) {
    pcg_t r;
    r.state = 0;
    r.inc = (seq << 1u) | 1u;
    pcg_next(&r);
    r.state += seed;
    pcg_next(&r);
    return r;
}
static inline float u01(uint32_t v) { return (float)v / 4294967296.0f; }
static inline float next_u(pcg_t* r) { return u01(pcg_next(r)); }
static inline uint64_t path_key(uint64_t seed, uint64_t seq) { return hash64(seq ^ (seed * 0x9e3779b97f4a7c15ull) ^ 0x243f6a8885a308d3ull); }
/* role 0 = primitive containing the ray origin, 1 = primitive entered */
static inline float prim_u(uint64_t key, uint32_t bounce, uint32_t role, uint32_t prim) {
    uint64_t x = key + 0x9e3779b97f4a7c15ull * (((uint64_t)((bounce << 1) | role) << 32) | prim);
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return u01((uint32_t)(x >> 32));
}
/* sampling.cuh:310-313 */
static inline float ff_tau(float xi) { return -vpo_logf(vpo_max(1.0f - xi, 1e-10f)); }

void vpo_pcg_stream(uint64_t seed, uint64_t seq, long n, uint32_t* out) {
    pcg_t r = pcg_init(seed, seq);
    for (long i = 0; i < n; ++i) out[i] = pcg_next(&r);
}
void vpo_pcg_raw(uint64_t state, uint64_t inc, long n, uint32_t* out) {
    pcg_t r = {state, inc};
    for (long i = 0; i < n; ++i) out[i] = pcg_next(&r);
}
uint64_t vpo_hash64(uint64_t x) { return hash64(x); }

/* ------------------------------------------------------------------------ */
/* Optional CPU bounding-volume hierarchy for the estimator (use_bvh).  Every  */
/* query result is consumed order-independently (minimum, fixed-point sum) or */
/* after sorting by primitive index, so renders are bit-identical to the      */
/* brute-force loops; it only makes the CPU baseline fair at large N.         */
/* ------------------------------------------------------------------------ */
typedef struct { float lo[3], hi[3]; int32_t left, right; } cnode_t; /* leaf: left = -1 - prim */
typedef struct { cnode_t* nodes; long nn; } cbvh_t;

/* Conservative box of the 3-sigma ellipsoid: half extent 3 sqrt(sum_j (R_ij s_j)^2)
 * along world axis i, widened by a relative 2^-8 and an absolute 2^-18 (|c| + h). */
static void prim_box(const prim_t* p, float lo[3], float hi[3]) {
    double qw = p->qw, qx = -(double)p->qx, qy = -(double)p->qy, qz = -(double)p->qz;
    double R[3][3] = {
        {1.0 - 2.0 * (qy * qy + qz * qz), 2.0 * (qx * qy - qw * qz), 2.0 * (qx * qz + qw * qy)},
        {2.0 * (qx * qy + qw * qz), 1.0 - 2.0 * (qx * qx + qz * qz), 2.0 * (qy * qz - qw * qx)},
        {2.0 * (qx * qz - qw * qy), 2.0 * (qy * qz + qw * qx), 1.0 - 2.0 * (qx * qx + qy * qy)},
    };
    double sv[3] = {1.0 / (double)p->rsx, 1.0 / (double)p->rsy, 1.0 / (double)p->rsz};
    double cv[3] = {p->cx, p->cy, p->cz};
    for (int a = 0; a < 3; ++a) {
        double h2 = 0.0;
        for (int j = 0; j < 3; ++j) h2 += (R[a][j] * sv[j]) * (R[a][j] * sv[j]);
        double h = 3.0 * sqrt(h2);
        h = h * (1.0 + 0x1p-8) + 0x1p-18 * (fabs(cv[a]) + h);
        lo[a] = (float)(cv[a] - h);
        hi[a] = (float)(cv[a] + h);
        if ((double)lo[a] > cv[a] - h) lo[a] = nextafterf(lo[a], -3.0e38f);
        if ((double)hi[a] < cv[a] + h) hi[a] = nextafterf(hi[a], 3.0e38f);
    }
}

static int32_t cbvh_build_rec(cbvh_t* B, const uint32_t* perm, long lo, long hi, const float* boxes) {
    int32_t id = (int32_t)B->nn++;
    cnode_t* nd = &B->nodes[id];
    if (hi - lo == 1) {
        uint32_t p = perm[lo];
        for (int a = 0; a < 3; ++a) { nd->lo[a] = boxes[6 * p + a]; nd->hi[a] = boxes[6 * p + 3 + a]; }
        nd->left = -1 - (int32_t)p;
        nd->right = -1;
        return id;
    }
    long mid = lo + (hi - lo) / 2;
    int32_t l = cbvh_build_rec(B, perm, lo, mid, boxes);
    int32_t r = cbvh_build_rec(B, perm, mid, hi, boxes);
    nd = &B->nodes[id];
    for (int a = 0; a < 3; ++a) {
        nd->lo[a] = vpo_min(B->nodes[l].lo[a], B->nodes[r].lo[a]);
        nd->hi[a] = vpo_max(B->nodes[l].hi[a], B->nodes[r].hi[a]);
    }
    nd->left = l;
    nd->right = r;
    return id;
}

/* Primitives in Morton order of their centres, split in halves. */
static void cbvh_build(cbvh_t* B, const prim_t* prims, long n) {
    B->nodes = NULL;
    B->nn = 0;
    if (n <= 0) return;
    float* boxes = (float*)malloc(sizeof(float) * 6 * (size_t)n);
    float* centers = (float*)malloc(sizeof(float) * 3 * (size_t)n);
    uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    for (long i = 0; i < n; ++i) {
        prim_box(prims + i, boxes + 6 * i, boxes + 6 * i + 3);
        centers[3 * i] = prims[i].cx;
        centers[3 * i + 1] = prims[i].cy;
        centers[3 * i + 2] = prims[i].cz;
    }
    vpo_morton_codes(centers, n, codes);
    vpo_sort_order(codes, n, perm);
    B->nodes = (cnode_t*)malloc(sizeof(cnode_t) * (size_t)(2 * n - 1));
    cbvh_build_rec(B, perm, 0, n, boxes);
    free(boxes);
    free(centers);
    free(codes);
    free(perm);
}

#define CBVH_STACK 128
typedef struct { const cbvh_t* B; v3 o, inv, x; int sp; int32_t st[CBVH_STACK]; } cbvh_it;

static inline float safe_rcp(float d) { return 1.0f / (fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d); }

static void cbvh_ray_begin(cbvh_it* it, const cbvh_t* B, v3 o, v3 d) {
    it->B = B;
    it->o = o;
    it->inv = V(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
    it->sp = 0;
    if (B->nn) it->st[it->sp++] = 0;
}
/* Next leaf whose box meets the ray over [0, tmax], or -1. */
static long cbvh_ray_next(cbvh_it* it, float tmax) {
    while (it->sp > 0) {
        const cnode_t* nd = &it->B->nodes[it->st[--it->sp]];
        float tx0 = (nd->lo[0] - it->o.x) * it->inv.x, tx1 = (nd->hi[0] - it->o.x) * it->inv.x;
        float ty0 = (nd->lo[1] - it->o.y) * it->inv.y, ty1 = (nd->hi[1] - it->o.y) * it->inv.y;
        float tz0 = (nd->lo[2] - it->o.z) * it->inv.z, tz1 = (nd->hi[2] - it->o.z) * it->inv.z;
        float tn = vpo_max(vpo_max(vpo_min(tx0, tx1), vpo_min(ty0, ty1)), vpo_max(vpo_min(tz0, tz1), 0.0f));
        float tf = vpo_min(vpo_min(vpo_max(tx0, tx1), vpo_max(ty0, ty1)), vpo_min(vpo_max(tz0, tz1), tmax));
        if (!(tn <= tf)) continue;
        if (nd->left < 0) return (long)(-1 - nd->left);
        it->st[it->sp++] = nd->right;
        it->st[it->sp++] = nd->left;
    }
    return -1;
}
static void cbvh_point_begin(cbvh_it* it, const cbvh_t* B, v3 x) {
    it->B = B;
    it->x = x;
    it->sp = 0;
    if (B->nn) it->st[it->sp++] = 0;
}
/* Next leaf whose box contains x, or -1. */
static long cbvh_point_next(cbvh_it* it) {
    while (it->sp > 0) {
        const cnode_t* nd = &it->B->nodes[it->st[--it->sp]];
        v3 x = it->x;
        if (!(x.x >= nd->lo[0] && x.x <= nd->hi[0] && x.y >= nd->lo[1] && x.y <= nd->hi[1] && x.z >= nd->lo[2] &&
              x.z <= nd->hi[2]))
            continue;
        if (nd->left < 0) return (long)(-1 - nd->left);
        it->st[it->sp++] = nd->right;
        it->st[it->sp++] = nd->left;
    }
    return -1;
}

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}
/* Sorted indices of the primitives whose box contains x; returns the count. */
static long cbvh_point_sorted(const cbvh_t* B, v3 x, uint32_t* out) {
    cbvh_it it;
    cbvh_point_begin(&it, B, x);
    long c = 0, p;
    while ((p = cbvh_point_next(&it)) >= 0) out[c++] = (uint32_t)p;
    qsort(out, (size_t)c, sizeof(uint32_t), cmp_u32);
    return c;
}

/* ------------------------------------------------------------------------ */
/* shadow-ray transmittance (sampling.cuh:548-583, anyhit.cuh:53-66)          */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t n_sh_active, n_sh_hit, n_arg_active, n_arg_hit, n_arg_accept;
    uint64_t n_rb_active, n_rb_hit, n_alb, n_scatter, n_sample, n_shadow_rays, n_trace;
    uint64_t active_overflow, max_hits, max_overlap;
} counters_t;

#define TAU_LIMIT_FIXED 96757023244288ull /* 88 * 2^40 */

static inline int tau_add(uint64_t* acc, float od) {
    float c = od > 0.0f ? od : 0.0f;
    if (c >= 88.0f) return 1;
    *acc += (uint64_t)(c * 1099511627776.0f);
    return *acc >= TAU_LIMIT_FIXED;
}

/* Fixed-point optical depth along a ray (sampling.cuh:548-583): origin set
 * (optical_depth from the origin to each member's exit) plus every entered
 * chord; the sum stops once it reaches `limit` (at most the tau = 88
 * saturation), which is order independent because every term is >= 0. */
static uint64_t shadow_depth(const prim_t* prims, long n, v3 o, v3 d, const uint32_t* set, int nset, counters_t* k,
                             const cbvh_t* B, uint64_t limit) {
    uint64_t acc = 0;
    if (acc >= limit) return acc;
    for (int s = 0; s < nset; ++s) {
        float tc = inside_chord(prims + set[s], o, d).tau;
        if (k) k->n_sh_active++;
        if (tc > 0.0f) {
            if (tau_add(&acc, tc)) return TAU_LIMIT_FIXED;
            if (acc >= limit) return acc;
        }
    }
    /* entered primitives: every index (brute force) or the BVH's leaves along the ray */
    cbvh_it it;
    if (B) cbvh_ray_begin(&it, B, o, d);
    for (long j = B ? cbvh_ray_next(&it, INF_F) : 0; B ? j >= 0 : j < n; j = B ? cbvh_ray_next(&it, INF_F) : j + 1) {
        crossing_t x = crossing(prims + j, o, d);
        if (x.t0 < 0.0f) continue;
        if (k) k->n_sh_hit++;
        if (tau_add(&acc, x.tau)) return TAU_LIMIT_FIXED;
        if (acc >= limit) return acc;
    }
    return acc;
}

static inline float depth_to_T(uint64_t acc) {
    if (acc >= TAU_LIMIT_FIXED) return 0.0f;
    float tau = (float)acc * (1.0f / 1099511627776.0f);
    return tau >= 88.0f ? 0.0f : vpo_expf(-tau);
}

static float transmittance(const prim_t* prims, long n, v3 o, v3 d, const uint32_t* set, int nset, counters_t* k,
                           const cbvh_t* B) {
    return depth_to_T(shadow_depth(prims, n, o, d, set, nset, k, B, TAU_LIMIT_FIXED));
}

/* Shadow-ray roulette (an option of this build; the reference walks every
 * shadow ray to the tau = 88 saturation).  A shadow ray whose contribution
 * is w * T(tau) plays a Russian roulette at each depth theta_j = tau_c +
 * j ln 4 it reaches (tau_c = log(w / eps), where w e^(-tau) falls below
 * eps), surviving each with probability 1/4; a ray that reached L levels and
 * survived them counts 4^L T, else 0.  The expectation is T, so the
 * estimator stays unbiased.  Whether tau passes a depth does not depend on
 * the order chords are summed in, and the survival draws are counter hashes
 * of (path, bounce, ray, level), so the walk can stop at the first level the
 * ray fails.  Depths are 24.40 fixed point; thresholds and the step are
 * multiples of 2^24. */
#define RR_GRAIN 0xffffffull
#define RR_STEP (1524246769571ull & ~RR_GRAIN) /* ln 4 * 2^40 */
#define RR_NO_FAIL 127u
static inline uint64_t rr_threshold(float w, float eps) {
    if (!(eps > 0.0f)) return TAU_LIMIT_FIXED;
    if (!(w > eps)) return 0;
    float tc = vpo_logf(w * vpo_rcp(eps));
    if (!(tc > 0.0f)) return 0;
    if (tc >= 88.0f) return TAU_LIMIT_FIXED;
    return ((uint64_t)(tc * 1099511627776.0f)) & ~RR_GRAIN;
}
static inline float rr_uniform(uint64_t key, uint32_t bounce, uint32_t ray, uint32_t level) {
    const uint64_t c = ((uint64_t)bounce << 16) | ((uint64_t)ray << 8) | level;
    uint64_t x = (key ^ 0xd1b54a32d192ed03ull) + 0x9e3779b97f4a7c15ull * c;
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return u01((uint32_t)(x >> 32));
}
/* first failing level of a ray's roulette (RR_NO_FAIL: none below tau = 88) */
static inline uint32_t rr_fail(uint64_t thr, uint64_t key, uint32_t bounce, uint32_t ray) {
    for (uint32_t j = 0; thr + j * RR_STEP < TAU_LIMIT_FIXED; ++j)
        if (!(rr_uniform(key, bounce, ray, j) < 0.25f)) return j;
    return RR_NO_FAIL;
}
static inline uint64_t rr_limit(uint64_t thr, uint32_t fail) {
    return fail == RR_NO_FAIL ? TAU_LIMIT_FIXED : thr + fail * RR_STEP;
}
/* T = depth_to_T(acc) under the roulette: 0 past the failing level, else
 * T * 4^(levels reached) in two exact power-of-two scalings */
static inline float rr_apply(uint64_t acc, uint64_t thr, uint32_t fail, float T) {
    if (acc >= rr_limit(thr, fail)) return 0.0f;
    if (acc < thr) return T;
    const int e = 2 * (int)((acc - thr) / RR_STEP + 1u);
    const int e1 = e >> 1;
    return (T * vpo_exp2i(e1)) * vpo_exp2i(e - e1);
}
static float shadow_T(const prim_t* prims, long n, v3 o, v3 d, const uint32_t* set, int nset, counters_t* k,
                      const cbvh_t* B, float w, float eps, uint64_t key, uint32_t bounce, uint32_t ray) {
    const uint64_t thr = rr_threshold(w, eps);
    const uint32_t fail = rr_fail(thr, key, bounce, ray);
    const uint64_t acc = shadow_depth(prims, n, o, d, set, nset, k, B, rr_limit(thr, fail));
    return rr_apply(acc, thr, fail, depth_to_T(acc));
}

/* Test access: the roulette's estimates of T for rays of weight w whose
 * depth is tau (walked in full), one per key, and their first failing
 * levels; the return value is the fixed-point threshold tau_c. */
uint64_t vpo_shadow_rr_T(float w, float eps, const uint64_t* keys, long m, uint32_t bounce, uint32_t ray, double tau,
                         float* T_out, uint32_t* fail_out) {
    const uint64_t thr = rr_threshold(w, eps);
    const double a = tau * 1099511627776.0;
    const uint64_t acc = a >= (double)TAU_LIMIT_FIXED ? TAU_LIMIT_FIXED : (uint64_t)a;
    const float T = depth_to_T(acc);
    for (long i = 0; i < m; ++i) {
        const uint32_t fail = rr_fail(thr, keys[i], bounce, ray);
        fail_out[i] = fail;
        T_out[i] = rr_apply(acc, thr, fail, T);
    }
    return thr;
}

/* Rays with origin-inside set = every primitive whose bound contains the
 * origin (the bounce-0 scan, sampling.cuh:322-341). */
void vpo_transmittance_rays(const float* prims16, long n, const float* rays, long m, float* T) {
    const prim_t* prims = (const prim_t*)prims16;
    uint32_t* set = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    for (long r = 0; r < m; ++r) {
        v3 o = V(rays[6 * r], rays[6 * r + 1], rays[6 * r + 2]);
        v3 d = V(rays[6 * r + 3], rays[6 * r + 4], rays[6 * r + 5]);
        int ns = 0;
        for (long i = 0; i < n; ++i)
            if (point_inside(prims + i, o)) set[ns++] = (uint32_t)i;
        T[r] = transmittance(prims, n, o, d, set, ns, NULL, NULL);
    }
    free(set);
}

/* Entry list of each ray sorted by (t, index): counts[r], then t/idx rows of
 * width cap.  Returns the largest count seen. */
long vpo_overlaps(const float* prims16, long n, const float* rays, long m, long cap, uint32_t* counts, float* tout, uint32_t* iout) {
    const prim_t* prims = (const prim_t*)prims16;
    long maxc = 0;
    for (long r = 0; r < m; ++r) {
        v3 o = V(rays[6 * r], rays[6 * r + 1], rays[6 * r + 2]);
        v3 d = V(rays[6 * r + 3], rays[6 * r + 4], rays[6 * r + 5]);
        long c = 0;
        float* tr = tout + r * cap;
        uint32_t* ir = iout + r * cap;
        for (long j = 0; j < n; ++j) {
            float t = crossing(prims + j, o, d).t0;
            if (t < 0.0f) continue;
            if (c < cap) {
                /* insertion in (t, index) order */
                long k = c;
                while (k > 0 && (tr[k - 1] > t)) { tr[k] = tr[k - 1]; ir[k] = ir[k - 1]; --k; }
                tr[k] = t;
                ir[k] = (uint32_t)j;
            }
            ++c;
        }
        counts[r] = (uint32_t)c;
        if (c > maxc) maxc = c;
    }
    return maxc;
}

/* ------------------------------------------------------------------------ */
/* environment                                                               */
/* ------------------------------------------------------------------------ */
typedef struct {
    const float* tex;
    const float* prob;
    const uint32_t* alias;
    const float* joint;
    uint32_t w, h;
    float total;
} env_t;

/* joint = lum709 * sin(theta_texel_centre), total = sum of row sums
 * (host/params/environment_map.h:45-101). */
static float env_joint(const float* rgba, int w, int h, float* joint) {
    float total = 0.0f;
    for (int v = 0; v < h; ++v) {
        float theta = ((float)v + 0.5f) / (float)h * VPO_PI;
        float st = vpo_sinf(theta);
        float row = 0.0f;
        for (int u = 0; u < w; ++u) {
            const float* t = rgba + 4 * ((size_t)v * w + u);
            float lum = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
            float wt = lum * st;
            joint[(size_t)v * w + u] = wt;
            row += wt;
        }
        total += row;
    }
    return total;
}

/* Vose alias table over p_i = joint_i / sum(joint) in double precision,
 * stacks popped from the back. */
void vpo_env_tables(const float* rgba, int w, int h, float* joint, float* prob, uint32_t* alias, float* total_out) {
    size_t n = (size_t)w * h;
    float total = env_joint(rgba, w, h, joint);
    *total_out = total;
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += (double)joint[i];
    double* q = (double*)malloc(sizeof(double) * n);
    uint32_t* small = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* large = (uint32_t*)malloc(sizeof(uint32_t) * n);
    size_t ns = 0, nl = 0;
    for (size_t i = 0; i < n; ++i) {
        q[i] = sum > 0.0 ? (double)joint[i] / sum * (double)n : 1.0;
        if (q[i] < 1.0) small[ns++] = (uint32_t)i;
        else large[nl++] = (uint32_t)i;
    }
    while (ns > 0 && nl > 0) {
        uint32_t s = small[--ns];
        uint32_t l = large[--nl];
        prob[s] = (float)q[s];
        alias[s] = l;
        q[l] = (q[l] + q[s]) - 1.0;
        if (q[l] < 1.0) small[ns++] = l;
        else large[nl++] = l;
    }
    while (nl > 0) { uint32_t l = large[--nl]; prob[l] = 1.0f; alias[l] = l; }
    while (ns > 0) { uint32_t s = small[--ns]; prob[s] = 1.0f; alias[s] = s; }
    free(q);
    free(small);
    free(large);
}

/* The reference's marginal/conditional CDFs, restated to pin the texel
 * distribution (host/params/environment_map.h:45-101). */
void vpo_env_ref_cdf(const float* rgba, int w, int h, float* marginal, float* conditional, float* joint, float* total_out) {
    for (int v = 0; v < h; ++v) {
        float theta = ((float)v + 0.5f) / (float)h * VPO_PI;
        float st = vpo_sinf(theta);
        float row = 0.0f;
        for (int u = 0; u < w; ++u) {
            const float* t = rgba + 4 * ((size_t)v * w + u);
            float lum = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
            float wt = lum * st;
            joint[(size_t)v * w + u] = wt;
            row += wt;
            conditional[(size_t)v * w + u] = row;
        }
        if (row > 0.0f) {
            float inv = 1.0f / row;
            for (int u = 0; u < w; ++u) conditional[(size_t)v * w + u] *= inv;
        } else {
            for (int u = 0; u < w; ++u) conditional[(size_t)v * w + u] = ((float)u + 1.0f) / (float)w;
        }
        marginal[v] = row;
    }
    float total = 0.0f;
    for (int v = 0; v < h; ++v) { total += marginal[v]; marginal[v] = total; }
    *total_out = total;
    if (total > 0.0f) {
        float inv = 1.0f / total;
        for (int v = 0; v < h; ++v) marginal[v] *= inv;
    } else {
        for (int v = 0; v < h; ++v) marginal[v] = ((float)v + 1.0f) / (float)h;
    }
}

static inline float lerp1(float a, float b, float t) { return fmaf(t, b - a, a); }

/* environment_map.h:36-61 with software bilinear (wrap u, clamp v) */
static v3 env_lookup(const env_t* e, v3 d) {
    float theta = vpo_atan2f(d.z, d.x);
    float phi = vpo_acosf(vpo_clamp(d.y, -1.0f, 1.0f));
    float u = fmaf(theta, 1.0f / (2.0f * VPO_PI), 0.5f);
    float v = phi * (1.0f / VPO_PI);
    float x = u * (float)e->w - 0.5f;
    float y = v * (float)e->h - 0.5f;
    float x0 = floorf(x), y0 = floorf(y);
    float fx = x - x0, fy = y - y0;
    int W = (int)e->w, H = (int)e->h;
    int i0 = (int)x0 % W;
    if (i0 < 0) i0 += W;
    int i1 = i0 + 1;
    if (i1 >= W) i1 -= W;
    int j0 = (int)y0, j1 = j0 + 1;
    j0 = j0 < 0 ? 0 : (j0 >= H ? H - 1 : j0);
    j1 = j1 < 0 ? 0 : (j1 >= H ? H - 1 : j1);
    const float* t00 = e->tex + 4 * ((size_t)j0 * W + i0);
    const float* t10 = e->tex + 4 * ((size_t)j0 * W + i1);
    const float* t01 = e->tex + 4 * ((size_t)j1 * W + i0);
    const float* t11 = e->tex + 4 * ((size_t)j1 * W + i1);
    float r[3];
    for (int c = 0; c < 3; ++c) {
        float a = lerp1(t00[c], t10[c], fx);
        float b = lerp1(t01[c], t11[c], fx);
        r[c] = lerp1(a, b, fy);
    }
    return V(r[0], r[1], r[2]);
}

static int env_table(const env_t* e) { return e->prob != NULL && e->total > 0.0f && e->w != 0 && e->h != 0; }

/* sampling.cuh:218-246 */
static float env_pdf(const env_t* e, v3 wo) {
    if (!env_table(e)) return ONE_OVER_FOUR_PI;
    int W = (int)e->w, H = (int)e->h;
    float theta = vpo_atan2f(wo.z, wo.x);
    float polar = vpo_acosf(vpo_clamp(wo.y, -1.0f, 1.0f));
    float un = fmaf(theta, 1.0f / (2.0f * VPO_PI), 0.5f);
    float vn = polar * (1.0f / VPO_PI);
    int u = (int)(un * (float)W);
    int v = (int)(vn * (float)H);
    u = u < 0 ? 0 : (u >= W ? W - 1 : u);
    v = v < 0 ? 0 : (v >= H ? H - 1 : v);
    float joint = e->joint[v * W + u];
    float sp = vpo_sinf(polar);
    if (joint <= 0.0f || sp <= 0.0f) return 0.0f;
    float p_uv = (joint * vpo_rcp(e->total)) * (float)(W * H);
    float jac = 2.0f * VPO_PI * VPO_PI * sp;
    return p_uv * vpo_rcp(jac);
}

/* sampling.cuh:173-216, texel through the alias table */
static v3 env_sample(const env_t* e, pcg_t* rng, float* pdf) {
    float ux = next_u(rng), uy = next_u(rng);
    if (!env_table(e)) {
        float z = fmaf(-2.0f, ux, 1.0f);
        float r = sqrtf(vpo_max(0.0f, fmaf(-z, z, 1.0f)));
        float sp, cp;
        vpo_sincosf(2.0f * VPO_PI * uy, &sp, &cp);
        *pdf = ONE_OVER_FOUR_PI;
        return V(r * cp, r * sp, z);
    }
    uint32_t n = e->w * e->h;
    uint32_t col = (uint32_t)(ux * (float)n);
    if (col >= n) col = n - 1;
    uint32_t texel = uy < e->prob[col] ? col : e->alias[col];
    uint32_t v = texel / e->w;
    uint32_t u = texel - v * e->w;
    float jx = next_u(rng), jy = next_u(rng);
    float un = ((float)u + jx) * vpo_rcp((float)e->w);
    float vn = ((float)v + jy) * vpo_rcp((float)e->h);
    float theta = fmaf(2.0f * VPO_PI, un, -VPO_PI);
    float polar = vn * VPO_PI;
    float sp, cpol, st, ct;
    vpo_sincosf(polar, &sp, &cpol);
    vpo_sincosf(theta, &st, &ct);
    v3 wo = V(sp * ct, cpol, sp * st);
    float joint = e->joint[texel];
    float p_uv = (joint * vpo_rcp(e->total)) * (float)(e->w * e->h);
    float jac = 2.0f * VPO_PI * VPO_PI * sp;
    *pdf = sp > 0.0f ? p_uv * vpo_rcp(jac) : 0.0f;
    return wo;
}

/* ------------------------------------------------------------------------ */
/* phase (sampling.cuh:74-134; host fold renderer.cpp:224-235)               */
/* ------------------------------------------------------------------------ */
typedef struct { int iso; float g, geff, opg2, omg2, ni2g, coeff; } hg_t;
static hg_t hg_make(float g) {
    hg_t h;
    float ag = g < 0.0f ? -g : g;
    h.iso = ag < 1e-3f;
    h.g = g;
    h.geff = -g;
    float g2 = g * g;
    h.opg2 = 1.0f + g2;
    h.omg2 = 1.0f - g2;
    h.ni2g = h.iso ? 0.0f : (-0.5f / h.geff);
    h.coeff = ONE_OVER_FOUR_PI * h.omg2;
    return h;
}
static float hg_eval_cos(const hg_t* h, float c) {
    if (h->iso) return ONE_OVER_FOUR_PI;
    float den = fmaf(-2.0f * h->g, c, h->opg2);
    float is = vpo_rsqrt(den);
    return h->coeff * (is * is * is);
}
static void onb(v3 n, v3* tx, v3* ty) {
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    *tx = V(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    *ty = V(b, sign + n.y * n.y * a, -n.y);
}
static v3 hg_sample(const hg_t* h, v3 wi, pcg_t* rng, float* pdf) {
    float ux = next_u(rng), uy = next_u(rng), c;
    if (h->iso) {
        c = fmaf(-2.0f, ux, 1.0f);
    } else {
        float inner = fmaf(-2.0f, ux, 1.0f);
        float den = fmaf(h->geff, inner, 1.0f);
        float sq = h->omg2 * vpo_rcp(den);
        c = h->ni2g * fmaf(-sq, sq, h->opg2);
    }
    float s = sqrtf(vpo_max(0.0f, 1.0f - c * c));
    float sp, cp;
    vpo_sincosf(2.0f * VPO_PI * uy, &sp, &cp);
    v3 ex, ey;
    onb(wi, &ex, &ey);
    v3 wo = vadd(vadd(vscale(ex, s * cp), vscale(ey, s * sp)), vscale(wi, c));
    *pdf = hg_eval_cos(h, c);
    return wo;
}

/* ------------------------------------------------------------------------ */
/* camera (device/params/camera.h:33-57; host/params/camera.h:40-111)         */
/* POD: eye[3] p00[3] du[3] dv[3] view_dir[3] ortho                            */
/* ------------------------------------------------------------------------ */
static v3 vld(const float* p) { return V(p[0], p[1], p[2]); }
static void vst(float* p, v3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

void vpo_camera_perspective(int w, int h, const float* from, const float* at, const float* up, float vfov, float* pod) {
    float theta = vfov * (VPO_PI / 180.0f);
    float hh = tanf(0.5f * theta);
    v3 lf = vld(from), la = vld(at), vu = vld(up);
    v3 W = vnormalize(vsub(lf, la));
    v3 U = vnormalize(vcross(vu, W));
    v3 Vv = vcross(W, U);
    v3 diff = vsub(lf, la);
    float focal = sqrtf(vdot(diff, diff));
    float aspect = (float)w * vpo_rcp((float)h);
    float vh = 2.0f * hh * focal;
    float vw = vh * aspect;
    v3 vpu = vscale(U, vw);
    v3 vpv = vscale(V(-Vv.x, -Vv.y, -Vv.z), vh);
    v3 pdu = vscale(vpu, vpo_rcp((float)w));
    v3 pdv = vscale(vpv, vpo_rcp((float)h));
    /* midpoint(a, b) = lerp(a, b, 0.5) = fma(0.5, b - a, a) (math.h:350-366) */
    v3 mid_uv = vfma(0.5f, vsub(vpv, vpu), vpu);
    v3 ul = vsub(vsub(lf, vscale(W, focal)), mid_uv);
    v3 mid_d = vfma(0.5f, vsub(pdv, pdu), pdu);
    v3 p00 = vadd(ul, mid_d);
    memset(pod, 0, 16 * sizeof(float));
    vst(pod, lf);
    vst(pod + 3, vsub(p00, lf));
    vst(pod + 6, pdu);
    vst(pod + 9, pdv);
    vst(pod + 12, V(0.0f, 0.0f, 1.0f));
    pod[15] = 0.0f;
}

void vpo_camera_orthographic(int w, int h, const float* from, const float* at, const float* up, float ortho_h, float* pod) {
    v3 lf = vld(from), la = vld(at), vu = vld(up);
    v3 vd = vnormalize(vsub(la, lf));
    v3 right = vnormalize(vcross(vd, vu));
    v3 upv = vcross(right, vd);
    float aspect = (float)w * vpo_rcp((float)h);
    float vh = ortho_h;
    float vw = vh * aspect;
    v3 vpu = vscale(right, vw);
    v3 vpv = vscale(upv, vh);
    v3 pdu = vscale(vpu, vpo_rcp((float)w));
    v3 pdv = vscale(vpv, vpo_rcp((float)h));
    v3 ul = vsub(lf, vfma(0.5f, vsub(vpv, vpu), vpu));
    v3 p00 = vadd(ul, vfma(0.5f, vsub(pdv, pdu), pdu));
    memset(pod, 0, 16 * sizeof(float));
    vst(pod, lf);
    vst(pod + 3, vsub(p00, lf));
    vst(pod + 6, pdu);
    vst(pod + 9, pdv);
    vst(pod + 12, vd);
    pod[15] = 1.0f;
}

static void cam_ray(const float* pod, uint32_t px, uint32_t py, float jx, float jy, v3* o, v3* d) {
    float x = (float)px + jx, y = (float)py + jy;
    v3 r = vfma(x, vld(pod + 6), vfma(y, vld(pod + 9), vld(pod + 3)));
    if (pod[15] != 0.0f) {
        *o = vadd(vld(pod), r);
        *d = vnormalize(vld(pod + 12));
    } else {
        *o = vld(pod);
        *d = vnormalize(r);
    }
}

/* ------------------------------------------------------------------------ */
/* the estimator (raygen.cuh:21-243, path_stages.cuh:25-226,                  */
/* sampling.cuh:322-583)                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t width, height, spp, seed;
    uint32_t max_depth, rr_depth;
    float rr_max_survival, firefly_clamp, filter_stddev, phase_g;
    uint32_t use_ris, ris_candidates, active_cap, measure;
    uint32_t x0, y0, x1, y1;         /* pixel crop [x0,x1) x [y0,y1) */
    uint32_t sample_begin, sample_end; /* samples [begin, end) of spp */
    uint32_t stride;                   /* render every stride-th row and column of the crop (0 = 1) */
    uint32_t use_bvh;                  /* queries through a CPU BVH instead of loops over all primitives */
    float shadow_rr;                   /* NEE shadow-ray roulette threshold (0 = off) */
    float adaptive_threshold;          /* adaptive sampling (0 = off); needs samples [0, spp) */
    uint32_t adaptive_min;             /* samples before a pixel may stop */
    uint32_t adaptive_batch;           /* samples per batch (0 = 16) */
} vpo_params;

typedef struct {
    const prim_t* prims;
    long n;
    env_t env;
    const float* cam;
    const vpo_params* p;
    const cbvh_t* bvh; /* NULL: brute force */
    hg_t hg;
    float* out_mean;
    float* out_samples;
    uint32_t* out_counts;
    uint32_t next_row;
    pthread_mutex_t mu;
    counters_t total;
} job_t;

/* set insert with the reference's CompactSet contract (drop at capacity) */
static inline void set_push(uint32_t* s, int* ns, int cap, uint32_t v, counters_t* k) {
    if (*ns < cap) s[(*ns)++] = v;
    else k->active_overflow++;
}

/* First scatter distance along (o, d): minimum over every primitive the ray
 * enters of inv_cdf_span from its entry, clamped up to the entry and
 * accepted if inside the exit, and over the origin set of inv_cdf, accepted
 * if inside the exit (sampling.cuh:310-313,350-415).  INF = escape. */
static float argmin_scatter(const prim_t* prims, long n, v3 o, v3 d, const uint32_t* act, int na, uint64_t key,
                            uint32_t bounce, uint32_t* hits_out, counters_t* k, const cbvh_t* B, int prune) {
    float ts_min = INF_F;
    uint32_t hits = 0;
    /* every index (brute force) or the BVH's leaves along the ray; with
     * pruning, boxes beyond the running minimum are skipped */
    cbvh_it it;
    if (B) cbvh_ray_begin(&it, B, o, d);
#define ARG_NEXT(j) (B ? cbvh_ray_next(&it, prune ? ts_min : INF_F) : (j) + 1)
    for (long j = B ? ARG_NEXT(-1) : 0; B ? j >= 0 : j < n; j = ARG_NEXT(j)) {
        const prim_t* pr = prims + j;
        crossing_t x = crossing(pr, o, d);
        if (x.t0 < 0.0f) continue;
        ++hits;
        k->n_arg_hit++;
        float tau = ff_tau(prim_u(key, bounce, 1u, (uint32_t)j));
        if (tau <= x.tau) {
            k->n_arg_accept++;
            float ts = entered_flight(pr, o, d, x, tau);
            if (ts < ts_min) ts_min = ts;
        }
    }
#undef ARG_NEXT
    for (int s = 0; s < na; ++s) {
        const prim_t* pr = prims + act[s];
        k->n_arg_active++;
        inside_t x = inside_chord(pr, o, d);
        float tau = ff_tau(prim_u(key, bounce, 0u, act[s]));
        if (tau <= x.tau) ts_min = vpo_min(ts_min, origin_flight(pr, o, d, x, tau));
    }
    *hits_out = hits;
    return ts_min;
}

/* First scatter distance of m independent paths (path keys keys[i], bounce 0)
 * along one ray; the origin set is every primitive containing o.  A test
 * entry point for the distribution P[t* > t] = exp(-tau(0, t)). */
void vpo_first_scatter(const float* prims16, long n, const float* ray6, const uint64_t* keys, long m, float* out) {
    const prim_t* prims = (const prim_t*)prims16;
    v3 o = V(ray6[0], ray6[1], ray6[2]), d = V(ray6[3], ray6[4], ray6[5]);
    uint32_t* act = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    int na = 0;
    for (long i = 0; i < n; ++i)
        if (point_inside(prims + i, o)) act[na++] = (uint32_t)i;
    counters_t k;
    memset(&k, 0, sizeof(k));
    for (long i = 0; i < m; ++i) {
        uint32_t hits;
        out[i] = argmin_scatter(prims, n, o, d, act, na, keys[i], 0u, &hits, &k, NULL, 0);
    }
    free(act);
}

static v3 trace_sample(const job_t* J, uint32_t px, uint32_t py, uint32_t sidx, uint32_t* buf, uint32_t* cand,
                       counters_t* k) {
    const vpo_params* P = J->p;
    const prim_t* prims = J->prims;
    const long n = J->n;
    const cbvh_t* B = J->bvh;
    const int cap = (int)P->active_cap;
    uint32_t* act = buf;
    uint32_t* nxt = buf + cap;

    /* raygen.cuh:82-99 */
    uint64_t pix = (uint64_t)py * P->width + px;
    uint64_t seq = hash64(pix * P->spp + sidx);
    pcg_t rng = pcg_init(P->seed, seq);
    uint64_t key = path_key(P->seed, seq);
    float jx, jy;
    if (P->filter_stddev > 0.0f) {
        float u1 = vpo_max(next_u(&rng), 1e-7f);
        float u2 = next_u(&rng);
        float r = P->filter_stddev * sqrtf(-2.0f * vpo_logf(u1));
        float s, c;
        vpo_sincosf(2.0f * VPO_PI * u2, &s, &c);
        jx = vpo_clamp(r * c, -2.0f, 2.0f);
        jy = vpo_clamp(r * s, -2.0f, 2.0f);
    } else {
        jx = next_u(&rng) - 0.5f;
        jy = next_u(&rng) - 0.5f;
    }
    v3 o, d;
    cam_ray(J->cam, px, py, jx, jy, &o, &d);
    k->n_sample++;

    v3 beta = V(1.0f, 1.0f, 1.0f), L = V(0.0f, 0.0f, 0.0f);

    /* bounce-0 origin-inside set (sampling.cuh:322-341) */
    int na = 0;
    uint32_t overlap = 0;
    if (B) {
        long c = cbvh_point_sorted(B, o, cand);
        for (long q = 0; q < c; ++q)
            if (point_inside(prims + cand[q], o)) {
                ++overlap;
                set_push(act, &na, cap, cand[q], k);
            }
    } else {
        for (long i = 0; i < n; ++i)
            if (point_inside(prims + i, o)) {
                ++overlap;
                set_push(act, &na, cap, (uint32_t)i, k);
            }
    }
    if (P->measure && overlap > k->max_overlap) k->max_overlap = overlap;

    for (uint32_t bounce = 0; bounce < P->max_depth; ++bounce) {
        /* argmin over entered primitives and the origin set (sampling.cuh:350-415) */
        uint32_t hits = 0;
        k->n_trace++;
        float ts_min = argmin_scatter(prims, n, o, d, act, na, key, bounce, &hits, k, B, !P->measure);
        if (P->measure && hits > k->max_hits) k->max_hits = hits;

        /* analytic direct term at bounce 0 (raygen.cuh:148-160), under the
         * shadow-ray roulette as ray 2 of bounce 0 (a no-op when it is off) */
        if (bounce == 0) {
            k->n_shadow_rays++;
            v3 kd = vmul(beta, env_lookup(&J->env, d));
            float T = shadow_T(prims, n, o, d, act, na, k, B, vmax3(kd), P->shadow_rr, key, 0, 2);
            L = vadd(L, vscale(kd, T));
        }
        if (ts_min >= INF_F) break; /* escape: nothing to add with NEE + analytic direct */

        /* rebuild the active set at the scatter point (sampling.cuh:422-458) */
        int nn = 0;
        v3 x = vadd(o, vscale(d, ts_min));
        {
            /* ascending indices: all of them (brute force) or the union of the
             * old set and the primitives whose box contains x */
            int ia = 0;
            long nc = B ? cbvh_point_sorted(B, x, cand) : 0, ic = 0, prevj = -1;
            for (long j = 0;; ++j) {
                if (B) {
                    while (ia < na && (long)act[ia] <= prevj) ++ia;
                    while (ic < nc && (long)cand[ic] <= prevj) ++ic;
                    uint32_t na_next = ia < na ? act[ia] : 0xffffffffu;
                    uint32_t nc_next = ic < nc ? cand[ic] : 0xffffffffu;
                    uint32_t nx = na_next < nc_next ? na_next : nc_next;
                    if (nx == 0xffffffffu) break;
                    j = nx;
                } else if (j >= n) {
                    break;
                }
                prevj = j;
                int in_act = 0;
                while (ia < na && act[ia] < (uint32_t)j) ++ia;
                if (ia < na && act[ia] == (uint32_t)j) in_act = 1;
                const prim_t* pr = prims + j;
                int keep = 0;
                if (in_act) {
                    k->n_rb_active++;
                    if (ts_min <= exit_inside(pr, o, d)) keep = 1;
                }
                if (!keep) {
                    crossing_t x = crossing(pr, o, d);
                    if (x.t0 >= 0.0f && x.t0 <= ts_min) {
                        k->n_rb_hit++;
                        if (ts_min <= x.t1) keep = 1;
                    }
                }
                if (keep) set_push(nxt, &nn, cap, (uint32_t)j, k);
            }
        }
        { uint32_t* t = act; act = nxt; nxt = t; na = nn; }

        if (P->measure) {
            uint32_t ov = 0;
            if (B) {
                long c = cbvh_point_sorted(B, x, cand);
                for (long q = 0; q < c; ++q) ov += (uint32_t)point_inside(prims + cand[q], x);
            } else {
                for (long i = 0; i < n; ++i) ov += (uint32_t)point_inside(prims + i, x);
            }
            if (ov > k->max_overlap) k->max_overlap = ov;
        }
        float phase_pdf;
        v3 wo = hg_sample(&J->hg, d, &rng, &phase_pdf);
        k->n_scatter++;

        /* evaluate_albedo (sampling.cuh:257-283) */
        v3 aa = V(0.0f, 0.0f, 0.0f);
        float aw = 0.0f;
        for (int s = 0; s < na; ++s) {
            const prim_t* pr = prims + act[s];
            k->n_alb++;
            float wgt = pr->mass * prim_pdf(pr, x);
            aa = vadd(aa, vscale(V(pr->ar, pr->ag, pr->ab), wgt));
            aw += wgt;
        }
        v3 albedo = aw <= 0.0f ? V(0.0f, 0.0f, 0.0f) : vscale(aa, vpo_rcp(aw));

        /* accumulate_nee_direct (path_stages.cuh:148-226) */
        v3 base = vmul(beta, albedo);
        if (P->use_ris) {
            float wsum = 0.0f;
            v3 ydir = V(0, 0, 0), yenv = V(0, 0, 0);
            for (uint32_t c = 0; c < P->ris_candidates; ++c) {
                float cpdf;
                v3 cwo = env_sample(&J->env, &rng, &cpdf);
                if (cpdf <= 0.0f) continue;
                v3 envc = env_lookup(&J->env, cwo);
                float lumc = vdot(envc, V(0.2126f, 0.7152f, 0.0722f));
                float wgt = hg_eval_cos(&J->hg, vdot(d, cwo)) * lumc * vpo_rcp(cpdf);
                if (!(wgt > 0.0f)) continue;
                wsum += wgt;
                if (next_u(&rng) * wsum < wgt) { ydir = cwo; yenv = envc; }
            }
            if (wsum > 0.0f) {
                float lumy = vdot(yenv, V(0.2126f, 0.7152f, 0.0722f));
                if (lumy > 0.0f) {
                    k->n_shadow_rays++;
                    float Wr = wsum * vpo_rcp((float)P->ris_candidates);
                    /* the factor T multiplies, formed before the walk */
                    v3 ky = vscale(vmul(base, vscale(yenv, vpo_rcp(lumy))), Wr);
                    float T = shadow_T(prims, n, x, ydir, act, na, k, B, vmax3(ky), P->shadow_rr, key, bounce, 0u);
                    L = vadd(L, vscale(ky, T));
                }
            }
        } else {
            float apdf;
            v3 awo = hg_sample(&J->hg, d, &rng, &apdf);
            float pb_a = env_pdf(&J->env, awo);
            float sa = apdf + pb_a;
            float wa = sa > 0.0f ? apdf / sa : 0.0f;
            v3 enva = env_lookup(&J->env, awo);
            k->n_shadow_rays++;
            v3 ka = vscale(vmul(base, enva), wa);
            float Ta = shadow_T(prims, n, x, awo, act, na, k, B, vmax3(ka), P->shadow_rr, key, bounce, 0u);
            L = vadd(L, vscale(ka, Ta));

            float bpdf;
            v3 bwo = env_sample(&J->env, &rng, &bpdf);
            float ph_b = hg_eval_cos(&J->hg, vdot(d, bwo));
            float sb = bpdf + ph_b;
            float wb = sb > 0.0f ? bpdf / sb : 0.0f;
            v3 envb = env_lookup(&J->env, bwo);
            k->n_shadow_rays++;
            v3 kb = vscale(vscale(vscale(vmul(base, envb), wb), ph_b), vpo_rcp(bpdf));
            float Tb = shadow_T(prims, n, x, bwo, act, na, k, B, vmax3(kb), P->shadow_rr, key, bounce, 1u);
            L = vadd(L, vscale(kb, Tb));
        }
        beta = vmul(beta, albedo);

        /* russian_roulette (path_stages.cuh:25-37) */
        if (bounce >= P->rr_depth) {
            float ps = vpo_min(P->rr_max_survival, vmax3(beta));
            if (next_u(&rng) > ps || ps <= 0.0f) break;
            beta = vscale(beta, vpo_rcp(ps));
        }
        if (vmax3(beta) < 1e-4f) break;
        {
            float sb = beta.x + beta.y + beta.z;
            if (!((vpo_f2u(sb) & 0x7f800000u) != 0x7f800000u)) break;
        }
        o = x;
        d = wo;
    }

    /* finalize_sample (path_stages.cuh:52-62) */
    if (P->firefly_clamp > 0.0f) {
        float lum = vdot(L, V(0.2126f, 0.7152f, 0.0722f));
        if (lum > P->firefly_clamp) L = vscale(L, P->firefly_clamp * vpo_rcp(lum));
    }
    {
        float s = L.x + L.y + L.z;
        if (!((vpo_f2u(s) & 0x7f800000u) != 0x7f800000u)) L = V(0.0f, 0.0f, 0.0f);
    }
    return L;
}

static void counters_add(counters_t* a, const counters_t* b) {
    uint64_t* pa = (uint64_t*)a;
    const uint64_t* pb = (const uint64_t*)b;
    const int nsum = 13; /* fields before max_hits */
    for (int i = 0; i < nsum; ++i) pa[i] += pb[i];
    if (b->max_hits > a->max_hits) a->max_hits = b->max_hits;
    if (b->max_overlap > a->max_overlap) a->max_overlap = b->max_overlap;
}

/* Exact per-pixel sum of sample radiance in 32.32 fixed point (the GPU's
 * radiance_fixed / accum_mean): each channel's low and high 32-bit halves are
 * summed in separate 64-bit words, so the total is an order-free integer. */
typedef struct { uint64_t lo[3], hi[3], n; } accum_t;
static inline uint64_t radiance_fixed(float v) {
    float c = vpo_min(vpo_max(v, 0.0f), 2147483648.0f);
    return (uint64_t)rintf(c * 4294967296.0f);
}
static inline void accum_add(accum_t* a, v3 L) {
    const float c[3] = {L.x, L.y, L.z};
    for (int i = 0; i < 3; ++i) {
        uint64_t q = radiance_fixed(c[i]);
        a->lo[i] += q & 0xffffffffu;
        a->hi[i] += q >> 32;
    }
    a->n += 1;
}
static inline float accum_mean(uint64_t lo, uint64_t hi, uint64_t n) {
    if (n == 0) return 0.0f;
    double s = (double)hi * 4294967296.0 + (double)lo;
    return (float)((s * (1.0 / 4294967296.0)) / (double)n);
}

/* Adaptive sampling of one pixel (raygen.cuh:37-69,214-233 with
 * ENABLE_ADAPTIVE_SAMPLING on; batches as in renderer.cpp:341-372): the
 * samples run in batches; before each batch after the first ones the pixel
 * stops once n >= the minimum and max over channels of sqrt(M2/(n-1)) /
 * max(mean, 1e-6) is below the threshold.  Each sample is folded into the
 * running Welford mean and M2 in sample order.  The mean is the pixel's
 * value and n its sample count. */
static void adaptive_pixel(job_t* J, uint32_t x, uint32_t y, uint32_t* buf, uint32_t* cand, counters_t* k) {
    const vpo_params* P = J->p;
    const uint32_t batch = P->adaptive_batch ? P->adaptive_batch : 16u;
    float mean[3] = {0.0f, 0.0f, 0.0f}, M2[3] = {0.0f, 0.0f, 0.0f};
    uint32_t n = 0;
    for (uint32_t s0 = 0; s0 < P->spp; s0 += batch) {
        if (n >= P->adaptive_min) {
            const float r = 1.0f / (float)(n - 1u);
            float err = 0.0f;
            for (int c = 0; c < 3; ++c) {
                const float e = sqrtf(M2[c] * r) / vpo_max(mean[c], 1e-6f);
                err = c == 0 ? e : vpo_max(err, e);
            }
            if (err < P->adaptive_threshold) break;
        }
        const uint32_t cnt = P->spp - s0 < batch ? P->spp - s0 : batch;
        for (uint32_t j = 0; j < cnt; ++j) {
            v3 Ls = trace_sample(J, x, y, s0 + j, buf, cand, k);
            const float xs[3] = {Ls.x, Ls.y, Ls.z};
            const float n_inv = 1.0f / (float)(n + j + 1u);
            for (int c = 0; c < 3; ++c) {
                const float d1 = xs[c] - mean[c];
                mean[c] = mean[c] + d1 * n_inv;
                M2[c] = M2[c] + d1 * (xs[c] - mean[c]);
            }
        }
        n += cnt;
    }
    const size_t idx = (size_t)(y - P->y0) * (P->x1 - P->x0) + (x - P->x0);
    if (J->out_mean) {
        J->out_mean[4 * idx] = mean[0];
        J->out_mean[4 * idx + 1] = mean[1];
        J->out_mean[4 * idx + 2] = mean[2];
        J->out_mean[4 * idx + 3] = 1.0f;
    }
    if (J->out_counts) J->out_counts[idx] = n;
}

static void* worker(void* arg) {
    job_t* J = (job_t*)arg;
    const vpo_params* P = J->p;
    uint32_t* buf = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (P->active_cap ? P->active_cap : 1));
    uint32_t* cand = J->bvh ? (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(J->n > 0 ? J->n : 1)) : NULL;
    counters_t k;
    memset(&k, 0, sizeof(k));
    for (;;) {
        pthread_mutex_lock(&J->mu);
        uint32_t y = J->next_row++;
        pthread_mutex_unlock(&J->mu);
        if (y >= P->y1) break;
        const uint32_t st = P->stride ? P->stride : 1u;
        if ((y - P->y0) % st) continue;
        for (uint32_t x = P->x0; x < P->x1; x += st) {
            if (P->adaptive_threshold > 0.0f) {
                adaptive_pixel(J, x, y, buf, cand, &k);
                continue;
            }
            /* mean over the samples (raygen.cuh:214-219), as an exact
             * fixed-point sum */
            accum_t acc;
            memset(&acc, 0, sizeof(acc));
            for (uint32_t s = P->sample_begin; s < P->sample_end; ++s) {
                v3 Ls = trace_sample(J, x, y, s, buf, cand, &k);
                if (J->out_samples) {
                    size_t idx = (((size_t)(s - P->sample_begin) * (P->y1 - P->y0) + (y - P->y0)) * (P->x1 - P->x0) + (x - P->x0)) * 4;
                    J->out_samples[idx] = Ls.x;
                    J->out_samples[idx + 1] = Ls.y;
                    J->out_samples[idx + 2] = Ls.z;
                    J->out_samples[idx + 3] = 1.0f;
                }
                accum_add(&acc, Ls);
            }
            v3 mean = V(accum_mean(acc.lo[0], acc.hi[0], acc.n), accum_mean(acc.lo[1], acc.hi[1], acc.n),
                        accum_mean(acc.lo[2], acc.hi[2], acc.n));
            if (J->out_mean) {
                size_t idx = ((size_t)(y - P->y0) * (P->x1 - P->x0) + (x - P->x0)) * 4;
                J->out_mean[idx] = mean.x;
                J->out_mean[idx + 1] = mean.y;
                J->out_mean[idx + 2] = mean.z;
                J->out_mean[idx + 3] = 1.0f;
            }
        }
    }
    pthread_mutex_lock(&J->mu);
    counters_add(&J->total, &k);
    pthread_mutex_unlock(&J->mu);
    free(buf);
    free(cand);
    return NULL;
}

/* prims16: records in scene order (the GPU's sorted order); env arrays from
 * vpo_env_tables (prob may be NULL for the uniform fallback); counters: 15
 * uint64 in counters_t order.  Returns 0 on success. */
int vpo_render(const float* prims16, long n, const float* env_rgba, const float* joint, const float* prob,
               const uint32_t* alias, float env_total, uint32_t ew, uint32_t eh, const float* cam_pod,
               const vpo_params* p, float* out_mean, float* out_samples, uint64_t* counters, int threads,
               uint32_t* out_counts) {
    if (p->active_cap == 0 || p->x1 > p->width || p->y1 > p->height || p->x0 >= p->x1 || p->y0 >= p->y1 ||
        p->sample_end > p->spp || p->sample_begin >= p->sample_end || ew == 0 || eh == 0)
        return 1;
    /* adaptive sampling covers all samples and writes no per-sample values */
    if (p->adaptive_threshold > 0.0f &&
        (p->sample_begin != 0 || p->sample_end != p->spp || out_samples || p->adaptive_min == 0))
        return 1;
    job_t J;
    memset(&J, 0, sizeof(J));
    J.prims = (const prim_t*)prims16;
    J.n = n;
    J.env.tex = env_rgba;
    J.env.joint = joint;
    J.env.prob = prob;
    J.env.alias = alias;
    J.env.total = env_total;
    J.env.w = ew;
    J.env.h = eh;
    J.cam = cam_pod;
    J.p = p;
    J.hg = hg_make(p->phase_g);
    J.out_mean = out_mean;
    J.out_samples = out_samples;
    J.out_counts = out_counts;
    J.next_row = p->y0;
    cbvh_t bvh = {NULL, 0};
    if (p->use_bvh) {
        cbvh_build(&bvh, J.prims, n);
        J.bvh = &bvh;
    }
    pthread_mutex_init(&J.mu, NULL);
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, worker, &J);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&J.mu);
    free(bvh.nodes);
    if (counters) memcpy(counters, &J.total, sizeof(counters_t));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* test access to the phase function and the environment sampler            */
/* ------------------------------------------------------------------------ */
void vpo_phase_eval(float g, const float* wi, const float* wo, long n, float* out) {
    hg_t h = hg_make(g);
    for (long i = 0; i < n; ++i) out[i] = hg_eval_cos(&h, vdot(vld(wi + 3 * i), vld(wo + 3 * i)));
}
void vpo_phase_sample(float g, const float* wi, uint64_t seed, long n, float* wo, float* pdf) {
    hg_t h = hg_make(g);
    pcg_t r = pcg_init(seed, 7u);
    for (long i = 0; i < n; ++i) {
        v3 o = hg_sample(&h, vld(wi + 3 * i), &r, pdf + i);
        vst(wo + 3 * i, o);
    }
}
void vpo_env_sample(const float* rgba, const float* joint, const float* prob, const uint32_t* alias, float total,
                    uint32_t w, uint32_t h, uint64_t seed, long n, float* wo, float* pdf) {
    env_t e = {rgba, prob, alias, joint, w, h, total};
    pcg_t r = pcg_init(seed, 11u);
    for (long i = 0; i < n; ++i) vst(wo + 3 * i, env_sample(&e, &r, pdf + i));
}
void vpo_env_pdf(const float* rgba, const float* joint, const float* prob, const uint32_t* alias, float total,
                 uint32_t w, uint32_t h, const float* wo, long n, float* pdf) {
    env_t e = {rgba, prob, alias, joint, w, h, total};
    for (long i = 0; i < n; ++i) pdf[i] = env_pdf(&e, vld(wo + 3 * i));
}
void vpo_env_lookup(const float* rgba, uint32_t w, uint32_t h, const float* dirs, long n, float* rgb) {
    env_t e = {rgba, NULL, NULL, NULL, w, h, 0.0f};
    for (long i = 0; i < n; ++i) vst(rgb + 3 * i, env_lookup(&e, vld(dirs + 3 * i)));
}
void vpo_camera_ray(const float* pod, uint32_t px, uint32_t py, float jx, float jy, float* o, float* d) {
    v3 oo, dd;
    cam_ray(pod, px, py, jx, jy, &oo, &dd);
    vst(o, oo);
    vst(d, dd);
}
float vpo_prim_uniform(uint64_t key, uint32_t bounce, uint32_t role, uint32_t prim) { return prim_u(key, bounce, role, prim); }
