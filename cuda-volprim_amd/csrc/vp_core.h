// vp_core.h — hot-path arithmetic of the Gaussian-mixture volumetric path
// tracer: primitive records, ray/ellipsoid tests, free-flight inversion,
// optical depth, RNG, phase function, environment lookup and importance
// sampling, camera rays.  Everything is VP_HD (host + device) and uses only
// vp_math.h, so a host build and the HIP kernels evaluate identical bits.
//
// Reference correspondence (kfernandez31/cuda-volprim):
//   PrimRecord, transform_*_local, pdf, inv_cdf, inv_cdf_span, optical_depth
//       include/thesis/device/params/primitive.h:27-318
//   exit_from_inside, compute_exit_from_entry, point_inside_bvh_bound
//       include/thesis/common/geometry/intersection.h:56-173
//   entry test (OptiX built-in sphere at 3 sigma, front face only)
//       src/thesis/host/app/renderer.cpp:142-157, device/entry/anyhit.cuh:68-86
//   quaternion rotate                 include/thesis/common/geometry/quat.h:67-71
//   PCG32 / hash / sample_uniform     device/core/random.cuh:16-72
//   HG phase                          device/core/sampling.cuh:74-134
//   env lookup                        include/thesis/device/params/environment_map.h:36-61
//   env importance sampling pdf       device/core/sampling.cuh:152-246
//   camera ray                        include/thesis/device/params/camera.h:33-57
#pragma once

#include "vp_math.h"

namespace vp {

// ---------------------------------------------------------------------------
// 3-vectors with the reference's operator semantics (component-wise, a - b
// evaluated as a + (-b), division as multiplication by a rounded reciprocal).
// ---------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};
VP_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
VP_HD f3 mk(float s) { return f3{s, s, s}; }
VP_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
VP_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
VP_HD f3 operator-(f3 a, f3 b) { return f3{a.x + (-b.x), a.y + (-b.y), a.z + (-b.z)}; }
VP_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
VP_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
VP_HD f3 operator*(float s, f3 a) { return f3{a.x * s, a.y * s, a.z * s}; }
VP_HD f3 fma3(float a, f3 b, f3 c) { return f3{fmaf_(a, b.x, c.x), fmaf_(a, b.y, c.y), fmaf_(a, b.z, c.z)}; }
VP_HD float dot(f3 a, f3 b) { return fmaf_(a.x, b.x, fmaf_(a.y, b.y, a.z * b.z)); }
VP_HD float length2(f3 a) { return dot(a, a); }
VP_HD f3 cross(f3 a, f3 b) {
    return f3{fmaf_(a.y, b.z, -(a.z * b.y)), fmaf_(a.z, b.x, -(a.x * b.z)), fmaf_(a.x, b.y, -(a.y * b.x))};
}
VP_HD float max3(f3 v) { return maxf(maxf(v.x, v.y), v.z); }
VP_HD float sum3(f3 v) { return v.x + v.y + v.z; }
VP_HD f3 normalize(f3 v) { return v * rsqrtf_(length2(v)); }

// ---------------------------------------------------------------------------
// Constants (reference device/core/constants.cuh and common/utils/math.h)
// ---------------------------------------------------------------------------
constexpr float kTMax = 1e20f;               // INF_F: ray extent and "no scatter"
constexpr float kExtent = 3.0f;              // GAUSSIAN_EXTENT_F (3 sigma)
constexpr float kExtent2 = 9.0f;             // GAUSSIAN_EXTENT_F^2
constexpr float kDiameter = 6.0f;            // GAUSSIAN_DIAMETER_F
constexpr float kSegMin = 1e-6f;             // RAY_SEGMENT_MIN_LENGTH
constexpr float kMaxTau = 88.0f;             // MAX_OPTICAL_DEPTH
constexpr float kDirMin2 = 1e-12f;           // RAY_DIRECTION_MIN_LENGTH2
constexpr float kDiscEps = 1e-12f;           // INTERSECTION_DISCRIMINANT_EPS
constexpr float kMinThroughput = 1e-4f;      // MIN_THROUGHPUT
constexpr float kHgIsoEps = 1e-3f;           // HG_ISOTROPIC_EPS
constexpr float kLumR = 0.2126f, kLumG = 0.7152f, kLumB = 0.0722f;

// Optical depth is summed in 24.40 fixed point so that the total along a
// shadow ray does not depend on the order primitives are visited in.
constexpr double kTauScale = 1099511627776.0;  // 2^40
constexpr float kTauScaleF = 1099511627776.0f;
constexpr uint64_t kTauLimitFixed = 96757023244288ull;  // 88 * 2^40

// ---------------------------------------------------------------------------
// Primitive record: 64 bytes, the reference's hot fields in its order.
// ---------------------------------------------------------------------------
struct alignas(16) PrimRecord {
    float cx, cy, cz;     // centre
    float qw, qx, qy, qz; // CONJUGATE of the forward rotation (world -> local)
    float rsx, rsy, rsz;  // 1 / scale
    float dnf;            // density_norm_factor = (2 pi)^(-3/2) / prod(scale)
    float icf;            // inv_cdf_factor = M * dnf * sqrt(pi/2)
    float ar, ag, ab;     // albedo
    float mass;           // optical_thickness M (total mass)
};
static_assert(sizeof(PrimRecord) == 64, "PrimRecord must stay 64 bytes");

// Build a record from a forward rotation (w, x, y, z), per-axis scale, albedo
// and mass.  Same arithmetic as Primitive::from_forward_quat and
// UnitQuaternion::from (primitive.h:70-87,108-111; quat.h:54-57).
VP_HD PrimRecord make_prim(f3 c, float w, float x, float y, float z, f3 scale, f3 albedo, float mass) {
    const float l2 = fmaf_(w, w, fmaf_(x, x, fmaf_(y, y, z * z)));
    const float r = 1.0f / sqrtf_(l2);
    PrimRecord p;
    p.cx = c.x;
    p.cy = c.y;
    p.cz = c.z;
    p.qw = w * r;
    p.qx = -(x * r);
    p.qy = -(y * r);
    p.qz = -(z * r);
    p.rsx = 1.0f / scale.x;
    p.rsy = 1.0f / scale.y;
    p.rsz = 1.0f / scale.z;
    p.dnf = kOneOverTwoPiPow32 * (p.rsx * p.rsy * p.rsz);
    p.icf = mass * p.dnf * kRootHalfPi;
    p.ar = albedo.x;
    p.ag = albedo.y;
    p.ab = albedo.z;
    p.mass = mass;
    return p;
}

VP_HD f3 prim_center(const PrimRecord& p) { return f3{p.cx, p.cy, p.cz}; }
VP_HD f3 prim_rcp_scale(const PrimRecord& p) { return f3{p.rsx, p.rsy, p.rsz}; }

// v' = v + 2 s (u x v) + 2 u x (u x v)  (quat.h:67-71)
VP_HD f3 quat_rotate(float s, f3 u, f3 v) {
    const f3 uv = cross(u, v);
    const f3 uuv = cross(u, uv);
    return fma3(2.0f, fma3(s, uv, uuv), v);
}
VP_HD f3 to_local_pos(const PrimRecord& p, f3 pos) {
    return quat_rotate(p.qw, f3{p.qx, p.qy, p.qz}, pos - prim_center(p)) * prim_rcp_scale(p);
}
VP_HD f3 to_local_dir(const PrimRecord& p, f3 dir) {
    return quat_rotate(p.qw, f3{p.qx, p.qy, p.qz}, dir) * prim_rcp_scale(p);
}

// Gaussian density at pos (primitive.h:127-135).
VP_HD float prim_pdf(const PrimRecord& p, f3 pos) {
    const f3 l = to_local_pos(p, pos);
    return expf_(-0.5f * length2(l)) * p.dnf;
}

// Free-flight inverse CDF from t = 0 (primitive.h:140-186).  Returns kTMax
// when the primitive's mass along the ray is below tau.
VP_HD float inv_cdf(const PrimRecord& pr, f3 o, f3 d, float tau) {
    const f3 w = to_local_dir(pr, d);
    const f3 p = to_local_pos(pr, o);
    const float w_len2 = length2(w);
    const float w_inv_len = rsqrtf_(w_len2);
    const float wp = dot(w, p) * w_inv_len;
    const float pp = length2(p);
    const float diff = fmaf_(-wp, wp, pp);
    const float K = w_inv_len * expf_(-0.5f * diff) * pr.icf;
    const float raw = erff_(wp * kOneOverRootTwo) + tau * rcpf(K);
    if (raw > 1.0f) return kTMax;
    const float v = erfinvf_(clampf(raw, -1.0f, 1.0f));
    return (kRootTwo * v - wp) * w_inv_len;
}

// Free-flight inverse CDF restricted to t >= t_off (primitive.h:203-246).
VP_HD float inv_cdf_span(const PrimRecord& pr, f3 o, f3 d, float t_off, float tau) {
    const f3 w = to_local_dir(pr, d);
    const f3 p = to_local_pos(pr, o);
    const float w_len2 = length2(w);
    const float w_inv_len = rsqrtf_(w_len2);
    const float w_len = w_len2 * w_inv_len;
    const float wp = dot(w, p) * w_inv_len;
    const float pp = length2(p);
    const float diff = fmaf_(-wp, wp, pp);
    const float K = w_inv_len * expf_(-0.5f * diff) * pr.icf;
    const float wp_off = fmaf_(t_off, w_len, wp);
    const float raw = erff_(wp_off * kOneOverRootTwo) + tau * rcpf(K);
    if (raw > 1.0f) return kTMax;
    const float v = erfinvf_(clampf(raw, -1.0f, 1.0f));
    return (kRootTwo * v - wp) * w_inv_len;
}

// Optical depth of one primitive over [t0, t1] (primitive.h:249-318).  The
// erf difference is evaluated with erf_diff_ (erfc on the tails).
VP_HD float optical_depth(const PrimRecord& pr, f3 o, f3 d, float t0, float t1) {
    if (t1 - t0 <= kSegMin) return 0.0f;
    const f3 w = to_local_dir(pr, d);
    const f3 p = to_local_pos(pr, o);
    const float w_inv_len = rsqrtf_(length2(w));
    const float t0s = maxf(0.0f, t0);
    const float t1s = maxf(t0s + kSegMin, t1);
    const f3 wn = w * w_inv_len;
    const f3 ps = p + t0s * w;
    const float t_limit = clampf((t1s - t0s) * rcpf(w_inv_len), 0.0f, kDiameter);
    const float B = dot(wn, ps);
    const float C = length2(ps);
    const float perp = fmaf_(-B, B, C);
    const float e = expf_(-0.5f * perp);
    const float G = kRootTwoPi * w_inv_len;
    const float bs = B * kOneOverRootTwo;
    const float term = erf_diff_(bs, fmaf_(t_limit, kOneOverRootTwo, bs)) * 0.5f;
    return pr.mass * G * e * term * pr.dnf;
}

// Exit of a ray whose origin is inside the 3 sigma bound (intersection.h:115-126).
VP_HD float exit_from_inside(const PrimRecord& pr, f3 o, f3 d) {
    const f3 w = to_local_dir(pr, d);
    const f3 p = to_local_pos(pr, o);
    const float a = length2(w);
    const float b = dot(p, w);
    const float c = length2(p) - kExtent2;
    const float disc = fmaf_(-a, c, b * b);
    return (-b + sqrtf_(disc)) * rcpf(a);
}

// 2 e^(-9/2): the chord depth's constant factor (chord_poly_ in vp_math.h).
constexpr float kChordScale = 0x1.6c0504p-6f;

// A ray's crossing of a primitive's 3 sigma bound from outside, over
// [0, kTMax): the contract of the reference's built-in-sphere traversal
// (front face only, no entry when the origin is inside or on the bound,
// tangents excluded; src/thesis/host/app/renderer.cpp:142-157,
// device/entry/anyhit.cuh:68-86, intersection.h:47-55), with the entry, the
// exit and the optical depth of the whole chord from one transform of the
// ray into the primitive's frame.  In that frame the ray is p + t w, and the
// quadratic |p + t w|^2 = 9 is solved in the stable form of
// intersection.h:66-101: b = -p.w, q0 = p + (b/a) w the point closest to the
// centre, delta = 9 - |q0|^2, q = b + sqrt(a delta); the entry is c/q and the
// exit the other root q/a (compute_exit_from_entry, intersection.h:142-173,
// gives the same root from the entry point).  The chord's optical depth
// (optical_depth over [entry, exit], primitive.h:249-318) is, with the
// Gaussian factored about q0,
//   M dnf sqrt(2 pi)/|w| e^(-|q0|^2/2) erf(sqrt(delta/2))
//     = 2 e^(-9/2) M dnf (sqrt(a delta)/a) S(delta),
// S from chord_poly_ (no exp, no erf).  Chords of length <= kSegMin have
// depth 0, as in optical_depth.  t0 < 0 when the ray does not enter.
struct Crossing {
    float t0, t1, tau;
};
// The first part of crossing(): false where crossing() finds no entry
// before solving for it (a direction too short, an origin inside or on the
// bound, a ray moving away, a miss or a tangent).  The same operations, so a
// leaf pass can sort out beforehand the leaves crossing() would reject.
struct CrossingSetup {
    float a, b, c, a_inv, delta;
};
VP_HD bool crossing_setup(const PrimRecord& pr, f3 o, f3 d, CrossingSetup& s) {
    const f3 w = to_local_dir(pr, d);
    const f3 p = to_local_pos(pr, o);
    s.a = length2(w);
    if (s.a < kDirMin2) return false;
    s.c = length2(p) - kExtent2;
    if (!(s.c > 0.0f)) return false;
    s.b = -dot(p, w);
    if (!(s.b > 0.0f)) return false;  // moving away: both roots behind the origin
    s.a_inv = rcpf(s.a);
    const f3 q0 = fma3(s.b * s.a_inv, w, p);
    s.delta = kExtent2 - length2(q0);
    return !(s.delta < kDiscEps);
}
VP_HD Crossing crossing(const PrimRecord& pr, f3 o, f3 d) {
    Crossing r{-1.0f, -1.0f, 0.0f};
    CrossingSetup s;
    if (!crossing_setup(pr, o, d, s)) return r;
    const float a = s.a, b = s.b, c = s.c, a_inv = s.a_inv, delta = s.delta;
    const float sq = sqrtf_(a * delta);
    const float q = b + sq;
    const float t = c * rcpf(q);
    if (!(t >= 0.0f && t < kTMax)) return r;
    r.t0 = t;
    r.t1 = q * a_inv;
    if (r.t1 - t > kSegMin && r.t1 < kTMax)
        r.tau = ((kChordScale * (pr.mass * pr.dnf)) * (sq * a_inv)) * chord_poly_(delta);
    return r;
}

// Entry distance alone (negative when the ray does not enter).
VP_HD float entry_t(const PrimRecord& pr, f3 o, f3 d) { return crossing(pr, o, d).t0; }

// e^(-9/2)
constexpr float kExpHalfExtent2 = 0x1.6c0504p-7f;

// The chord of a ray whose origin lies inside the 3 sigma bound, from the
// origin to the exit, from one transform: the exit of exit_from_inside
// (intersection.h:115-126) and the optical depth over [0, exit]
// (optical_depth, primitive.h:249-318).  With b = p.w, u0 = b/|w| the
// origin's coordinate along the ray from the closest point q0,
// delta = 9 - |q0|^2 = (b^2 - a c)/a and |p|^2 = |q0|^2 + u0^2, the depth
//   M dnf sqrt(pi/2)/|w| e^(-|q0|^2/2) [erf(sqrt(delta/2)) - erf(u0/sqrt 2)]
// is, by e^(x^2) erf(x) = (2/sqrt(pi)) x S(2x^2),
//   M dnf [e^(-9/2) (sqrt(a delta)/a) S(delta) - e^(-|p|^2/2) (b/a) S(b^2/a)]
// with S from chord_poly_: one exp and no erf.  Exits at or below kSegMin (and
// origins the rounding puts outside) have depth 0, as optical_depth gives.
struct InsideChord {
    float t1, tau;
};
VP_HD InsideChord inside_chord(const PrimRecord& pr, f3 o, f3 d) {
    const f3 w = to_local_dir(pr, d);
    const f3 p = to_local_pos(pr, o);
    const float a = length2(w);
    const float b = dot(p, w);
    const float pp = length2(p);
    const float c = pp - kExtent2;
    const float disc = fmaf_(-a, c, b * b);
    const float sq = sqrtf_(disc);
    const float a_inv = rcpf(a);
    InsideChord r;
    r.t1 = (-b + sq) * a_inv;
    r.tau = 0.0f;
    if (r.t1 > kSegMin && r.t1 < kTMax) {
        const float far = (kExpHalfExtent2 * (sq * a_inv)) * chord_poly_(minf(disc * a_inv, kExtent2));
        const float near = (expf_(-0.5f * pp) * (b * a_inv)) * chord_poly_(minf((b * b) * a_inv, kExtent2));
        r.tau = maxf((pr.mass * pr.dnf) * (far - near), 0.0f);
    }
    return r;
}

// Free flight from an origin inside a primitive (argmin over the origin set,
// sampling.cuh:356-370): it ends inside the chord exactly when the sampled
// depth tau is at most the chord's; the distance is then inv_cdf
// (primitive.h:140-186) kept inside [0, exit].  kTMax when it passes.
VP_HD float origin_flight(const PrimRecord& pr, f3 o, f3 d, const InsideChord& x, float tau) {
    if (!(tau <= x.tau)) return kTMax;
    return clampf(inv_cdf(pr, o, d, tau), 0.0f, x.t1);
}

// Free flight of a ray that entered a primitive (argmin over the entered
// primitives, sampling.cuh:372-412): with tau the path's optical-depth sample
// for the primitive, the flight ends inside its chord exactly when tau is at
// most the chord's depth; the distance is then the inverse CDF from the entry
// (inv_cdf_span, primitive.h:203-246), kept inside [entry, exit] as the
// reference keeps it (a result below the entry is clamped up to it; the test
// against the exit is the depth comparison).  kTMax when the flight passes
// through the chord.
VP_HD float entered_flight(const PrimRecord& pr, f3 o, f3 d, const Crossing& x, float tau) {
    if (!(tau <= x.tau)) return kTMax;
    return clampf(inv_cdf_span(pr, o, d, x.t0, tau), x.t0, x.t1);
}

// Point inside the 3 sigma bound (intersection.h:56-61).
VP_HD bool point_inside(const PrimRecord& pr, f3 x) { return length2(to_local_pos(pr, x)) <= kExtent2; }

// ---------------------------------------------------------------------------
// RNG (random.cuh:16-53) and the per-primitive counter draw.
// ---------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state, inc;
};
VP_HD uint32_t pcg_next(Pcg32& r) {
    const uint64_t old = r.state;
    r.state = old * 6364136223846793005ull + r.inc;
    const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
VP_HD uint64_t hash64(uint64_t x) {
    x = x * 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x = x * 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}
VP_HD Pcg32 pcg_init(uint64_t seed, uint64_t seq) {
    Pcg32 r;
    r.state = 0;
    r.inc = (seq << 1u) | 1u;
    pcg_next(r);
    r.state += seed;
    pcg_next(r);
    return r;
}
VP_HD float u32_to_unit(uint32_t v) { return static_cast<float>(v) / 4294967296.0f; }
VP_HD float next_uniform(Pcg32& r) { return u32_to_unit(pcg_next(r)); }

// Key for a path's per-primitive draws, fixed by (seed, sequence).
VP_HD uint64_t path_key(uint64_t seed, uint64_t seq) { return hash64(seq ^ (seed * 0x9e3779b97f4a7c15ull) ^ 0x243f6a8885a308d3ull); }

// Uniform for primitive `prim` at bounce `bounce` of the path with key k, in
// role 0 (the primitive contains the ray origin) or 1 (the ray enters it): a
// SplitMix64 finaliser of a Weyl step over the packed (bounce, role, prim)
// counter.  The draw depends only on these, never on visiting order.
VP_HD float prim_uniform(uint64_t key, uint32_t bounce, uint32_t role, uint32_t prim) {
    uint64_t x = key + 0x9e3779b97f4a7c15ull * ((static_cast<uint64_t>((bounce << 1) | role) << 32) | prim);
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return u32_to_unit(static_cast<uint32_t>(x >> 32));
}

// tau = -log(1 - xi) (sampling.cuh:310-313)
VP_HD float free_flight_tau(float xi) { return -logf_(maxf(1.0f - xi, 1e-10f)); }

// ---------------------------------------------------------------------------
// Shadow-ray roulette, an estimator option of this build (the reference walks
// every shadow ray to the tau = 88 saturation, sampling.cuh:548-583).  A NEE
// shadow ray whose contribution is w * T(tau) (w: the largest channel of the
// term T multiplies) plays a Russian roulette at each of the depths
// theta_j = tau_c + j ln 4 (j = 0, 1, ...) it reaches, tau_c = log(w / eps)
// being where w e^(-tau) falls below eps: it survives each with probability
// 1/4, and a ray that reached L levels and survived them all counts
// 4^L T, else 0.  E[4^L T 1(survived L)] = T, so the estimator stays
// unbiased; each level holds the surviving contribution 4^L w e^(-tau) near
// eps, so the variance added per ray stays of order eps^2.  The survival
// draws are counter hashes of (path, bounce, ray, level), so the first level
// a ray fails is known before its walk starts: the walk stops at that depth
// (or at the tau = 88 saturation), and since whether a sum reaches a depth
// does not depend on the order chords are summed in, the result matches a
// full walk bit for bit.  eps = 0 turns it off.  Depths are 24.40
// fixed-point; thresholds and the level step are multiples of 2^24.
// ---------------------------------------------------------------------------
constexpr uint64_t kRrGrain = 0xffffffull;                      // thresholds are multiples of 2^24
constexpr uint64_t kRrStep = 1524246769571ull & ~kRrGrain;      // ln 4 * 2^40, rounded down
constexpr uint32_t kRrNoFail = 127u;                             // no failing level below the saturation
VP_HD uint64_t rr_threshold(float w, float eps) {
    if (!(eps > 0.0f)) return kTauLimitFixed;
    if (!(w > eps)) return 0ull;
    const float tc = logf_(w * rcpf(eps));
    if (!(tc > 0.0f)) return 0ull;
    if (tc >= kMaxTau) return kTauLimitFixed;
    return static_cast<uint64_t>(tc * kTauScaleF) & ~kRrGrain;
}
VP_HD float rr_uniform(uint64_t key, uint32_t bounce, uint32_t ray, uint32_t level) {
    const uint64_t c = (static_cast<uint64_t>(bounce) << 16) | (static_cast<uint64_t>(ray) << 8) | level;
    uint64_t x = (key ^ 0xd1b54a32d192ed03ull) + 0x9e3779b97f4a7c15ull * c;
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return u32_to_unit(static_cast<uint32_t>(x >> 32));
}
// The roulette of one shadow ray: its first level tau_c (kTauLimitFixed when
// the roulette is off) and the first level it fails, fail (kRrNoFail when it
// survives every level below the saturation); its walk may stop at limit().
struct ShadowRr {
    uint64_t thr;
    uint32_t fail;
    VP_HD uint64_t limit() const { return fail == kRrNoFail ? kTauLimitFixed : thr + fail * kRrStep; }
};
VP_HD ShadowRr shadow_rr(float w, float eps, uint64_t key, uint32_t bounce, uint32_t ray) {
    ShadowRr r;
    r.thr = rr_threshold(w, eps);
    r.fail = kRrNoFail;
    for (uint32_t j = 0; r.thr + j * kRrStep < kTauLimitFixed; ++j)
        if (!(rr_uniform(key, bounce, ray, j) < 0.25f)) {
            r.fail = j;
            break;
        }
    return r;
}
// Transmittance from the summed depth acc (T = depth_T(acc)) under the
// roulette rr: 0 past the failing level, else T times 4 per level reached
// (two exact power-of-two scalings: at most 64 levels below tau = 88).
VP_HD float shadow_rr_T(uint64_t acc, const ShadowRr& rr, float T) {
    if (acc >= rr.limit()) return 0.0f;
    if (acc < rr.thr) return T;
    const int e = 2 * static_cast<int>((acc - rr.thr) / kRrStep + 1u);
    const int e1 = e >> 1;
    return (T * exp2i_(e1)) * exp2i_(e - e1);
}

// ---------------------------------------------------------------------------
// Per-pixel radiance accumulator.  The reference folds each sample into a
// running Welford mean in sample order (raygen.cuh:214-238).  Here every
// sample's radiance is added to an exact integer sum in 32.32 fixed point, so
// the total does not depend on the order in which paths finish (any launch
// shape, job mix or number of concurrent launches gives the same bits), and
// the mean is formed once at the end.  Each channel is kept as two 64-bit
// words, the low and high 32 bits of every sample's value summed separately,
// which cannot overflow before 2^32 samples.  Eight words per pixel:
//   [0..2] low sums r, g, b   [3] sample count   [4..6] high sums   [7] unused
// A sample value is clamped to [0, 2^31] (finalize_sample already maps
// non-finite radiance to 0); values below 2^-33 round to 0.
// ---------------------------------------------------------------------------
constexpr uint32_t kAccumWords = 8;
VP_HD uint64_t radiance_fixed(float v) {
    const float c = minf(maxf(v, 0.0f), 2147483648.0f);
    return static_cast<uint64_t>(rintf_(c * 4294967296.0f));
}
// Mean of one channel from its low and high sums over n samples.
VP_HD float accum_mean(uint64_t lo, uint64_t hi, uint64_t n) {
    if (n == 0) return 0.0f;
    const double s = static_cast<double>(hi) * 4294967296.0 + static_cast<double>(lo);
    return static_cast<float>((s * (1.0 / 4294967296.0)) / static_cast<double>(n));
}

// ---------------------------------------------------------------------------
// Henyey–Greenstein phase (sampling.cuh:74-134, constants folded on the host
// as renderer.cpp:224-235 does).
// ---------------------------------------------------------------------------
struct PhaseParams {
    int isotropic;
    float g, g_eff, one_plus_g2, one_minus_g2, neg_inv_2g, coeff;
};
VP_HD PhaseParams make_phase(float g) {
    PhaseParams h;
    const float abs_g = g < 0.0f ? -g : g;
    h.isotropic = abs_g < kHgIsoEps ? 1 : 0;
    h.g = g;
    h.g_eff = -g;
    const float g2 = g * g;
    h.one_plus_g2 = 1.0f + g2;
    h.one_minus_g2 = 1.0f - g2;
    h.neg_inv_2g = h.isotropic ? 0.0f : (-0.5f / h.g_eff);
    h.coeff = kOneOverFourPi * h.one_minus_g2;
    return h;
}
VP_HD float phase_eval_cos(const PhaseParams& h, float c) {
    if (h.isotropic) return kOneOverFourPi;
    const float denom = fmaf_(-2.0f * h.g, c, h.one_plus_g2);
    const float is = rsqrtf_(denom);
    return h.coeff * (is * is * is);
}
VP_HD float phase_eval(const PhaseParams& h, f3 wi, f3 wo) { return phase_eval_cos(h, dot(wi, wo)); }

VP_HD void onb(f3 n, f3& tx, f3& ty) {
    const float sign = copysignf_(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    tx = f3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    ty = f3{b, sign + n.y * n.y * a, -n.y};
}

// Returns wo; *pdf receives the phase value at the sampled cosine.
VP_HD f3 phase_sample(const PhaseParams& h, f3 wi, Pcg32& rng, float* pdf) {
    const float ux = next_uniform(rng);
    const float uy = next_uniform(rng);
    float c;
    if (h.isotropic) {
        c = fmaf_(-2.0f, ux, 1.0f);
    } else {
        const float inner = fmaf_(-2.0f, ux, 1.0f);
        const float denom = fmaf_(h.g_eff, inner, 1.0f);
        const float sq = h.one_minus_g2 * rcpf(denom);
        c = h.neg_inv_2g * fmaf_(-sq, sq, h.one_plus_g2);
    }
    const float s = sqrtf_(maxf(0.0f, 1.0f - c * c));
    const float phi = kTwoPi * uy;
    float sp, cp;
    sincosf_(phi, &sp, &cp);
    f3 ex, ey;
    onb(wi, ex, ey);
    const f3 wo = (ex * (s * cp) + ey * (s * sp)) + wi * c;
    *pdf = phase_eval_cos(h, c);
    return wo;
}

// ---------------------------------------------------------------------------
// Environment: lat-long RGBA32F texture, bilinear, wrap in u, clamp in v.
// ---------------------------------------------------------------------------
struct AliasEntry {
    float prob;      // acceptance threshold of the texel's own column
    uint32_t alias;  // texel taken when the threshold is not met
};

struct EnvView {
    const float* tex;         // W*H*4 floats
    const AliasEntry* alias;  // W*H entries, null when no importance table
    const float* joint;       // W*H luminance * sin(theta) weights
    uint32_t w, h;
    float total;              // sum of joint weights
};

VP_HD float lerpf(float a, float b, float t) { return fmaf_(t, b - a, a); }

VP_HD void dir_to_uv(f3 d, float* u, float* v) {
    const float theta = atan2f_(d.z, d.x);
    const float phi = acosf_(clampf(d.y, -1.0f, 1.0f));
    *u = fmaf_(theta, kOneOverTwoPi, 0.5f);
    *v = phi * kOneOverPi;
}

VP_HD f3 env_texel(const EnvView& e, int i, int j) {
    const float* t = e.tex + 4u * (static_cast<uint32_t>(j) * e.w + static_cast<uint32_t>(i));
    return f3{t[0], t[1], t[2]};
}

VP_HD f3 env_lookup(const EnvView& e, f3 d) {
    float u, v;
    dir_to_uv(d, &u, &v);
    const float x = u * static_cast<float>(e.w) - 0.5f;
    const float y = v * static_cast<float>(e.h) - 0.5f;
    const float x0 = floorf_(x);
    const float y0 = floorf_(y);
    const float fx = x - x0;
    const float fy = y - y0;
    const int W = static_cast<int>(e.w), H = static_cast<int>(e.h);
    int i0 = static_cast<int>(x0) % W;
    if (i0 < 0) i0 += W;
    int i1 = i0 + 1;
    if (i1 >= W) i1 -= W;
    int j0 = static_cast<int>(y0);
    int j1 = j0 + 1;
    j0 = j0 < 0 ? 0 : (j0 >= H ? H - 1 : j0);
    j1 = j1 < 0 ? 0 : (j1 >= H ? H - 1 : j1);
    const f3 t00 = env_texel(e, i0, j0), t10 = env_texel(e, i1, j0);
    const f3 t01 = env_texel(e, i0, j1), t11 = env_texel(e, i1, j1);
    const f3 a = f3{lerpf(t00.x, t10.x, fx), lerpf(t00.y, t10.y, fx), lerpf(t00.z, t10.z, fx)};
    const f3 b = f3{lerpf(t01.x, t11.x, fx), lerpf(t01.y, t11.y, fx), lerpf(t01.z, t11.z, fx)};
    return f3{lerpf(a.x, b.x, fy), lerpf(a.y, b.y, fy), lerpf(a.z, b.z, fy)};
}

VP_HD bool env_has_table(const EnvView& e) { return e.alias != nullptr && e.total > 0.0f && e.w != 0 && e.h != 0; }

// Solid-angle pdf of env importance sampling (sampling.cuh:218-246).
VP_HD float env_pdf(const EnvView& e, f3 wo) {
    if (!env_has_table(e)) return kOneOverFourPi;
    const int W = static_cast<int>(e.w), H = static_cast<int>(e.h);
    const float theta = atan2f_(wo.z, wo.x);
    const float polar = acosf_(clampf(wo.y, -1.0f, 1.0f));
    const float un = fmaf_(theta, kOneOverTwoPi, 0.5f);
    const float vn = polar * kOneOverPi;
    int u = static_cast<int>(un * static_cast<float>(W));
    int v = static_cast<int>(vn * static_cast<float>(H));
    u = u < 0 ? 0 : (u >= W ? W - 1 : u);
    v = v < 0 ? 0 : (v >= H ? H - 1 : v);
    const float joint = e.joint[v * W + u];
    const float sp = sinf_(polar);
    if (joint <= 0.0f || sp <= 0.0f) return 0.0f;
    const float p_uv = (joint * rcpf(e.total)) * static_cast<float>(W * H);
    const float jac = 2.0f * kPi * kPi * sp;
    return p_uv * rcpf(jac);
}

// Importance-sample a direction from the texel distribution joint/total with
// a continuous position inside the texel (sampling.cuh:173-216), the texel
// chosen through the alias table.  Consumes 4 uniforms (2 without a table).
VP_HD f3 env_sample(const EnvView& e, Pcg32& rng, float* pdf) {
    const float ux = next_uniform(rng);
    const float uy = next_uniform(rng);
    if (!env_has_table(e)) {
        const float z = fmaf_(-2.0f, ux, 1.0f);
        const float r = sqrtf_(maxf(0.0f, fmaf_(-z, z, 1.0f)));
        float sp, cp;
        sincosf_(kTwoPi * uy, &sp, &cp);
        *pdf = kOneOverFourPi;
        return f3{r * cp, r * sp, z};
    }
    const uint32_t n = e.w * e.h;
    uint32_t col = static_cast<uint32_t>(ux * static_cast<float>(n));
    if (col >= n) col = n - 1;
    const AliasEntry ae = e.alias[col];
    const uint32_t texel = uy < ae.prob ? col : ae.alias;
    const uint32_t v = texel / e.w;
    const uint32_t u = texel - v * e.w;
    const float jx = next_uniform(rng);
    const float jy = next_uniform(rng);
    const float un = (static_cast<float>(u) + jx) * rcpf(static_cast<float>(e.w));
    const float vn = (static_cast<float>(v) + jy) * rcpf(static_cast<float>(e.h));
    const float theta = fmaf_(2.0f * kPi, un, -kPi);
    const float polar = vn * kPi;
    float sp, cpol, st, ct;
    sincosf_(polar, &sp, &cpol);
    sincosf_(theta, &st, &ct);
    const f3 wo = f3{sp * ct, cpol, sp * st};
    const float joint = e.joint[texel];
    const float p_uv = (joint * rcpf(e.total)) * static_cast<float>(e.w * e.h);
    const float jac = 2.0f * kPi * kPi * sp;
    *pdf = sp > 0.0f ? p_uv * rcpf(jac) : 0.0f;
    return wo;
}

// ---------------------------------------------------------------------------
// Camera (device/params/camera.h:17-57)
// ---------------------------------------------------------------------------
struct CameraPod {
    float eye[3];
    float p00[3];  // pixel (0,0) centre relative to eye
    float du[3];
    float dv[3];
    float view_dir[3];
    int orthographic;
};

VP_HD void camera_ray(const CameraPod& c, uint32_t px, uint32_t py, float jx, float jy, f3* o, f3* d) {
    const float x = static_cast<float>(px) + jx;
    const float y = static_cast<float>(py) + jy;
    const f3 du = f3{c.du[0], c.du[1], c.du[2]};
    const f3 dv = f3{c.dv[0], c.dv[1], c.dv[2]};
    const f3 p00 = f3{c.p00[0], c.p00[1], c.p00[2]};
    const f3 eye = f3{c.eye[0], c.eye[1], c.eye[2]};
    const f3 r = fma3(x, du, fma3(y, dv, p00));
    if (c.orthographic) {
        *o = eye + r;
        *d = normalize(f3{c.view_dir[0], c.view_dir[1], c.view_dir[2]});
    } else {
        *o = eye;
        *d = normalize(r);
    }
}

// Gaussian pixel-filter jitter, clamped to +-2 px (random.cuh:64-72).
VP_HD void gaussian_jitter(Pcg32& rng, float stddev, float* jx, float* jy) {
    const float u1 = maxf(next_uniform(rng), 1e-7f);
    const float u2 = next_uniform(rng);
    const float r = stddev * sqrtf_(-2.0f * logf_(u1));
    float s, c;
    sincosf_(kTwoPi * u2, &s, &c);
    *jx = clampf(r * c, -2.0f, 2.0f);
    *jy = clampf(r * s, -2.0f, 2.0f);
}

}  // namespace vp
