// vp_device.h — device building blocks shared by the render kernels:
// record loads, the stackless BVH walks (plain and with the LDS leaf queue),
// fixed-point optical depth, shadow-ray transmittance, event counters and
// the job lookup.  Included by vp_render.hip (megakernel, test kernels) and
// vp_wavefront.hip (wavefront pipeline).
#pragma once

#include "vp_internal.h"

namespace vp {
namespace {

constexpr float kTauInvScale = 1.0f / 1099511627776.0f;

__device__ __forceinline__ f3 safe_inv(f3 d) {
    const float x = fabsf_(d.x) < 1e-20f ? copysignf_(1e-20f, d.x) : d.x;
    const float y = fabsf_(d.y) < 1e-20f ? copysignf_(1e-20f, d.y) : d.y;
    const float z = fabsf_(d.z) < 1e-20f ? copysignf_(1e-20f, d.z) : d.z;
    return f3{1.0f / x, 1.0f / y, 1.0f / z};
}

// Scene arrays are reached through pointers stored in device memory (the job
// table), which the compiler can only treat as generic; they always point to
// global memory, and saying so turns flat loads into global loads (which do
// not wait on the LDS counter).
// (The host pass of the compiler only parses these device functions; the
// qualifier is applied in the device pass.)
#if defined(__HIP_DEVICE_COMPILE__)
#define VP_GLOBAL_AS __attribute__((address_space(1)))
#else
#define VP_GLOBAL_AS
#endif
template <typename T>
__device__ __forceinline__ const VP_GLOBAL_AS T* as_global(const T* p) {
    return (const VP_GLOBAL_AS T*)(p);
}

__device__ __forceinline__ PrimRecord load_prim(const PrimRecord* prims, uint32_t i) {
    const auto* p = as_global(reinterpret_cast<const float4*>(prims + i));
    const float4 a = p[0], b = p[1], c = p[2], d = p[3];
    PrimRecord r;
    r.cx = a.x; r.cy = a.y; r.cz = a.z; r.qw = a.w;
    r.qx = b.x; r.qy = b.y; r.qz = b.z; r.rsx = b.w;
    r.rsy = c.x; r.rsz = c.y; r.dnf = c.z; r.icf = c.w;
    r.ar = d.x; r.ag = d.y; r.ab = d.z; r.mass = d.w;
    return r;
}

// number of set bits of m below the calling lane
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Hand-off point for data exchanged between lanes of one wavefront through
// LDS (queues, compaction, per-ray tables, steal slots): the wavefront-scope
// release/acquire pair keeps the compiler from moving LDS accesses across it
// (a lane's reads of entries other lanes write, and the reverse), and the
// wave barrier keeps the active lanes converged there.  The hardware already
// executes a wavefront's LDS operations in order.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- walk-tree traversal (tree layout and Walk4 in vp_walk.h) ----

// Children of node n that the ray crosses within [0, tmax], as a 4-bit mask.
// Slab distances are fma(bound, 1/d, -o/d).  Which boxes a walk enters only
// decides what is visited (leaf boxes are conservative and the test is
// monotone in the box), never a result.
__device__ __forceinline__ bool slab_hit(float lx, float ly, float lz, float hx, float hy, float hz, f3 inv, f3 oi,
                                         float tmax) {
    const float tx0 = fmaf_(lx, inv.x, -oi.x), tx1 = fmaf_(hx, inv.x, -oi.x);
    const float ty0 = fmaf_(ly, inv.y, -oi.y), ty1 = fmaf_(hy, inv.y, -oi.y);
    const float tz0 = fmaf_(lz, inv.z, -oi.z), tz1 = fmaf_(hz, inv.z, -oi.z);
    const float tnear = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    const float tfar = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    return tnear <= tfar;
}
__device__ __forceinline__ uint32_t w4_ray_mask(const float4* w4, uint32_t n, f3 inv, f3 oi, float tmax) {
    const auto* r = as_global(w4) + static_cast<size_t>(kW4RecordF4) * n;
    const float4 lx = r[0], ly = r[1], lz = r[2], hx = r[3], hy = r[4], hz = r[5];
    return (slab_hit(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, inv, oi, tmax) ? 1u : 0u) |
           (slab_hit(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, inv, oi, tmax) ? 2u : 0u) |
           (slab_hit(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, inv, oi, tmax) ? 4u : 0u) |
           (slab_hit(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, inv, oi, tmax) ? 8u : 0u);
}

// The same mask with the near and far bound of each axis picked by the sign
// of the direction component, through the load address: since t is monotone
// in the bound, the near distance is exactly min(t(min), t(max)) and needs
// no min/max per axis.  A third fewer instructions per node, but six 64-bit
// addresses: faster where registers are spare (the wavefront walk kernels,
// +5 % on the scattering workload; the megakernel at 4 waves per SIMD),
// slower where they are not (the megakernel at 6 waves per SIMD: -20 %).
__device__ __forceinline__ uint32_t w4_ray_mask_nf(const float4* w4, uint32_t n, f3 inv, f3 oi, float tmax) {
    // byte offsets of the near bounds (min.x at 0, min.y 16, min.z 32,
    // max.x 48, max.y 64, max.z 80)
    const uint32_t fx = (__float_as_uint(inv.x) >> 31) * 48u;
    const uint32_t fy = (__float_as_uint(inv.y) >> 31) * 48u;
    const uint32_t fz = (__float_as_uint(inv.z) >> 31) * 48u;
    const auto* rec = reinterpret_cast<const VP_GLOBAL_AS char*>(as_global(w4) + static_cast<size_t>(kW4RecordF4) * n);
    const float4 nx = *reinterpret_cast<const VP_GLOBAL_AS float4*>(rec + fx);
    const float4 ny = *reinterpret_cast<const VP_GLOBAL_AS float4*>(rec + 16u + fy);
    const float4 nz = *reinterpret_cast<const VP_GLOBAL_AS float4*>(rec + 32u + fz);
    const float4 gx = *reinterpret_cast<const VP_GLOBAL_AS float4*>(rec + 48u - fx);
    const float4 gy = *reinterpret_cast<const VP_GLOBAL_AS float4*>(rec + 64u - fy);
    const float4 gz = *reinterpret_cast<const VP_GLOBAL_AS float4*>(rec + 80u - fz);
    auto hit = [&](float ax, float ay, float az, float bx, float by, float bz) {
        const float tn = fmaxf(fmaxf(fmaf_(ax, inv.x, -oi.x), fmaf_(ay, inv.y, -oi.y)),
                               fmaxf(fmaf_(az, inv.z, -oi.z), 0.0f));
        // (the clamp to tmax also rejects the +inf boxes of empty children
        // for a direction with all components positive)
        const float tf = fminf(fminf(fmaf_(bx, inv.x, -oi.x), fmaf_(by, inv.y, -oi.y)),
                               fminf(fmaf_(bz, inv.z, -oi.z), tmax));
        return tn <= tf;
    };
    return (hit(nx.x, ny.x, nz.x, gx.x, gy.x, gz.x) ? 1u : 0u) | (hit(nx.y, ny.y, nz.y, gx.y, gy.y, gz.y) ? 2u : 0u) |
           (hit(nx.z, ny.z, nz.z, gx.z, gy.z, gz.z) ? 4u : 0u) | (hit(nx.w, ny.w, nz.w, gx.w, gy.w, gz.w) ? 8u : 0u);
}

// ---- the same tests on 16-bit records (layout in vp_walk.h) ----
// The low and high binary16 of word w as float (exact).  As an FMA operand
// the conversion folds into v_fma_mix_f32, so a slab distance costs one
// instruction, as on 32-bit records, for half the bytes per node.
__device__ __forceinline__ float h16lo(uint32_t w) {
    return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w)));
}
__device__ __forceinline__ float h16hi(uint32_t w) {
    return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w >> 16)));
}
// record n of a 16-bit tree (kW4hRecordB = 4 float4)
__device__ __forceinline__ const VP_GLOBAL_AS uint4* w4h_record(const float4* w4, uint32_t n) {
    return reinterpret_cast<const VP_GLOBAL_AS uint4*>(as_global(w4) + 4u * static_cast<size_t>(n));
}
__device__ __forceinline__ uint32_t w4h_ray_mask(const float4* w4, uint32_t n, f3 inv, f3 oi, float tmax) {
    const auto* r = w4h_record(w4, n);
    // a = min.x (children 0,1 | 2,3), min.y; b = min.z, max.x; c = max.y, max.z
    const uint4 a = r[0], b = r[1], c = r[2];
    return (slab_hit(h16lo(a.x), h16lo(a.z), h16lo(b.x), h16lo(b.z), h16lo(c.x), h16lo(c.z), inv, oi, tmax) ? 1u
                                                                                                             : 0u) |
           (slab_hit(h16hi(a.x), h16hi(a.z), h16hi(b.x), h16hi(b.z), h16hi(c.x), h16hi(c.z), inv, oi, tmax) ? 2u
                                                                                                             : 0u) |
           (slab_hit(h16lo(a.y), h16lo(a.w), h16lo(b.y), h16lo(b.w), h16lo(c.y), h16lo(c.w), inv, oi, tmax) ? 4u
                                                                                                             : 0u) |
           (slab_hit(h16hi(a.y), h16hi(a.w), h16hi(b.y), h16hi(b.w), h16hi(c.y), h16hi(c.w), inv, oi, tmax) ? 8u
                                                                                                             : 0u);
}
// near/far form: the bounds of one axis and side are 8 bytes (min.x at 0,
// min.y 8, min.z 16, max.x 24, max.y 32, max.z 40)
__device__ __forceinline__ uint32_t w4h_ray_mask_nf(const float4* w4, uint32_t n, f3 inv, f3 oi, float tmax) {
    const uint32_t fx = (__float_as_uint(inv.x) >> 31) * 24u;
    const uint32_t fy = (__float_as_uint(inv.y) >> 31) * 24u;
    const uint32_t fz = (__float_as_uint(inv.z) >> 31) * 24u;
    const auto* rec = reinterpret_cast<const VP_GLOBAL_AS char*>(w4h_record(w4, n));
    const uint2 nx = *reinterpret_cast<const VP_GLOBAL_AS uint2*>(rec + fx);
    const uint2 ny = *reinterpret_cast<const VP_GLOBAL_AS uint2*>(rec + 8u + fy);
    const uint2 nz = *reinterpret_cast<const VP_GLOBAL_AS uint2*>(rec + 16u + fz);
    const uint2 gx = *reinterpret_cast<const VP_GLOBAL_AS uint2*>(rec + 24u - fx);
    const uint2 gy = *reinterpret_cast<const VP_GLOBAL_AS uint2*>(rec + 32u - fy);
    const uint2 gz = *reinterpret_cast<const VP_GLOBAL_AS uint2*>(rec + 40u - fz);
    auto hit = [&](float ax, float ay, float az, float bx, float by, float bz) {
        const float tn = fmaxf(fmaxf(fmaf_(ax, inv.x, -oi.x), fmaf_(ay, inv.y, -oi.y)),
                               fmaxf(fmaf_(az, inv.z, -oi.z), 0.0f));
        const float tf = fminf(fminf(fmaf_(bx, inv.x, -oi.x), fmaf_(by, inv.y, -oi.y)),
                               fminf(fmaf_(bz, inv.z, -oi.z), tmax));
        return tn <= tf;
    };
    return (hit(h16lo(nx.x), h16lo(ny.x), h16lo(nz.x), h16lo(gx.x), h16lo(gy.x), h16lo(gz.x)) ? 1u : 0u) |
           (hit(h16hi(nx.x), h16hi(ny.x), h16hi(nz.x), h16hi(gx.x), h16hi(gy.x), h16hi(gz.x)) ? 2u : 0u) |
           (hit(h16lo(nx.y), h16lo(ny.y), h16lo(nz.y), h16lo(gx.y), h16lo(gy.y), h16lo(gz.y)) ? 4u : 0u) |
           (hit(h16hi(nx.y), h16hi(ny.y), h16hi(nz.y), h16hi(gx.y), h16hi(gy.y), h16hi(gz.y)) ? 8u : 0u);
}

// The ray tests on either record form (half = the scene's w4_half).
__device__ __forceinline__ uint32_t ray_mask(const float4* w4, uint32_t half, uint32_t n, f3 inv, f3 oi,
                                             float tmax) {
    return half ? w4h_ray_mask(w4, n, inv, oi, tmax) : w4_ray_mask(w4, n, inv, oi, tmax);
}
__device__ __forceinline__ uint32_t ray_mask_nf(const float4* w4, uint32_t half, uint32_t n, f3 inv, f3 oi,
                                                float tmax) {
    return half ? w4h_ray_mask_nf(w4, n, inv, oi, tmax) : w4_ray_mask_nf(w4, n, inv, oi, tmax);
}

// The node test of the megakernel's shared-queue walks: 1 selects the
// near/far form (vp_render.hip sets it).
#ifndef VP_MEGA_NEAR_FAR
#define VP_MEGA_NEAR_FAR 0
#endif
__device__ __forceinline__ uint32_t mega_ray_mask(const float4* w4, uint32_t half, uint32_t n, f3 inv, f3 oi,
                                                  float tmax) {
#if VP_MEGA_NEAR_FAR
    return ray_mask_nf(w4, half, n, inv, oi, tmax);
#else
    return ray_mask(w4, half, n, inv, oi, tmax);
#endif
}

// children of node n whose box contains x
__device__ __forceinline__ uint32_t point_mask(const float4* w4, uint32_t half, uint32_t n, f3 x) {
    auto in = [&](float l0, float l1, float l2, float h0, float h1, float h2) {
        return x.x >= l0 && x.x <= h0 && x.y >= l1 && x.y <= h1 && x.z >= l2 && x.z <= h2;
    };
    if (half) {
        const auto* r = w4h_record(w4, n);
        const uint4 a = r[0], b = r[1], c = r[2];
        return (in(h16lo(a.x), h16lo(a.z), h16lo(b.x), h16lo(b.z), h16lo(c.x), h16lo(c.z)) ? 1u : 0u) |
               (in(h16hi(a.x), h16hi(a.z), h16hi(b.x), h16hi(b.z), h16hi(c.x), h16hi(c.z)) ? 2u : 0u) |
               (in(h16lo(a.y), h16lo(a.w), h16lo(b.y), h16lo(b.w), h16lo(c.y), h16lo(c.w)) ? 4u : 0u) |
               (in(h16hi(a.y), h16hi(a.w), h16hi(b.y), h16hi(b.w), h16hi(c.y), h16hi(c.w)) ? 8u : 0u);
    }
    const auto* r = as_global(w4) + static_cast<size_t>(kW4RecordF4) * n;
    const float4 lx = r[0], ly = r[1], lz = r[2], hx = r[3], hy = r[4], hz = r[5];
    return (in(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x) ? 1u : 0u) | (in(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y) ? 2u : 0u) |
           (in(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z) ? 4u : 0u) | (in(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w) ? 8u : 0u);
}

// Walk over [0, tmax]; leaf(prim) returns false to stop.  tmax may be
// lowered by the caller between leaves.
template <typename Leaf>
__device__ __forceinline__ void walk_ray(const SceneView& s, f3 o, f3 inv, const float& tmax, uint32_t& visits,
                                         Leaf&& leaf) {
    const f3 oi = o * inv;
    Walk4 w = walk4_begin();
    while (w.n != kWalkEnd) {
        ++visits;
        const uint32_t m = ray_mask(s.w4, s.w4_half, w.n, inv, oi, tmax);
        uint32_t first;
        for (uint32_t lm = walk4_advance(w, s.w4_levels, m, first); lm != 0u; lm &= lm - 1u)
            if (!leaf(first + static_cast<uint32_t>(__builtin_ctz(lm)))) return;
    }
}

// Walk visiting every leaf whose box contains x, in ascending stored-primitive
// order.
template <typename Leaf>
__device__ __forceinline__ void walk_point(const SceneView& s, f3 x, uint32_t& visits, Leaf&& leaf) {
    Walk4 w = walk4_begin();
    while (w.n != kWalkEnd) {
        ++visits;
        const uint32_t m = point_mask(s.w4, s.w4_half, w.n, x);
        uint32_t first;
        for (uint32_t lm = walk4_advance(w, s.w4_levels, m, first); lm != 0u; lm &= lm - 1u)
            leaf(first + static_cast<uint32_t>(__builtin_ctz(lm)));
    }
}

// Leaf queue depth per lane.  The walks below postpone leaves into a per-lane
// LDS queue (row `thread` of a [blockDim][kLeafQ + 1] array: the odd row
// length puts the 64 lanes' entries k in distinct banks) and keep walking
// until every lane of the wavefront has a full queue or has finished; then all
// lanes process their queued leaves together.  Node steps stay convergent and
// the expensive leaf work runs with most lanes active (the "while-while"
// scheme of Aila and Laine 2009, with speculative postponing).
#ifndef VP_LEAF_Q
#define VP_LEAF_Q 16
#endif
constexpr uint32_t kLeafQ = VP_LEAF_Q;
constexpr uint32_t kLeafRow = kLeafQ + 1;

// Wave-shared leaf queue.  The LDS of a wavefront holds either its threads'
// queue rows or one queue for the whole wavefront (the two are used at
// different times): node phases append crossed leaves with their owner lane,
// and in the leaf phase every active lane takes the next entry, whichever
// lane's ray it belongs to.  Leaf work then runs at full width however
// unevenly the lanes' walks found leaves.  Per-ray results are combined with
// LDS atomics whose outcome does not depend on order (fixed-point sums, and
// minima).
#ifndef VP_WAVE_Q
#define VP_WAVE_Q 224
#endif
constexpr uint32_t kWaveQ = VP_WAVE_Q;
// A node phase of the shared-queue walks ends early, to share out the queued
// leaves, once the queue holds kWaveQ / 4 entries and fewer than
// VP_SHARED_STEP_MIN lanes still step.
// Shadow leaf phases in two passes (entry tests, then the chords of the
// entered leaves), each at full width: +2 % on the scattering workload.
#ifndef VP_SHADOW_TWO_STAGE
#define VP_SHADOW_TWO_STAGE 1
#endif
#ifndef VP_SHARED_STEP_MIN
#define VP_SHARED_STEP_MIN VP_WALK_MIN
#endif
// (the queue fill, as a fraction 1 / VP_SHARED_END_DIV, from which that rule applies)
#ifndef VP_SHARED_END_DIV
#define VP_SHARED_END_DIV 4u
#endif
// Rays are numbered per wavefront: ray k of a one-ray-per-lane walk belongs
// to lane k; the MIS shadow rays of lane l are rays 2l and 2l + 1.
struct WaveQueue {
    uint32_t* prim;                // [kWaveQ] queued leaf
    float* tj;                     // [kWaveQ] entry distances of the leaves that passed the entry test
    uint8_t* owner;                // [kWaveQ] the ray each leaf belongs to
    float* ro;                     // [64 * 3] ray origin of each lane
    const PrimRecord** prims;      // [64] primitive array of each lane's scene (lanes of one
                                   //      wavefront may run jobs over different scenes)
    float* rd;                     // [128 * 3] ray directions
    unsigned long long* acc;       // [128] fixed-point optical depth of each ray
    // MIS shadow-ray pool
    const float4** nodes;          // [64] walk-tree records of each lane's scene
    uint32_t* nn;                  // [64] its level count (or, in the wavefront argmin
                                   //      kernel, the pool slot of each lane's path)
    uint8_t* list;                 // [128] the rays still to be walked
    uint32_t* lim;                 // [128] depth at which each ray's walk may stop, >> 16 (shadow roulette)
    // argmin walks (one ray per lane); they lie over rd[64..128) and acc[64..128)
    unsigned long long* key;       // [64] path key of each lane
    uint32_t* ts;                  // [64] running minimum of each lane's argmin, as float bits
    uint32_t* nc;                  // [64] candidates recorded for each lane's argmin
    uint32_t* meta;                // [64] bounce | direct-term flag << 31
};
constexpr uint32_t kWaveQueueBytes = kWaveQ * 9 + 64 * 12 + 64 * 8 + 128 * 12 + 128 * 8 + 64 * 8 + 64 * 4 + 128 + 128 * 4;
static_assert(kWaveQ % 8 == 0, "queue arrays keep the per-lane arrays 8-byte aligned");
static_assert(64 * (8 + 4) <= 64 * 12 && 64 * (4 + 4) <= 64 * 8, "argmin arrays fit over the upper rays");
// LDS bytes per wavefront: the larger of the two uses, in 16-byte units
constexpr uint32_t kWaveBytes =
    ((64 * kLeafRow * 4 > kWaveQueueBytes ? 64 * kLeafRow * 4 : kWaveQueueBytes) + 15) / 16 * 16;

// the calling wavefront's queue
__device__ __forceinline__ WaveQueue wave_queue(uint32_t* lds) {
    uint8_t* base = reinterpret_cast<uint8_t*>(lds) + (threadIdx.x >> 6) * kWaveBytes;
    WaveQueue q;
    q.prim = reinterpret_cast<uint32_t*>(base);
    q.tj = reinterpret_cast<float*>(base + 4 * kWaveQ);
    q.owner = base + 8 * kWaveQ;
    q.ro = reinterpret_cast<float*>(base + 9 * kWaveQ);
    q.prims = reinterpret_cast<const PrimRecord**>(q.ro + 64 * 3);
    q.rd = reinterpret_cast<float*>(q.prims + 64);
    q.acc = reinterpret_cast<unsigned long long*>(q.rd + 128 * 3);
    q.nodes = reinterpret_cast<const float4**>(q.acc + 128);
    q.nn = reinterpret_cast<uint32_t*>(q.nodes + 64);
    q.list = reinterpret_cast<uint8_t*>(q.nn + 64);
    q.lim = reinterpret_cast<uint32_t*>(q.list + 128);
    q.key = reinterpret_cast<unsigned long long*>(q.rd + 64 * 3);
    q.ts = reinterpret_cast<uint32_t*>(q.key + 64);
    q.nc = reinterpret_cast<uint32_t*>(q.acc + 64);
    q.meta = q.nc + 64;
    return q;
}

__device__ __forceinline__ f3 lds_load3(const float* p) { return f3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void lds_store3(float* p, f3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

// Node phase of the per-lane ray walks: a lane steps while its leaf queue has
// room for a step (up to four leaf blocks) and its walk has nodes left.  With
// VP_WALK_MIN > 0 the loop runs for the whole wavefront, which leaves it early
// once fewer than VP_WALK_MIN lanes are stepping and some lane's queue is
// full, so full lanes do not wait for the last few; the stepping lanes resume
// after the leaf phase.  A node phase that starts with every queue empty can
// only end early after some lane has queued leaves, so walks always progress.
// Every query consumes leaves order-independently, so this changes only the
// schedule.
#ifndef VP_WALK_MIN
#define VP_WALK_MIN 8
#endif

// room in a lane's queue for one more node step
__device__ __forceinline__ bool queue_room(uint32_t cnt) { return cnt + 4u <= kLeafQ; }
// ... for one more step of a walk of leaf blocks (up to four blocks)
static_assert(4u * kLeafBlock <= kLeafQ, "a lane's queue must hold one step of leaf blocks");
__device__ __forceinline__ bool block_room(uint32_t cnt) { return cnt + 4u * kLeafBlock <= kLeafQ; }

// true while this lane (VP_WALK_MIN == 0) or the wavefront stays in the node
// phase (room: the lane's queue has room for its next step)
__device__ __forceinline__ bool node_phase(bool step, bool room) {
#if VP_WALK_MIN > 0
    const unsigned long long sm = __builtin_amdgcn_ballot_w64(step);
    if (sm == 0) return false;
    return __popcll(sm) >= VP_WALK_MIN || __builtin_amdgcn_ballot_w64(!room) == 0;
#else
    (void)room;
    return step;
#endif
}

struct LeafQueue {
    uint32_t* row;  // &lds[thread][0]
    __device__ __forceinline__ uint32_t& at(uint32_t k) const { return row[k]; }
};

// the calling thread's queue in a dynamic LDS array of leaf_queue_bytes(blockDim.x)
__device__ __forceinline__ LeafQueue leaf_queue(uint32_t* lds) {
    return LeafQueue{lds + (threadIdx.x >> 6) * (kWaveBytes / 4) + (threadIdx.x & 63u) * kLeafRow};
}
inline size_t leaf_queue_bytes(uint32_t block) { return static_cast<size_t>(kWaveBytes) * ((block + 63) / 64); }

// Appends the leaves of mask lm (over first .. first + 3) to a lane's queue at
// cnt, which has room for four.  Slot cnt + (set bits of lm below k) is
// written for every k, set or not, and a later set bit overwrites it, so the
// step has no branches; the writes stay below cnt + 4.
__device__ __forceinline__ void queue_leaves(const LeafQueue& q, uint32_t& cnt, uint32_t lm, uint32_t first) {
    for (uint32_t k = 0; k < 4u; ++k) q.at(cnt + static_cast<uint32_t>(__popc(lm & ((1u << k) - 1u)))) = first + k;
    cnt += static_cast<uint32_t>(__popc(lm));
}

// One node step of a ray walk of leaf blocks (walk4_block_levels) over
// [0, tmax]: the leaves of the crossed blocks are queued, in ascending order,
// and the walk moves on.  The same walk as the wave-shared ones, so the
// counting build counts the node steps and leaf tests the timed kernel makes.
__device__ __forceinline__ void ray_node_step(const SceneView& s, f3 inv, f3 oi, float tmax, Walk4& w, uint32_t& cnt,
                                              const LeafQueue& q) {
    const uint32_t m = ray_mask(s.w4, s.w4_half, w.n, inv, oi, tmax);
    uint32_t first;
    const uint32_t bm = walk4_advance(w, walk4_block_levels(s.w4_levels), m, first);
    for (uint32_t b = bm; b != 0u; b &= b - 1u) {
        const uint32_t l0 = kLeafBlock * (first + static_cast<uint32_t>(__builtin_ctz(b)));
        for (uint32_t k = 0; k < kLeafBlock; ++k) q.at(cnt++) = l0 + k;
    }
}

template <typename Leaf>
__device__ __forceinline__ void walk_ray_q(const SceneView& s, f3 o, f3 inv, const float& tmax, uint32_t& visits,
                                           const LeafQueue& q, Leaf&& leaf) {
    const f3 oi = o * inv;
    Walk4 w = walk4_begin();
    while (true) {
        uint32_t cnt = 0;
        while (true) {
            const bool room = block_room(cnt);
            const bool step = w.n != kWalkEnd && room;
            if (!node_phase(step, room)) break;
            if (!step) continue;
            ++visits;
            ray_node_step(s, inv, oi, tmax, w, cnt, q);
        }
        for (uint32_t k = 0; k < cnt; ++k)
            if (!leaf(q.at(k))) return;
        if (w.n == kWalkEnd) return;
    }
}

template <typename Leaf>
__device__ __forceinline__ void walk_point_q(const SceneView& s, f3 x, uint32_t& visits, const LeafQueue& q,
                                             Leaf&& leaf) {
    Walk4 w = walk4_begin();
    while (true) {
        uint32_t cnt = 0;
        while (w.n != kWalkEnd && queue_room(cnt)) {
            ++visits;
            const uint32_t m = point_mask(s.w4, s.w4_half, w.n, x);
            uint32_t first;
            const uint32_t lm = walk4_advance(w, s.w4_levels, m, first);
            if (lm != 0u) queue_leaves(q, cnt, lm, first);
        }
        for (uint32_t k = 0; k < cnt; ++k) leaf(q.at(k));
        if (w.n == kWalkEnd) return;
    }
}

// Appends the leaves of the calling lanes' crossed leaf blocks (bm: a mask
// over the blocks first .. first + 3 of a walk of leaf blocks,
// walk4_block_levels; block c holds the stored primitives kLeafBlock c ..
// kLeafBlock (c + 1) - 1, records past the scene's last primitive being zero
// records, which the leaf tests reject) to the wave queue at count, in lane
// order, as far as they fit in kWaveQ.  Returns false for a lane whose leaves
// did not fit: its node step is to be repeated after the leaf phase.  The
// lanes that fit form a prefix in lane order, so the queue has no gaps; count
// advances by the leaves written and stays a multiple of 4, so each group of
// four leaves is one 16-byte LDS write of ids and one 4-byte write of owners.
// Every lane of the wavefront that is active here must call it.
__device__ __forceinline__ bool wave_push_blocks(const WaveQueue& wq, uint32_t& count, uint32_t bm, uint32_t first,
                                                 uint32_t owner) {
    static_assert(kWaveQ % 4 == 0, "blocks are queued four leaves at a time");
    const uint32_t nb = static_cast<uint32_t>(__popc(bm));  // 0 .. 4 blocks
    const unsigned long long b0 = __builtin_amdgcn_ballot_w64((nb & 1u) != 0u);
    const unsigned long long b1 = __builtin_amdgcn_ballot_w64((nb & 2u) != 0u);
    const unsigned long long b2 = __builtin_amdgcn_ballot_w64((nb & 4u) != 0u);
    if ((b0 | b1 | b2) == 0ull) return true;  // no blocks this step
    const uint32_t pos = count + kLeafBlock * (lanes_below(b0) + 2u * lanes_below(b1) + 4u * lanes_below(b2));
    const bool fits = nb == 0u || pos + kLeafBlock * nb <= kWaveQ;
    const unsigned long long f = __builtin_amdgcn_ballot_w64(fits);
    if (fits && nb != 0u) {
        uint32_t p = pos;
        for (uint32_t b = bm; b != 0u; b &= b - 1u) {
            const uint32_t l0 = kLeafBlock * (first + static_cast<uint32_t>(__builtin_ctz(b)));
#pragma unroll
            for (uint32_t k = 0; k < kLeafBlock; k += 4u, p += 4u) {
                *reinterpret_cast<uint4*>(wq.prim + p) = make_uint4(l0 + k, l0 + k + 1u, l0 + k + 2u, l0 + k + 3u);
                *reinterpret_cast<uint32_t*>(wq.owner + p) = owner * 0x01010101u;
            }
        }
    }
    count += kLeafBlock * static_cast<uint32_t>(__popcll(b0 & f) + 2 * __popcll(b1 & f) + 4 * __popcll(b2 & f));
    return fits;
}

// Add one chord to the fixed-point optical depth; true once the total
// reaches 88 (exp(-88) underflows), which is order independent because every
// term is non-negative.
__device__ __forceinline__ bool tau_add(uint64_t& acc, float od) {
    const float c = od > 0.0f ? od : 0.0f;
    if (c >= kMaxTau) return true;
    acc += static_cast<uint64_t>(c * kTauScaleF);
    return acc >= kTauLimitFixed;
}

__device__ __forceinline__ float tau_to_T(uint64_t acc) {
    const float tau = static_cast<float>(acc) * kTauInvScale;
    return tau >= kMaxTau ? 0.0f : expf_(-tau);
}
// Transmittance of a fixed-point depth (0 once saturated).
__device__ __forceinline__ float depth_T(uint64_t acc) { return acc >= kTauLimitFixed ? 0.0f : tau_to_T(acc); }

struct EvCount {
    uint32_t v[kNumCounters];
};

template <bool kCount>
struct Ev {
    EvCount c;
    __device__ __forceinline__ void reset() {
        if constexpr (kCount)
            for (int i = 0; i < kNumCounters; ++i) c.v[i] = 0;
    }
    __device__ __forceinline__ void add(int slot, uint32_t n = 1) {
        if constexpr (kCount) c.v[slot] += n;
    }
    __device__ __forceinline__ void flush(unsigned long long* g) {
        if constexpr (kCount) {
            for (int i = kCntPaths; i < kNumCounters; ++i)
                if (c.v[i]) atomicAdd(&g[i], static_cast<unsigned long long>(c.v[i]));
            reset();
        }
    }
};

// compute_transmittance_to_env (sampling.cuh:548-583) with the fused
// any-hit integration (anyhit.cuh:53-66), as a fixed-point optical depth: the
// walk stops once the sum reaches limit (at most the tau = 88 saturation,
// which it returns as kTauLimitFixed).
template <bool kCount>
__device__ uint64_t shadow_depth(const SceneView& s, f3 o, f3 d, const uint32_t* act, uint32_t na, uint32_t stride,
                                 const LeafQueue& lq, Ev<kCount>& ev, uint64_t limit) {
    uint64_t acc = 0;
    if (acc >= limit) return acc;
    for (uint32_t k = 0; k < na; ++k) {
        const PrimRecord pr = load_prim(s.prims, act[k * stride]);
        const InsideChord x = inside_chord(pr, o, d);
        ev.add(kCntShActive);
        if (x.tau > 0.0f) {
            if (tau_add(acc, x.tau)) return kTauLimitFixed;
            if (acc >= limit) return acc;
        }
    }
    const f3 inv = safe_inv(d);
    const float tmax = kTMax;
    uint32_t visits = 0, leaves = 0, sh = 0;
    walk_ray_q(s, o, inv, tmax, visits, lq, [&](uint32_t prim) -> bool {
        const PrimRecord pr = load_prim(s.prims, prim);
        ++leaves;
        const Crossing x = crossing(pr, o, d);
        if (x.t0 < 0.0f) return true;
        ++sh;
        if (tau_add(acc, x.tau)) {
            acc = kTauLimitFixed;
            return false;
        }
        return acc < limit;
    });
    ev.add(kCntNodeVisits, visits);
    ev.add(kCntVisitsShadow, visits);
    ev.add(kCntLeafTests, leaves);
    ev.add(kCntShHit, sh);
    return acc;
}

template <bool kCount>
__device__ float transmittance(const SceneView& s, f3 o, f3 d, const uint32_t* act, uint32_t na, uint32_t stride,
                               const LeafQueue& lq, Ev<kCount>& ev) {
    return depth_T(shadow_depth(s, o, d, act, na, stride, lq, ev, kTauLimitFixed));
}

// one chord in fixed point, as tau_add adds it; a chord of tau >= 88 alone
// adds the saturation limit
__device__ __forceinline__ unsigned long long tau_fixed(float od) {
    const float c = od > 0.0f ? od : 0.0f;
    return c >= kMaxTau ? kTauLimitFixed : static_cast<unsigned long long>(c * kTauScaleF);
}

// Work sharing between the walks of one wavefront, for queries whose result
// does not depend on visiting order.  Lanes without a walk (idle) each take
// the shallowest pending subtree of a running walk (can_give): the donor
// clears that child from its trail and the thief starts a walk at the child
// with an empty trail, which ends once the subtree is done, so every subtree
// is still walked exactly once.  The handoff uses the first 128 words of
// wq.tj, which are free between leaf phases.  Returns true for a lane that
// received a subtree; its ray and walk are set.
// A node phase ends to share out the running walks once at least
// VP_STEAL_IDLE lanes have nothing to step (scattering workload: 8, 16 and 32
// measured 27.7, 28.3 and 28.5 Msamples/s).
#ifndef VP_STEAL_IDLE
#define VP_STEAL_IDLE 32
#endif
// ... and only after at least VP_STEAL_GAP node steps in the phase, so each
// handoff (a leaf phase, the handoff and the lanes' ray setup) is paid for by
// that many steps of the lanes it feeds.
#ifndef VP_STEAL_GAP
#define VP_STEAL_GAP 1u
#endif
__device__ __forceinline__ bool wave_steal(const WaveQueue& wq, bool idle, bool can_give, uint32_t& ray, Walk4& w) {
    const uint64_t pend = can_give ? walk4_pending(w) : 0ull;
    const unsigned long long im = __builtin_amdgcn_ballot_w64(idle);
    const unsigned long long dm = __builtin_amdgcn_ballot_w64(pend != 0ull);
    const uint32_t ni = static_cast<uint32_t>(__popcll(im)), nd = static_cast<uint32_t>(__popcll(dm));
    const uint32_t pairs = ni < nd ? ni : nd;
    if (pairs == 0u) return false;
    uint32_t* slot = reinterpret_cast<uint32_t*>(wq.tj);
    if (pend != 0ull) {
        const uint32_t r = lanes_below(dm);
        if (r < pairs) {
            Walk4 part;
            walk4_split(w, part);
            slot[r] = part.n;
            slot[64 + r] = ray | (part.lvl << 8);
        }
    }
    wave_sync();
    bool got = false;
    if (idle) {
        const uint32_t r = lanes_below(im);
        if (r < pairs) {
            const uint32_t meta = slot[64 + r];
            ray = meta & 0xffu;
            w = Walk4{slot[r], meta >> 8, 0ull};
            got = true;
        }
    }
    return got;
}

// The MIS shadow rays of the calling lanes, with the wave-shared leaf queue
// and a wave-wide ray pool: every lane starts on one ray of the wavefront's
// list (all rays a, then all rays b) and, when its ray is done, takes the
// next one from the list, whichever lane the ray belongs to.  The node
// phases then stay full however the walk lengths are spread, until the list
// runs out.  Each ray's chords are summed in fixed point in LDS (a ray is
// saturated once its sum reaches the tau = 88 limit, in whatever order the
// chords arrive), so Ta and Tb equal two transmittance() calls bit for bit.
// Every lane of the calling wavefront that is active here takes part.
// With two = false only ray a is walked (the one shadow ray of RIS).  Each
// ray's walk may stop once its depth reaches its limit (lim_a, lim_b: the
// tau = 88 saturation, or a shadow-roulette threshold; multiples of 2^16);
// acc_a and acc_b receive the fixed-point depths.
__device__ __forceinline__ void transmittance_pool(const SceneView& s, f3 o, f3 da, f3 db, const uint32_t* act,
                                                   uint32_t na, uint32_t stride, const WaveQueue& wq,
                                                   uint64_t lim_a, uint64_t lim_b, uint64_t& acc_a,
                                                   uint64_t& acc_b, bool two = true) {
    const uint32_t lane = __lane_id();
    lds_store3(wq.ro + 3 * lane, o);
    wq.prims[lane] = s.prims;
    wq.nodes[lane] = s.w4;
    wq.nn[lane] = s.w4_levels | (s.w4_half << 8);
    lds_store3(wq.rd + 3 * (2 * lane), da);
    lds_store3(wq.rd + 3 * (2 * lane + 1), db);
    wq.acc[2 * lane] = 0ull;
    wq.acc[2 * lane + 1] = two ? 0ull : kTauLimitFixed;
    wq.lim[2 * lane] = static_cast<uint32_t>(lim_a >> 16);
    wq.lim[2 * lane + 1] = two ? static_cast<uint32_t>(lim_b >> 16) : 0u;
    // a ray is done once its depth reaches its limit
    auto reached = [&](uint32_t r) -> bool {
        return wq.acc[r] >= (static_cast<unsigned long long>(wq.lim[r]) << 16);
    };
    wave_sync();
    // Chords of the origin set (every member's optical depth from the origin
    // to its exit), for all lanes' rays together: the wavefront's (member,
    // lane) pairs are numbered through a prefix sum of the set sizes and
    // shared out over the lanes, so the loop runs sum(na) / lanes times
    // instead of max(na).  Each chord goes into its ray's fixed-point sum; a
    // sum that passes the tau = 88 limit gives T = 0 however far it passes,
    // as in transmittance().  The owners' set pointers and prefix sums sit in
    // wq.tj and their lane ids in wq.list, both free until the walks start.
    {
        const unsigned long long m0 = __builtin_amdgcn_ballot_w64(true);
        const uint32_t rk = lanes_below(m0), nl = static_cast<uint32_t>(__popcll(m0));
        uint32_t pre = 0, total = 0;
        for (uint32_t bit = 0; bit < 16u; ++bit) {
            const unsigned long long bm = __builtin_amdgcn_ballot_w64(((na >> bit) & 1u) != 0u);
            pre += lanes_below(bm) << bit;
            total += static_cast<uint32_t>(__popcll(bm)) << bit;
        }
        static_assert(kWaveQ * sizeof(float) >= 64 * (sizeof(void*) + sizeof(uint32_t)),
                      "the shared queue must hold the origin-set tables");
        const uint32_t** set_of = reinterpret_cast<const uint32_t**>(wq.tj);  // [64] by rank
        uint32_t* pre_of = reinterpret_cast<uint32_t*>(wq.tj) + 2 * 64;     // [64] by rank
        set_of[rk] = act;
        pre_of[rk] = pre;
        wq.list[rk] = static_cast<uint8_t>(lane);
        wave_sync();
        for (uint32_t g = rk; g < total; g += nl) {
            // the owner: the last rank whose prefix is at most g
            uint32_t lo = 0, hi = nl - 1u;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if (pre_of[mid] <= g) lo = mid;
                else hi = mid - 1u;
            }
            const uint32_t ol = wq.list[lo];
            const PrimRecord pr = load_prim(wq.prims[ol], set_of[lo][(g - pre_of[lo]) * stride]);
            const f3 xo = lds_load3(wq.ro + 3 * ol);
            const f3 dra = lds_load3(wq.rd + 3 * (2 * ol));
            const float ta = inside_chord(pr, xo, dra).tau;
            if (ta > 0.0f) atomicAdd(&wq.acc[2 * ol], tau_fixed(ta));
            if (two) {
                const f3 drb = lds_load3(wq.rd + 3 * (2 * ol + 1));
                const float tb = inside_chord(pr, xo, drb).tau;
                if (tb > 0.0f) atomicAdd(&wq.acc[2 * ol + 1], tau_fixed(tb));
            }
        }
    }
    wave_sync();
    const bool full_a = reached(2 * lane);
    const bool full_b = reached(2 * lane + 1);
    // the rays still to walk: every ray a, then every ray b (a saturated ray
    // is done already)
    const unsigned long long am = __builtin_amdgcn_ballot_w64(true);
    const unsigned long long pa = __builtin_amdgcn_ballot_w64(!full_a);
    const unsigned long long pb = __builtin_amdgcn_ballot_w64(!full_b);
    const uint32_t npa = static_cast<uint32_t>(__popcll(pa));
    const uint32_t nr = npa + static_cast<uint32_t>(__popcll(pb));
    if (!full_a) wq.list[lanes_below(pa)] = static_cast<uint8_t>(2 * lane);
    if (!full_b) wq.list[npa + lanes_below(pb)] = static_cast<uint8_t>(2 * lane + 1);
    wave_sync();
    const uint32_t nact = static_cast<uint32_t>(__popcll(am));
    uint32_t next = nact < nr ? nact : nr;  // list entries handed out
    constexpr uint32_t kNoRay = 0xffu;
    // the ray this lane walks, and its walk state
    const uint32_t rank = lanes_below(am);
    uint32_t ray = rank < nr ? wq.list[rank] : kNoRay;
    const float4* w4 = nullptr;
    uint32_t levels = 0, half = 0;
    Walk4 w = walk4_begin();
    f3 inv = mk(0.0f), oi = mk(0.0f);
    // the walk setup of the lane's ray (any lane may run it: an idle lane
    // reads ray 127's data and does not use it)
    auto setup = [&]() {
        const uint32_t r = ray & 127u, own = r >> 1;
        const f3 ro = lds_load3(wq.ro + 3 * own), rd = lds_load3(wq.rd + 3 * r);
        inv = safe_inv(rd);
        oi = ro * inv;
        w4 = wq.nodes[own];
        levels = walk4_block_levels(wq.nn[own] & 0xffu);
        half = wq.nn[own] >> 8;
    };
    setup();
    while (__builtin_amdgcn_ballot_w64(ray != kNoRay) != 0) {
        // node phase: every lane with a ray steps; crossed leaves are queued
        // for the wavefront.  It ends when no lane steps, when the queue
        // could not take some lane's leaves, or, once a quarter of the queue
        // is filled, when fewer than VP_SHARED_STEP_MIN lanes still step.
        uint32_t count = 0, nsteps = 0;
        while (true) {
            const bool step = ray != kNoRay && w.n != kWalkEnd;
            const unsigned long long sm = __builtin_amdgcn_ballot_w64(step);
            if (sm == 0) break;
            const uint32_t ns = static_cast<uint32_t>(__popcll(sm));
            if (count >= kWaveQ / VP_SHARED_END_DIV && ns < VP_SHARED_STEP_MIN) break;
            // the list is used up and many lanes are idle: end the phase so
            // the running walks can be shared out (after at least one step,
            // so every phase progresses)
            if (nsteps >= VP_STEAL_GAP && next >= nr && ns + VP_STEAL_IDLE <= nact &&
                __builtin_amdgcn_ballot_w64(step && walk4_pending(w) != 0ull) != 0ull)
                break;
            ++nsteps;
            uint32_t lm = 0, first = 0;
            Walk4 nw = w;
            if (step) lm = walk4_advance(nw, levels, mega_ray_mask(w4, half, w.n, inv, oi, kTMax), first);
            const bool fits = wave_push_blocks(wq, count, lm, first, ray);
            if (step && fits) w = nw;
            if (__builtin_amdgcn_ballot_w64(!fits) != 0) break;
        }
        wave_sync();
        // leaf phase, shared by all lanes: each queued leaf's crossing (entry
        // test and whole-chord depth from one transform) goes into its ray's
        // fixed-point sum
        // (two leaves per lane and round, both records loaded before either
        // crossing, so each lane has two loads in flight)
        for (uint32_t g = rank; g < count; g += 2u * nact) {
            const uint32_t g2 = g + nact;
            const bool two_leaves = g2 < count;
            const uint32_t own = wq.owner[g], own2 = two_leaves ? wq.owner[g2] : own;
            const PrimRecord pr = load_prim(wq.prims[own >> 1], wq.prim[g]);
            const PrimRecord pr2 = load_prim(wq.prims[own2 >> 1], two_leaves ? wq.prim[g2] : wq.prim[g]);
            const Crossing x = crossing(pr, lds_load3(wq.ro + 3 * (own >> 1)), lds_load3(wq.rd + 3 * own));
            if (x.tau > 0.0f) atomicAdd(&wq.acc[own], tau_fixed(x.tau));
            if (two_leaves) {
                const Crossing x2 = crossing(pr2, lds_load3(wq.ro + 3 * (own2 >> 1)), lds_load3(wq.rd + 3 * own2));
                if (x2.tau > 0.0f) atomicAdd(&wq.acc[own2], tau_fixed(x2.tau));
            }
        }
        wave_sync();
        // lanes whose ray is done (walked to the end, or saturated) take the
        // next ray of the list
        const bool fin = ray != kNoRay && (w.n == kWalkEnd || reached(ray & 127u));
        const unsigned long long fm = __builtin_amdgcn_ballot_w64(fin);
        const uint32_t k = next + lanes_below(fm);
        const uint32_t listed = k < nr ? wq.list[k < 128u ? k : 0u] : kNoRay;
        ray = fin ? listed : ray;
        if (fin) w = walk4_begin();
        next += static_cast<uint32_t>(__popcll(fm));
        // once the list is used up, idle lanes take over subtrees of running
        // walks.  Every lane then redoes its ray setup (the same values for a
        // lane that kept its ray): the refill and the handoff stay free of
        // branches around the setup, so every lane of the wavefront stays in
        // the loop and takes part in its ballots and leaf passes.
        const bool can_give = ray != kNoRay && w.n != kWalkEnd && !reached(ray & 127u);
        if (next >= nr) wave_steal(wq, ray == kNoRay, can_give, ray, w);
        setup();
        wave_sync();
    }
    wave_sync();
    acc_a = wq.acc[2 * lane];
    acc_b = wq.acc[2 * lane + 1];
}

__device__ __forceinline__ EnvView env_view(const EnvDev& e) {
    EnvView v;
    v.tex = e.tex;
    v.alias = e.has_table ? e.alias : nullptr;
    v.joint = e.joint;
    v.w = e.w;
    v.h = e.h;
    v.total = e.total;
    return v;
}

__device__ __forceinline__ uint32_t find_job(const uint64_t* begin, uint32_t njobs, uint64_t p) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (begin[mid] <= p) lo = mid;
        else hi = mid;
    }
    return lo;
}


// finalize_sample (path_stages.cuh:52-62): optional luminance clamp,
// non-finite radiance to zero.
__device__ __forceinline__ float4 finalize_radiance(const JobDesc& J, f3 L) {
    if (J.firefly_clamp > 0.0f) {
        const float lum = dot(L, f3{kLumR, kLumG, kLumB});
        if (lum > J.firefly_clamp) L = L * (J.firefly_clamp * rcpf(lum));
    }
    if (!isfinitef_(sum3(L))) L = mk(0.0f);
    return make_float4(L.x, L.y, L.z, 1.0f);
}

// Path setup (raygen.cuh:72-113) for path q of job J's current slice: pixel
// and sample from the 8x8-tile numbering, RNG stream, counter-hash key and
// jittered camera ray; pix receives the pixel index (y * width + x) and sl the
// sample's index within the slice.  False for the padding paths of partial
// tiles and for the paths of pixels that adaptive sampling has stopped.
__device__ __forceinline__ bool path_begin(const JobDesc& J, uint64_t q, f3& o, f3& d, Pcg32& rng, uint64_t& key,
                                           uint32_t& pix_out, uint32_t& sl_out) {
    // the slice's samples of one pixel are consecutive paths, so a
    // wavefront's lanes start on few pixels (nearly the same camera rays)
    const uint32_t r = static_cast<uint32_t>(q / J.s_count);
    const uint32_t sl = static_cast<uint32_t>(q - static_cast<uint64_t>(r) * J.s_count);
    // 8x8 tiles, in bands of kTileBand tile rows, column by column within a
    // band, so the paths in flight on one XCD (a run of consecutive path
    // indices) cover a compact patch of the image
    constexpr uint32_t kTileBand = 16;
    const uint32_t tile = r >> 6, w = r & 63u;
    const uint32_t tiles_y = (J.height + 7u) / 8u;
    const uint32_t band = tile / (kTileBand * J.tiles_x), c = tile - band * (kTileBand * J.tiles_x);
    const uint32_t rows = tiles_y - band * kTileBand < kTileBand ? tiles_y - band * kTileBand : kTileBand;
    const uint32_t px = (c / rows) * 8u + (w & 7u);
    const uint32_t py = (band * kTileBand + c % rows) * 8u + (w >> 3);
    if (px >= J.width || py >= J.height) return false;
    const uint32_t sample = J.s_begin + sl;
    const uint64_t pix = static_cast<uint64_t>(py) * J.width + px;
    // adaptive sampling: a converged pixel traces nothing more
    // (raygen.cuh:53-69); its paths end like padding
    if (J.adapt_state != nullptr && (J.adapt_state[pix] >> 31) != 0u) return false;
    const uint64_t seq = hash64(pix * J.spp + sample);
    rng = pcg_init(J.seed, seq);
    key = path_key(J.seed, seq);
    float jx, jy;
    if (J.filter_stddev > 0.0f) {
        gaussian_jitter(rng, J.filter_stddev, &jx, &jy);
    } else {
        jx = next_uniform(rng) - 0.5f;
        jy = next_uniform(rng) - 0.5f;
    }
    camera_ray(J.cam, px, py, jx, jy, &o, &d);
    pix_out = static_cast<uint32_t>(pix);
    sl_out = sl;
    return true;
}

// A finished path (raygen.cuh:214-238): its radiance is added to its pixel's
// exact fixed-point sums (fire-and-forget atomics: integer adds commute, so
// the total does not depend on which path finishes first), written to its
// per-sample slot if the caller asked for per-sample output, and counted
// against its job-table entry if the entry is tracked (the server).  The
// count follows the completion of the sums' atomics, so a host that sees the
// entry complete and then copies the sums reads finished ones.
__device__ __forceinline__ void path_finish(const JobDesc& J, uint32_t pix, uint32_t sl, f3 L) {
    const float4 v = finalize_radiance(J, L);
    unsigned long long* a = J.accum + static_cast<size_t>(pix) * kAccumWords;
    const float c[3] = {v.x, v.y, v.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint64_t q = radiance_fixed(c[k]);
        __hip_atomic_fetch_add(a + k, static_cast<unsigned long long>(q & 0xffffffffull), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if ((q >> 32) != 0ull)
            __hip_atomic_fetch_add(a + 4 + k, static_cast<unsigned long long>(q >> 32), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
    __hip_atomic_fetch_add(a + 3, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (J.samples != nullptr) J.samples[static_cast<size_t>(sl) * (static_cast<size_t>(J.width) * J.height) + pix] = v;
    if (J.done != nullptr) {
        // The sums are agent-scope atomics, which the device performs where
        // every compute unit sees them; the count only has to follow their
        // completion, so the wave waits for all its outstanding memory
        // operations (s_waitcnt with every counter at 0) before adding it.
        // An agent-scope release fence would also write back the XCD's
        // whole L2 for every finished path.  Entries with per-sample output
        // are not counted (the server asks for none).
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_fetch_add(J.done, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}


}  // namespace
}  // namespace vp
