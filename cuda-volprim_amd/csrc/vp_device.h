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

// Stackless walk of the rope layout over [0, tmax]; leaf(prim) returns false
// to stop.  tmax may be lowered by the caller between leaves.
template <typename Leaf>
__device__ __forceinline__ void walk_ray(const SceneView& s, f3 o, f3 inv, const float& tmax, uint32_t& visits,
                                         Leaf&& leaf) {
    // slab distances as fma(bound, 1/d, -o/d): which boxes a walk enters
    // only decides what is visited (leaf boxes are conservative), never a
    // result
    const f3 oi = o * inv;
    uint32_t i = 0;
    const uint32_t nn = s.n_nodes;
    while (i < nn) {
        const float4 a = as_global(s.nodes)[2 * i];
        const float4 b = as_global(s.nodes)[2 * i + 1];
        ++visits;
        const float tx0 = fmaf_(a.x, inv.x, -oi.x), tx1 = fmaf_(b.x, inv.x, -oi.x);
        const float ty0 = fmaf_(a.y, inv.y, -oi.y), ty1 = fmaf_(b.y, inv.y, -oi.y);
        const float tz0 = fmaf_(a.z, inv.z, -oi.z), tz1 = fmaf_(b.z, inv.z, -oi.z);
        const float tnear = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
        const float tfar = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
        const uint32_t skip = __float_as_uint(a.w);
        if (tnear <= tfar) {
            const uint32_t prim = __float_as_uint(b.w);
            if (prim != kInnerNode) {
                if (!leaf(prim)) return;
                i = skip;
            } else {
                i = i + 1;
            }
        } else {
            i = skip;
        }
    }
}

// Stackless walk visiting every leaf whose box contains x, in ascending
// stored-primitive order.
template <typename Leaf>
__device__ __forceinline__ void walk_point(const SceneView& s, f3 x, uint32_t& visits, Leaf&& leaf) {
    uint32_t i = 0;
    const uint32_t nn = s.n_nodes;
    while (i < nn) {
        const float4 a = as_global(s.nodes)[2 * i];
        const float4 b = as_global(s.nodes)[2 * i + 1];
        ++visits;
        const bool in = x.x >= a.x && x.x <= b.x && x.y >= a.y && x.y <= b.y && x.z >= a.z && x.z <= b.z;
        const uint32_t skip = __float_as_uint(a.w);
        if (in) {
            const uint32_t prim = __float_as_uint(b.w);
            if (prim != kInnerNode) {
                leaf(prim);
                i = skip;
            } else {
                i = i + 1;
            }
        } else {
            i = skip;
        }
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
#define VP_WAVE_Q 256
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
struct WaveQueue {
    uint32_t* prim;                // [kWaveQ]
    uint8_t* owner;                // [kWaveQ]
    float4* ro;                    // [64] ray origin of each lane
    float4* rd;                    // [64] ray direction of each lane
    unsigned long long* acc;       // [64] fixed-point optical depth of each lane's ray
    unsigned long long* key;       // [64] path key of each lane (argmin walks)
    const PrimRecord** prims;      // [64] primitive array of each lane's scene (lanes of one
                                   //      wavefront may run jobs over different scenes)
    uint32_t* ts;                  // [64] running minimum of each lane's argmin, as float bits
    uint32_t* nc;                  // [64] candidates recorded for each lane's argmin
    uint32_t* meta;                // [64] bounce | direct-term flag << 31 (argmin walks)
    float* tj;                     // [kWaveQ] entry distances of the leaves that passed the entry test
};
constexpr uint32_t kWaveQueueBytes = kWaveQ * 9 + 64 * (16 + 16 + 8 + 8 + 8 + 4 + 4 + 4);
// LDS bytes per wavefront: the larger of the two uses, in 16-byte units
constexpr uint32_t kWaveBytes =
    ((64 * kLeafRow * 4 > kWaveQueueBytes ? 64 * kLeafRow * 4 : kWaveQueueBytes) + 15) / 16 * 16;

// the calling wavefront's queue
__device__ __forceinline__ WaveQueue wave_queue(uint32_t* lds) {
    uint8_t* base = reinterpret_cast<uint8_t*>(lds) + (threadIdx.x >> 6) * kWaveBytes;
    WaveQueue q;
    q.prim = reinterpret_cast<uint32_t*>(base);
    q.owner = base + 4 * kWaveQ;
    q.ro = reinterpret_cast<float4*>(base + 5 * kWaveQ);
    q.rd = q.ro + 64;
    q.acc = reinterpret_cast<unsigned long long*>(q.rd + 64);
    q.key = q.acc + 64;
    q.prims = reinterpret_cast<const PrimRecord**>(q.key + 64);
    q.ts = reinterpret_cast<uint32_t*>(q.prims + 64);
    q.nc = q.ts + 64;
    q.meta = q.nc + 64;
    q.tj = reinterpret_cast<float*>(q.meta + 64);
    return q;
}

// Node phase of the walks: a lane steps while its leaf queue has room and its
// walk has nodes left.  With VP_WALK_MIN > 0 the loop runs for the whole
// wavefront, which leaves it early once fewer than VP_WALK_MIN lanes are
// stepping and some lane's queue is full, so full lanes do not wait for the
// last few; the stepping lanes resume after the leaf phase.  A node phase
// that starts with every queue empty can only end early after some lane has
// taken kLeafQ steps, so walks always progress.  Every query consumes leaves
// order-independently, so this changes only the schedule.  On the scattering
// workload (4096 Gaussians, 512x512, 64 spp) 4 and 8 measure alike, about
// 1.5 % faster than 0; 16 and 32 are slower.
#ifndef VP_WALK_MIN
#define VP_WALK_MIN 8
#endif

// true while this lane (VP_WALK_MIN == 0) or the wavefront stays in the node
// phase
__device__ __forceinline__ bool node_phase(bool step, uint32_t cnt) {
#if VP_WALK_MIN > 0
    const unsigned long long sm = __builtin_amdgcn_ballot_w64(step);
    if (sm == 0) return false;
    return __popcll(sm) >= VP_WALK_MIN || __builtin_amdgcn_ballot_w64(cnt == kLeafQ) == 0;
#else
    (void)cnt;
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

// One node step of a ray walk over [0, tmax]: a crossed leaf is queued, a
// crossed inner node is entered, anything else is skipped.  The queue slot
// cnt (< kLeafQ) is written whether or not the node is a crossed leaf; cnt
// only advances for one, so a later leaf overwrites the slot.  This keeps the
// step free of branches (with the padded rows, 1 % faster on the scattering
// workload than a conditional write into [slot][lane] columns).
__device__ __forceinline__ void ray_node_step(const SceneView& s, f3 inv, f3 oi, float tmax, uint32_t& i,
                                              uint32_t& cnt, const LeafQueue& q) {
    const float4 a = as_global(s.nodes)[2 * i];
    const float4 b = as_global(s.nodes)[2 * i + 1];
    const float tx0 = fmaf_(a.x, inv.x, -oi.x), tx1 = fmaf_(b.x, inv.x, -oi.x);
    const float ty0 = fmaf_(a.y, inv.y, -oi.y), ty1 = fmaf_(b.y, inv.y, -oi.y);
    const float tz0 = fmaf_(a.z, inv.z, -oi.z), tz1 = fmaf_(b.z, inv.z, -oi.z);
    const float tnear = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    const float tfar = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    const uint32_t skip = __float_as_uint(a.w);
    const uint32_t prim = __float_as_uint(b.w);
    const bool hit = tnear <= tfar, leaf_node = prim != kInnerNode;
    q.at(cnt) = prim;
    cnt += (hit && leaf_node) ? 1u : 0u;
    i = (hit && !leaf_node) ? i + 1 : skip;
}

template <typename Leaf>
__device__ __forceinline__ void walk_ray_q(const SceneView& s, f3 o, f3 inv, const float& tmax, uint32_t& visits,
                                           const LeafQueue& q, Leaf&& leaf) {
    const f3 oi = o * inv;
    uint32_t i = 0;
    const uint32_t nn = s.n_nodes;
    while (true) {
        uint32_t cnt = 0;
        while (true) {
            const bool step = i < nn && cnt < kLeafQ;
            if (!node_phase(step, cnt)) break;
            if (!step) continue;
            ++visits;
            ray_node_step(s, inv, oi, tmax, i, cnt, q);
        }
        for (uint32_t k = 0; k < cnt; ++k)
            if (!leaf(q.at(k))) return;
        if (i >= nn) return;
    }
}

template <typename Leaf>
__device__ __forceinline__ void walk_point_q(const SceneView& s, f3 x, uint32_t& visits, const LeafQueue& q,
                                             Leaf&& leaf) {
    uint32_t i = 0;
    const uint32_t nn = s.n_nodes;
    while (true) {
        uint32_t cnt = 0;
        while (i < nn && cnt < kLeafQ) {
            const float4 a = as_global(s.nodes)[2 * i];
            const float4 b = as_global(s.nodes)[2 * i + 1];
            ++visits;
            const bool in = x.x >= a.x && x.x <= b.x && x.y >= a.y && x.y <= b.y && x.z >= a.z && x.z <= b.z;
            const uint32_t skip = __float_as_uint(a.w);
            const uint32_t prim = __float_as_uint(b.w);
            const bool leaf_node = prim != kInnerNode;
            if (in && leaf_node) q.at(cnt++) = prim;
            i = (in && !leaf_node) ? i + 1 : skip;
        }
        for (uint32_t k = 0; k < cnt; ++k) leaf(q.at(k));
        if (i >= nn) return;
    }
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
// any-hit integration (anyhit.cuh:53-66).
template <bool kCount>
__device__ float transmittance(const SceneView& s, f3 o, f3 d, const uint32_t* act, uint32_t na, uint32_t stride,
                               const LeafQueue& lq, Ev<kCount>& ev) {
    uint64_t acc = 0;
    for (uint32_t k = 0; k < na; ++k) {
        const PrimRecord pr = load_prim(s.prims, act[k * stride]);
        const float te = exit_from_inside(pr, o, d);
        ev.add(kCntShActive);
        if (te > 0.0f && te < kTMax)
            if (tau_add(acc, optical_depth(pr, o, d, 0.0f, te))) return 0.0f;
    }
    const f3 inv = safe_inv(d);
    const float tmax = kTMax;
    bool full = false;
    uint32_t visits = 0, leaves = 0, sh = 0;
    walk_ray_q(s, o, inv, tmax, visits, lq, [&](uint32_t prim) -> bool {
        const PrimRecord pr = load_prim(s.prims, prim);
        ++leaves;
        float wl2;
        const float tj = entry_t(pr, o, d, &wl2);
        if (tj < 0.0f) return true;
        const float te = exit_from_entry(pr, o, d, tj, wl2);
        ++sh;
        if (te > tj && te < kTMax) {
            if (tau_add(acc, optical_depth(pr, o, d, tj, te))) {
                full = true;
                return false;
            }
        }
        return true;
    });
    ev.add(kCntNodeVisits, visits);
    ev.add(kCntVisitsShadow, visits);
    ev.add(kCntLeafTests, leaves);
    ev.add(kCntShHit, sh);
    if (full) return 0.0f;
    return tau_to_T(acc);
}

// number of set bits of m below the calling lane
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// one chord in fixed point, as tau_add adds it; a chord of tau >= 88 alone
// adds the saturation limit
__device__ __forceinline__ unsigned long long tau_fixed(float od) {
    const float c = od > 0.0f ? od : 0.0f;
    return c >= kMaxTau ? kTauLimitFixed : static_cast<unsigned long long>(c * kTauScaleF);
}

// The two shadow rays of one scatter vertex (MIS), each lane walking its ray
// a and then its ray b without waiting for the wavefront in between: a lane
// whose ray a ends starts ray b at once, so the wavefront's walk lasts as
// long as its longest a-plus-b pair instead of its longest a plus its longest
// b.  Both rays leave o through the same active set, whose chords are summed
// in one loop over the set.  Ta and Tb equal two transmittance() calls bit
// for bit (the same chords in the same fixed-point sums, and a ray that
// reaches tau >= 88 gives 0).
__device__ __forceinline__ void transmittance_two(const SceneView& s, f3 o, f3 da, f3 db, const uint32_t* act,
                                                  uint32_t na, uint32_t stride, const LeafQueue& q, float& Ta,
                                                  float& Tb) {
    uint64_t acc_a = 0, acc_b = 0;
    bool full_a = false, full_b = false;
    for (uint32_t k = 0; k < na && !(full_a && full_b); ++k) {
        const PrimRecord pr = load_prim(s.prims, act[k * stride]);
        if (!full_a) {
            const float te = exit_from_inside(pr, o, da);
            if (te > 0.0f && te < kTMax) full_a = tau_add(acc_a, optical_depth(pr, o, da, 0.0f, te));
        }
        if (!full_b) {
            const float te = exit_from_inside(pr, o, db);
            if (te > 0.0f && te < kTMax) full_b = tau_add(acc_b, optical_depth(pr, o, db, 0.0f, te));
        }
    }
    Ta = 0.0f;
    Tb = 0.0f;
    // current ray r: 0 = a, 1 = b, 2 = both done; a saturated ray is done
    uint32_t r = full_a ? (full_b ? 2u : 1u) : 0u;
    f3 d = r == 0 ? da : db;
    uint64_t acc = r == 0 ? acc_a : acc_b;
    bool full = false;
    f3 inv = safe_inv(d), oi = o * inv;
    uint32_t i = 0;
    const uint32_t nn = s.n_nodes;
    while (__ballot(r < 2u) != 0) {
        uint32_t cnt = 0;
        while (true) {
            const bool step = r < 2u && i < nn && cnt < kLeafQ;
            if (!node_phase(step, cnt)) break;
            if (!step) continue;
            ray_node_step(s, inv, oi, kTMax, i, cnt, q);
        }
        for (uint32_t k = 0; k < cnt && !full; ++k) {
            const PrimRecord pr = load_prim(s.prims, q.at(k));
            float wl2;
            const float tj = entry_t(pr, o, d, &wl2);
            if (tj < 0.0f) continue;
            const float te = exit_from_entry(pr, o, d, tj, wl2);
            if (te > tj && te < kTMax) full = tau_add(acc, optical_depth(pr, o, d, tj, te));
        }
        if (r < 2u && (full || i >= nn)) {
            const float T = full ? 0.0f : tau_to_T(acc);
            if (r == 0u) {
                Ta = T;
                r = full_b ? 2u : 1u;
                d = db;
                acc = acc_b;
                full = false;
                inv = safe_inv(d);
                oi = o * inv;
                i = 0;
            } else {
                Tb = T;
                r = 2u;
            }
        }
    }
}

// transmittance_two with the wave-shared leaf queue: the same rays, chords
// and fixed-point sums (a ray is saturated once its sum reaches the tau = 88
// limit, whatever the order its chords arrive in), so Ta and Tb equal two
// transmittance() calls bit for bit.  Every lane of the calling wavefront
// that is active here takes part in the leaf phases.
__device__ __forceinline__ void transmittance_two_shared(const SceneView& s, f3 o, f3 da, f3 db,
                                                         const uint32_t* act, uint32_t na, uint32_t stride,
                                                         const WaveQueue& wq, float& Ta, float& Tb) {
    uint64_t acc_a = 0, acc_b = 0;
    bool full_a = false, full_b = false;
    for (uint32_t k = 0; k < na && !(full_a && full_b); ++k) {
        const PrimRecord pr = load_prim(s.prims, act[k * stride]);
        if (!full_a) {
            const float te = exit_from_inside(pr, o, da);
            if (te > 0.0f && te < kTMax) full_a = tau_add(acc_a, optical_depth(pr, o, da, 0.0f, te));
        }
        if (!full_b) {
            const float te = exit_from_inside(pr, o, db);
            if (te > 0.0f && te < kTMax) full_b = tau_add(acc_b, optical_depth(pr, o, db, 0.0f, te));
        }
    }
    Ta = 0.0f;
    Tb = 0.0f;
    const uint32_t lane = __lane_id();
    // current ray r: 0 = a, 1 = b, 2 = both done; a saturated ray is done
    uint32_t r = full_a ? (full_b ? 2u : 1u) : 0u;
    f3 d = r == 0u ? da : db;
    wq.ro[lane] = make_float4(o.x, o.y, o.z, 0.0f);
    wq.rd[lane] = make_float4(d.x, d.y, d.z, 0.0f);
    wq.acc[lane] = r == 0u ? acc_a : acc_b;
    wq.prims[lane] = s.prims;
    f3 inv = safe_inv(d), oi = o * inv;
    uint32_t i = 0;
    const uint32_t nn = s.n_nodes;
    while (__ballot(r < 2u) != 0) {
        // node phase: every lane with a ray steps; crossed leaves are queued
        // for the wavefront.  It ends when no lane steps, when the queue
        // might not take one more step of every lane, or, once a quarter of
        // the queue is filled, when fewer than VP_WALK_MIN lanes still step.
        uint32_t count = 0;
        while (true) {
            const bool step = r < 2u && i < nn;
            const unsigned long long sm = __builtin_amdgcn_ballot_w64(step);
            if (sm == 0 || count > kWaveQ - 64u) break;
            if (count >= kWaveQ / 4u && static_cast<uint32_t>(__popcll(sm)) < VP_SHARED_STEP_MIN) break;
            bool push = false;
            uint32_t prim = 0;
            if (step) {
                const float4 a = as_global(s.nodes)[2 * i];
                const float4 b = as_global(s.nodes)[2 * i + 1];
                const float tx0 = fmaf_(a.x, inv.x, -oi.x), tx1 = fmaf_(b.x, inv.x, -oi.x);
                const float ty0 = fmaf_(a.y, inv.y, -oi.y), ty1 = fmaf_(b.y, inv.y, -oi.y);
                const float tz0 = fmaf_(a.z, inv.z, -oi.z), tz1 = fmaf_(b.z, inv.z, -oi.z);
                const float tnear = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
                const float tfar = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), kTMax));
                const uint32_t skip = __float_as_uint(a.w);
                prim = __float_as_uint(b.w);
                const bool hit = tnear <= tfar, leaf_node = prim != kInnerNode;
                push = hit && leaf_node;
                i = (hit && !leaf_node) ? i + 1 : skip;
            }
            const unsigned long long pm = __builtin_amdgcn_ballot_w64(push);
            if (push) {
                const uint32_t pos = count + lanes_below(pm);
                wq.prim[pos] = prim;
                wq.owner[pos] = static_cast<uint8_t>(lane);
            }
            count += static_cast<uint32_t>(__popcll(pm));
        }
        // leaf phase: the active lanes share the queued leaves
        const unsigned long long am = __builtin_amdgcn_ballot_w64(true);
        const uint32_t nact = static_cast<uint32_t>(__popcll(am));
#if VP_SHADOW_TWO_STAGE
        // entry tests first; the leaves that are entered move down to the
        // front of the queue (an entry is only overwritten after it has been
        // read), then their chords are integrated at full width
        uint32_t nhit = 0;
        for (uint32_t base = 0; base < count; base += nact) {
            const uint32_t g = base + lanes_below(am);
            bool entered = false;
            uint32_t own = 0, prim = 0;
            float tj = -1.0f;
            if (g < count) {
                own = wq.owner[g];
                prim = wq.prim[g];
                const float4 o4 = wq.ro[own], d4 = wq.rd[own];
                const PrimRecord pr = load_prim(wq.prims[own], prim);
                float wl2;
                tj = entry_t(pr, f3{o4.x, o4.y, o4.z}, f3{d4.x, d4.y, d4.z}, &wl2);
                entered = tj >= 0.0f;
            }
            const unsigned long long hm = __builtin_amdgcn_ballot_w64(entered);
            if (entered) {
                const uint32_t h = nhit + lanes_below(hm);
                wq.prim[h] = prim;
                wq.owner[h] = static_cast<uint8_t>(own);
                wq.tj[h] = tj;
            }
            nhit += static_cast<uint32_t>(__popcll(hm));
        }
        for (uint32_t g = lanes_below(am); g < nhit; g += nact) {
            const uint32_t own = wq.owner[g];
            const float4 o4 = wq.ro[own], d4 = wq.rd[own];
            const f3 ro = f3{o4.x, o4.y, o4.z}, rd = f3{d4.x, d4.y, d4.z};
            const PrimRecord pr = load_prim(wq.prims[own], wq.prim[g]);
            const float tj = wq.tj[g];
            const float te = exit_from_entry(pr, ro, rd, tj, length2(to_local_dir(pr, rd)));
            if (te > tj && te < kTMax) atomicAdd(&wq.acc[own], tau_fixed(optical_depth(pr, ro, rd, tj, te)));
        }
#else
        for (uint32_t g = lanes_below(am); g < count; g += nact) {
            const uint32_t own = wq.owner[g];
            const float4 o4 = wq.ro[own], d4 = wq.rd[own];
            const f3 ro = f3{o4.x, o4.y, o4.z}, rd = f3{d4.x, d4.y, d4.z};
            const PrimRecord pr = load_prim(wq.prims[own], wq.prim[g]);
            float wl2;
            const float tj = entry_t(pr, ro, rd, &wl2);
            if (tj < 0.0f) continue;
            const float te = exit_from_entry(pr, ro, rd, tj, wl2);
            if (te > tj && te < kTMax) atomicAdd(&wq.acc[own], tau_fixed(optical_depth(pr, ro, rd, tj, te)));
        }
#endif
        // each lane looks at its own ray
        if (r < 2u) {
            const unsigned long long acc = wq.acc[lane];
            const bool full = acc >= kTauLimitFixed;
            if (full || i >= nn) {
                const float T = full ? 0.0f : tau_to_T(acc);
                if (r == 0u) {
                    Ta = T;
                    r = full_b ? 2u : 1u;
                    d = db;
                    wq.rd[lane] = make_float4(d.x, d.y, d.z, 0.0f);
                    wq.acc[lane] = acc_b;
                    inv = safe_inv(d);
                    oi = o * inv;
                    i = 0;
                } else {
                    Tb = T;
                    r = 2u;
                }
            }
        }
    }
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


// Path setup (raygen.cuh:72-113) for path q of job J's current slice: pixel
// and sample from the 8x8-tile numbering, RNG stream, counter-hash key,
// jittered camera ray and the per-sample output slot.  False for the padding
// paths of partial tiles.
__device__ __forceinline__ bool path_begin(const JobDesc& J, uint64_t q, f3& o, f3& d, Pcg32& rng, uint64_t& key,
                                           uint32_t& slot) {
    const uint32_t sl = static_cast<uint32_t>(q / J.npad);
    const uint32_t r = static_cast<uint32_t>(q - static_cast<uint64_t>(sl) * J.npad);
    const uint32_t tile = r >> 6, w = r & 63u;
    const uint32_t px = (tile % J.tiles_x) * 8u + (w & 7u);
    const uint32_t py = (tile / J.tiles_x) * 8u + (w >> 3);
    if (px >= J.width || py >= J.height) return false;
    const uint32_t sample = J.s_begin + sl;
    const uint64_t pix = static_cast<uint64_t>(py) * J.width + px;
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
    slot = static_cast<uint32_t>(static_cast<uint64_t>(sl) * J.width * J.height + pix);
    return true;
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

}  // namespace
}  // namespace vp
