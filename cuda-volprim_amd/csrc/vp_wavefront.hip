// vp_wavefront.hip — the wavefront-queue variant of the renderer.
//
// A pool of P path slots lives in HBM (structure of arrays).  One iteration
// advances every live path by one bounce with six kernels, each over a
// compacted queue so its wavefronts do one kind of work:
//   k_wf_shade        per slot: add the previous bounce's shadow-ray terms to
//                     L in the estimator's order, retire finished paths
//                     into their pixels' sums (counted per job), queue free and
//                     live slots (ballot + popcount prefix compaction, one
//                     atomic per wavefront);
//   k_wf_claim        one thread: hands out min(free slots, unclaimed paths)
//                     path numbers, so jobs can be added between iterations;
//   k_wf_regen        per claimed path: setup and origin set, queued live;
//   k_wf_extend       per live path: argmin free flight (origin set + one
//                     pruned BVH walk);
//   k_wf_scatter      per live path: bounce-0 analytic direct ray, active-set
//                     rebuild at the scatter vertex, albedo, phase sampling,
//                     NEE direction sampling, Russian roulette; shadow rays
//                     are appended to the shadow queue;
//   k_wf_shadow       per shadow ray: fixed-point transmittance.
// Each path draws its random numbers in the same order as the megakernel
// (camera jitter, phase, NEE, roulette) and adds its radiance terms in the
// same order, so both variants produce identical per-sample radiance.
//
// Queue counters come in two parity sets; the first kernel of an iteration
// clears the set the next iteration will use, so a two-iteration sequence
// can be captured once as a HIP graph and replayed (no memset nodes).
//
// Reference (kfernandez31/cuda-volprim): device/entry/raygen.cuh:21-243,
// device/entry/path_stages.cuh:25-226, device/core/sampling.cuh:257-583.
#include "vp_device.h"

namespace vp {
namespace {

enum : uint32_t { kSlotFree = 0, kSlotAlive = 1, kSlotDone = 2 };

// misc.w bit fields
__device__ __forceinline__ uint32_t f_cur(uint32_t w) { return w & 1u; }
__device__ __forceinline__ uint32_t f_state(uint32_t w) { return (w >> 1) & 3u; }
__device__ __forceinline__ uint32_t f_npend(uint32_t w) { return (w >> 3) & 3u; }
__device__ __forceinline__ uint32_t f_pack(uint32_t cur, uint32_t state, uint32_t npend) {
    return cur | (state << 1) | (npend << 3);
}

// Wave-aggregated append: lanes with `want` get consecutive indices from one
// atomicAdd on *counter per wavefront.
__device__ __forceinline__ uint32_t wave_append(unsigned int* counter, bool want) {
    const unsigned long long m = __ballot(want);
    if (m == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint32_t leader = __ffsll(static_cast<long long>(m)) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, static_cast<unsigned int>(__popcll(m)));
    base = __shfl(base, leader);
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    return base + static_cast<uint32_t>(__popcll(m & lt));
}

__device__ __forceinline__ void wave_count(unsigned long long* counter, bool flag) {
    const unsigned long long m = __ballot(flag);
    if (m != 0 && __lane_id() == static_cast<uint32_t>(__ffsll(static_cast<long long>(m)) - 1))
        atomicAdd(counter, static_cast<unsigned long long>(__popcll(m)));
}

#define VP_WF_LB __launch_bounds__(256)

// Per-key counts of the flagged lanes of a wavefront, one atomic per distinct
// key (lanes usually share one job).  Wave-uniform control flow.
__device__ __forceinline__ void wave_add_per_key(unsigned long long* arr, bool flag, uint32_t key) {
    unsigned long long pending = __ballot(flag);
    while (pending) {
        const uint32_t leader = __ffsll(static_cast<long long>(pending)) - 1;
        const uint32_t k = __shfl(key, leader);
        const unsigned long long same = __ballot(flag && key == k) & pending;
        if (__lane_id() == leader) atomicAdd(&arr[k], static_cast<unsigned long long>(__popcll(same)));
        pending &= ~same;
    }
}

// queue counters of parity par: [0] live, [1] shadow, [2] free, [3] claimed,
// [4] live paths taken by k_wf_extend, [5] shadow rays taken by k_wf_shadow
constexpr uint32_t kQc = 8;
__device__ __forceinline__ unsigned int* qcs(const WfPool& w, uint32_t par) { return w.qc + kQc * par; }

__global__ void VP_WF_LB k_wf_init(WfPool w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.P) w.misc[i] = make_uint4(0u, 0u, 0u, f_pack(0u, kSlotFree, 0u));
    if (i < 2 * kQc) w.qc[i] = 0u;
    if (i < 2) w.claim_base[i] = 0ull;
    if (i == 0) *w.finished = 0ull;
}

// Stage 1 of an iteration, one thread per slot: add the finished shadow-ray
// terms, retire finished paths, queue free and live slots.
__global__ void VP_WF_LB k_wf_shade(const JobDesc* __restrict__ jobs, WfPool w, uint32_t par) {
    const uint32_t P = w.P;
    if (blockIdx.x == 0 && threadIdx.x < kQc) qcs(w, par ^ 1u)[threadIdx.x] = 0u;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < P;
    uint4 m = in ? w.misc[i] : make_uint4(0u, 0u, 0u, f_pack(0u, kSlotFree, 0u));
    uint32_t state = f_state(m.w);
    const uint32_t npend = f_npend(m.w);
    bool retired = false;
    uint32_t job = 0;
    if (in && state != kSlotFree) {
        const float4 lr = w.lr[i];
        f3 L = f3{lr.x, lr.y, lr.z};
        for (uint32_t k = 0; k < npend; ++k) {
            const float T = w.pd[k * P + i].w;
            const float4 a = w.pw[k * P + i];
            L = L + f3{a.x, a.y, a.z} * T;
        }
        if (state == kSlotDone) {
            job = __float_as_uint(w.dj[i].w);
            const JobDesc& J = jobs[job];
            path_finish(J, m.z, 0u, L);
            state = kSlotFree;
            retired = true;
        } else {
            w.lr[i] = make_float4(L.x, L.y, L.z, lr.w);
        }
        m.w = f_pack(f_cur(m.w), state, 0u);
        w.misc[i].w = m.w;
    }
    wave_count(w.finished, retired);
    wave_add_per_key(w.job_done, retired, job);
    const bool is_free = in && state == kSlotFree;
    const uint32_t qf = wave_append(&qcs(w, par)[2], is_free);
    if (is_free) w.free_q[qf] = i;
    const bool live = in && state == kSlotAlive;
    const uint32_t ql = wave_append(&qcs(w, par)[0], live);
    if (live) w.live[ql] = i;
}

// Stage 2, one thread: hand out min(free slots, unclaimed paths) path
// numbers, so the counter never passes the installed total (jobs may be
// added between iterations).
__global__ void k_wf_claim(const uint64_t* __restrict__ launch_desc, unsigned long long* __restrict__ path_counter,
                           WfPool w, uint32_t par) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint64_t total = launch_desc[1];
    const unsigned long long c = *path_counter;
    const uint64_t avail = total > c ? total - c : 0ull;
    const uint32_t nfree = qcs(w, par)[2];
    const uint32_t n = static_cast<uint32_t>(avail < nfree ? avail : nfree);
    w.claim_base[par] = c;
    *path_counter = c + n;
    qcs(w, par)[3] = n;
}

// Stage 3: start the claimed paths in the first free slots, trace their
// origin sets, queue them as live.
__global__ void VP_WF_LB k_wf_regen(const JobDesc* __restrict__ jobs, const uint64_t* __restrict__ job_begin,
                                    const uint64_t* __restrict__ launch_desc, WfPool w, uint32_t par,
                                    unsigned long long* __restrict__ counters) {
    const uint32_t njobs = static_cast<uint32_t>(launch_desc[0]);
    const uint32_t n = qcs(w, par)[3];
    const uint64_t base = w.claim_base[par];
    const uint32_t P = w.P;
    uint32_t overflow = 0;
    const uint32_t span = gridDim.x * blockDim.x;
    for (uint32_t first = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); first < n; first += span) {
        const uint32_t idx = first + __lane_id();
        bool live = false, padded = false;
        uint32_t i = 0, j = 0;
        if (idx < n) {
            i = w.free_q[idx];
            const uint64_t p = base + idx;
            j = find_job(job_begin, njobs, p);
            const JobDesc& J = jobs[j];
            f3 o, d;
            Pcg32 rng;
            uint64_t key;
            uint32_t slot, sl;
            if (path_begin(J, p - J.path_begin, o, d, rng, key, slot, sl)) {
                // origin-inside set (sampling.cuh:322-341) by point query, into buffer 0
                uint32_t na = 0, visits = 0;
                uint32_t* act = w.slab + static_cast<size_t>(i) * kSlabBuffers * w.cap;
                const SceneView sc = J.scene;
                walk_point(sc, o, visits, [&](uint32_t prim) {
                    if (point_inside(load_prim(sc.prims, prim), o)) {
                        const bool fits = na < w.cap;
                        if (fits) act[na] = prim;
                        na += fits ? 1u : 0u;
                        overflow += fits ? 0u : 1u;
                    }
                });
                w.ob[i] = make_float4(o.x, o.y, o.z, __uint_as_float(0u));
                w.dj[i] = make_float4(d.x, d.y, d.z, __uint_as_float(j));
                w.bt[i] = make_float4(1.0f, 1.0f, 1.0f, kTMax);
                w.lr[i] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(na));
                w.rng[i] = make_uint4(static_cast<uint32_t>(rng.state), static_cast<uint32_t>(rng.state >> 32),
                                      static_cast<uint32_t>(rng.inc), static_cast<uint32_t>(rng.inc >> 32));
                w.misc[i] = make_uint4(static_cast<uint32_t>(key), static_cast<uint32_t>(key >> 32), slot,
                                       f_pack(0u, kSlotAlive, 0u));
                live = true;
            } else {
                padded = true;  // a partial tile's padding path: done at once, the slot stays free
            }
        }
        wave_count(w.finished, padded);
        wave_add_per_key(w.job_done, padded, j);
        const uint32_t q = wave_append(&qcs(w, par)[0], live);
        if (live) w.live[q] = i;
    }
    if (overflow) atomicAdd(&counters[kCntActiveOverflow], static_cast<unsigned long long>(overflow));
}

// Work fetch for the walk kernels: every idle lane of a wavefront takes the
// next queue entry (one atomic per wavefront), so lanes whose walk ended
// start another one instead of idling until the wavefront's longest walk is
// done.  Returns the entry for this lane or n (none).
__device__ __forceinline__ uint32_t wave_fetch(unsigned int* head, bool want, uint32_t n, uint32_t fetch_min) {
    const unsigned long long m = __ballot(want);
    if (m == 0) return n;
    // fetch once fetch_min lanes are idle, or when none is busy
    if (static_cast<uint32_t>(__popcll(m)) < fetch_min && __ballot(!want) != 0) return n;
    const uint32_t lane = __lane_id();
    const uint32_t leader = __ffsll(static_cast<long long>(m)) - 1;
    uint32_t base = n;
    if (lane == leader) {
        // the head only grows; once it has passed n nobody adds to it again,
        // so it stays far from wrapping around
        if (__hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n)
            base = atomicAdd(head, static_cast<unsigned int>(__popcll(m)));
    }
    base = __shfl(base, leader);
    if (base >= n) return n;
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t q = base + static_cast<uint32_t>(__popcll(m & lt));
    return want && q < n ? q : n;
}

// Argmin free flight (sampling.cuh:350-415) for every live path: origin-set
// candidates first, then one pruned walk over entered primitives; lanes
// refill from the live queue as their walks end.  The leaves all lanes' walks
// cross go to the wave-shared queue and are evaluated by whichever lanes are
// free: each path's running minimum is an LDS atomic minimum over float bits
// (every candidate is >= 0) and its rebuild candidates go through an LDS
// counter, so the results do not depend on which lane handles which leaf.
__global__ void VP_WF_LB k_wf_extend(const JobDesc* __restrict__ jobs, WfPool w, uint32_t par, uint32_t fetch_min) {
    extern __shared__ __align__(16) uint32_t s_leafq[];
    const WaveQueue wq = wave_queue(s_leafq);
    uint32_t* slot = wq.nn;  // the pool slot of each lane's path
    const uint32_t lane = __lane_id();
    unsigned int* qc = qcs(w, par);
    const uint32_t n = qc[0];
    bool have = false;
    uint32_t i = 0;
    Walk4 w4w = walk4_begin();
    float ts_origin = kTMax, tmax = kTMax;
    f3 inv = mk(0.0f), oi = mk(0.0f);
    SceneView sc{};
    while (true) {
        const uint32_t idx = wave_fetch(&qc[4], !have, n, fetch_min);
        if (!have && idx < n) {
            i = w.live[idx];
            const float4 ob = w.ob[i], dj = w.dj[i];
            const f3 o = f3{ob.x, ob.y, ob.z};
            const f3 d = f3{dj.x, dj.y, dj.z};
            const uint32_t b = __float_as_uint(ob.w);
            sc = jobs[__float_as_uint(dj.w)].scene;
            const uint4 m = w.misc[i];
            const uint64_t key = static_cast<uint64_t>(m.x) | (static_cast<uint64_t>(m.y) << 32);
            const uint32_t na = __float_as_uint(w.lr[i].w);
            const uint32_t* act = w.slab + (static_cast<size_t>(i) * kSlabBuffers + f_cur(m.w)) * w.cap;
            // argmin over the origin set (sampling.cuh:356-370)
            float ts = kTMax;
            for (uint32_t k = 0; k < na; ++k) {
                const uint32_t pi = act[k];
                const PrimRecord pr = load_prim(sc.prims, pi);
                const InsideChord x = inside_chord(pr, o, d);
                const float tau = free_flight_tau(prim_uniform(key, b, 0u, pi));
                if (tau <= x.tau) ts = minf(ts, origin_flight(pr, o, d, x, tau));
            }
            lds_store3(wq.ro + 3 * lane, o);
            lds_store3(wq.rd + 3 * lane, d);
            wq.key[lane] = key;
            wq.meta[lane] = b;
            wq.prims[lane] = sc.prims;
            wq.ts[lane] = ts == 0.0f ? 0u : __float_as_uint(ts);
            wq.nc[lane] = 0u;
            slot[lane] = i;
            ts_origin = ts;
            tmax = ts;
            inv = safe_inv(d);
            oi = o * inv;
            w4w = walk4_begin();
            have = true;
        }
        wave_sync();
        if (__ballot(have) == 0) break;
        // node phase: crossed leaves go to the wave queue
        uint32_t count = 0;
        while (true) {
            const bool step = have && w4w.n != kWalkEnd;
            const unsigned long long sm = __builtin_amdgcn_ballot_w64(step);
            if (sm == 0) break;
            if (count >= kWaveQ / VP_SHARED_END_DIV && static_cast<uint32_t>(__popcll(sm)) < VP_SHARED_STEP_MIN) break;
            uint32_t lm = 0, first = 0;
            Walk4 nw = w4w;
            if (step)
                lm = walk4_advance(nw, walk4_block_levels(sc.w4_levels),
                                   ray_mask_nf(sc.w4, sc.w4_half, w4w.n, inv, oi, tmax), first);
            const bool fits = wave_push_blocks(wq, count, lm, first, lane);
            if (step && fits) w4w = nw;
            if (__builtin_amdgcn_ballot_w64(!fits) != 0) break;
        }
        wave_sync();
        // leaf phase, shared by all lanes (sampling.cuh:372-412)
        const unsigned long long am = __builtin_amdgcn_ballot_w64(true);
        const uint32_t nact = static_cast<uint32_t>(__popcll(am));
        for (uint32_t g = lanes_below(am); g < count; g += nact) {
            const uint32_t own = wq.owner[g];
            const uint32_t prim = wq.prim[g];
            const f3 ro = lds_load3(wq.ro + 3 * own), rd = lds_load3(wq.rd + 3 * own);
            const PrimRecord pr = load_prim(wq.prims[own], prim);
            const Crossing x = crossing(pr, ro, rd);
            // entered before the running minimum: a rebuild candidate.
            // Beyond it nothing can change (t >= entry > minimum).
            if (x.t0 < 0.0f || x.t0 > __uint_as_float(wq.ts[own])) continue;
            const uint32_t pos = atomicAdd(&wq.nc[own], 1u);
            if (pos < w.cap) w.slab[(static_cast<size_t>(slot[own]) * kSlabBuffers + 2) * w.cap + pos] = prim;
            const float tau = free_flight_tau(prim_uniform(wq.key[own], wq.meta[own], 1u, prim));
            if (tau <= x.tau) {
                const float t = entered_flight(pr, ro, rd, x, tau);
                atomicMin(&wq.ts[own], t == 0.0f ? 0u : __float_as_uint(t));
            }
        }
        wave_sync();
        if (have) {
            const float wmin = __uint_as_float(wq.ts[lane]);
            tmax = wmin;
            if (w4w.n == kWalkEnd) {
                w.bt[i].w = wmin < ts_origin ? wmin : ts_origin;
                w.ncand[i] = wq.nc[lane];
                have = false;
            }
        }
    }
}

// Pending radiance term k of slot i: contribution kf * T once the shadow
// kernel has written T (kf: the factor T multiplies).  Until then pd.w
// carries the ray's shadow
// roulette: its first threshold >> 24 in bits 0..22 (at most 88 * 2^16) and
// its failing level in bits 23..29.
__device__ __forceinline__ uint32_t rr_bits(const ShadowRr& rr) {
    return static_cast<uint32_t>(rr.thr >> 24) | (rr.fail << 23);
}
__device__ __forceinline__ ShadowRr rr_unpack(uint32_t bits) {
    ShadowRr rr;
    rr.thr = static_cast<uint64_t>(bits & 0x7fffffu) << 24;
    rr.fail = bits >> 23;
    return rr;
}
__device__ __forceinline__ void put_pending(const WfPool& w, uint32_t i, uint32_t k, f3 org, uint32_t set,
                                            uint32_t count, f3 dir, f3 kf, uint32_t rrbits) {
    const uint32_t e = k * w.P + i;
    w.po[e] = make_float4(org.x, org.y, org.z, __uint_as_float(set | (count << 1)));
    w.pd[e] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(rrbits));
    w.pw[e] = make_float4(kf.x, kf.y, kf.z, 0.0f);
}
static_assert((kTauLimitFixed >> 24) < (1u << 23) && (kTauLimitFixed & kRrGrain) == 0, "roulette bits fit pd.w");

__global__ void VP_WF_LB k_wf_scatter(const JobDesc* __restrict__ jobs, WfPool w, uint32_t par,
                                      unsigned long long* __restrict__ counters) {
    extern __shared__ uint32_t s_leafq[];
    const LeafQueue lq = leaf_queue(s_leafq);
    const uint32_t n = qcs(w, par)[0];
    const uint32_t P = w.P;
    uint32_t overflow = 0;
    // every lane of a wavefront runs the same trip count so the queue
    // appends below stay wave-uniform
    const uint32_t span = gridDim.x * blockDim.x;
    const uint32_t first = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    for (uint32_t base = first; base < n; base += span) {
        const uint32_t idx = base + __lane_id();
        const bool active = idx < n;
        uint32_t npend = 0;
        uint32_t i = 0;
        if (active) {
            i = w.live[idx];
            const float4 ob = w.ob[i], dj = w.dj[i], bt = w.bt[i];
            const f3 o = f3{ob.x, ob.y, ob.z}, d = f3{dj.x, dj.y, dj.z};
            const uint32_t b = __float_as_uint(ob.w);
            const JobDesc& J = jobs[__float_as_uint(dj.w)];
            const SceneView sc = J.scene;
            const EnvView env = env_view(J.env);
            uint4 m = w.misc[i];
            const uint64_t key = static_cast<uint64_t>(m.x) | (static_cast<uint64_t>(m.y) << 32);
            uint32_t cur = f_cur(m.w);
            const uint32_t na = __float_as_uint(w.lr[i].w);
            f3 beta = f3{bt.x, bt.y, bt.z};
            const float ts = bt.w;
            uint4 rs = w.rng[i];
            Pcg32 rng;
            rng.state = static_cast<uint64_t>(rs.x) | (static_cast<uint64_t>(rs.y) << 32);
            rng.inc = static_cast<uint64_t>(rs.z) | (static_cast<uint64_t>(rs.w) << 32);
            bool finish = false;
            uint32_t nn = na;

            // analytic direct term at bounce 0 (raygen.cuh:148-160): T along the
            // camera ray with the origin set, under its own roulette (ray 2)
            if (b == 0) {
                const f3 kd = beta * env_lookup(env, d);
                put_pending(w, i, npend++, o, cur, na, d, kd, rr_bits(shadow_rr(max3(kd), J.shadow_rr, key, 0u, 2u)));
            }

            if (ts >= kTMax) {
                finish = true;
            } else {
                // rebuild the active set at the scatter point (sampling.cuh:422-458)
                const uint32_t* act = w.slab + (static_cast<size_t>(i) * kSlabBuffers + cur) * w.cap;
                uint32_t* nxt = w.slab + (static_cast<size_t>(i) * kSlabBuffers + (cur ^ 1u)) * w.cap;
                uint32_t* cnd = w.slab + (static_cast<size_t>(i) * kSlabBuffers + 2) * w.cap;
                const uint32_t nc = w.ncand[i];
                const f3 x = o + ts * d;
                uint32_t cnt = 0, ia = 0, visits = 0;
                auto push = [&](uint32_t v) {
                    const bool fits = cnt < w.cap;
                    if (fits) nxt[cnt] = v;
                    cnt += fits ? 1u : 0u;
                    overflow += fits ? 0u : 1u;
                };
                auto drain_below = [&](uint32_t c) {
                    while (ia < na) {
                        const uint32_t a = act[ia];
                        if (a >= c) break;
                        if (ts <= exit_from_inside(load_prim(sc.prims, a), o, d)) push(a);
                        ++ia;
                    }
                };
                auto visit = [&](uint32_t c) {
                    const PrimRecord pr = load_prim(sc.prims, c);
                    bool keep = false;
                    if (ia < na && act[ia] == c) {
                        keep = ts <= exit_from_inside(pr, o, d);
                        ++ia;
                    }
                    if (!keep) {
                        const Crossing cx = crossing(pr, o, d);
                        keep = cx.t0 >= 0.0f && cx.t0 <= ts && ts <= cx.t1;
                    }
                    if (keep) push(c);
                };
                if (nc <= w.cap) {
                    // the candidates the extend walk recorded, in ascending index order
                    for (uint32_t q = 1; q < nc; ++q) {
                        const uint32_t v = cnd[q];
                        uint32_t j = q;
                        while (j > 0 && cnd[j - 1] > v) {
                            cnd[j] = cnd[j - 1];
                            --j;
                        }
                        cnd[j] = v;
                    }
                    for (uint32_t q = 0; q < nc; ++q) {
                        drain_below(cnd[q]);
                        visit(cnd[q]);
                    }
                } else {
                    // more candidates than fit: one BVH point query at the vertex
                    walk_point_q(sc, x, visits, lq, [&](uint32_t c) {
                        drain_below(c);
                        visit(c);
                    });
                }
                while (ia < na) {
                    const uint32_t a = act[ia];
                    if (ts <= exit_from_inside(load_prim(sc.prims, a), o, d)) push(a);
                    ++ia;
                }
                cur ^= 1u;
                nn = cnt;
                const uint32_t* nact = nxt;

                float phase_pdf;
                const f3 wo = phase_sample(J.hg, d, rng, &phase_pdf);

                // evaluate_albedo (sampling.cuh:257-283)
                f3 aa = mk(0.0f);
                float aw = 0.0f;
                for (uint32_t k = 0; k < nn; ++k) {
                    const PrimRecord pr = load_prim(sc.prims, nact[k]);
                    const float wgt = pr.mass * prim_pdf(pr, x);
                    aa = aa + f3{pr.ar, pr.ag, pr.ab} * wgt;
                    aw += wgt;
                }
                const f3 albedo = aw <= 0.0f ? mk(0.0f) : aa * rcpf(aw);

                // accumulate_nee_direct (path_stages.cuh:148-226): directions and
                // weights now, transmittance in the shadow kernel
                const f3 base3 = beta * albedo;
                const f3 luma = f3{kLumR, kLumG, kLumB};
                if (J.use_ris) {
                    float wsum = 0.0f;
                    f3 ydir = mk(0.0f), yenv = mk(0.0f);
                    for (uint32_t c = 0; c < J.ris_candidates; ++c) {
                        float cpdf;
                        const f3 cwo = env_sample(env, rng, &cpdf);
                        if (cpdf <= 0.0f) continue;
                        const f3 envc = env_lookup(env, cwo);
                        const float lumc = dot(envc, luma);
                        const float wgt = phase_eval(J.hg, d, cwo) * lumc * rcpf(cpdf);
                        if (!(wgt > 0.0f)) continue;
                        wsum += wgt;
                        if (next_uniform(rng) * wsum < wgt) {
                            ydir = cwo;
                            yenv = envc;
                        }
                    }
                    if (wsum > 0.0f) {
                        const float lumy = dot(yenv, luma);
                        if (lumy > 0.0f) {
                            const float Wr = wsum * rcpf(static_cast<float>(J.ris_candidates));
                            const f3 ky = (base3 * (yenv * rcpf(lumy))) * Wr;
                            put_pending(w, i, npend++, x, cur, nn, ydir, ky,
                                        rr_bits(shadow_rr(max3(ky), J.shadow_rr, key, b, 0u)));
                        }
                    }
                } else {
                    float apdf;
                    const f3 awo = phase_sample(J.hg, d, rng, &apdf);
                    const float pb_a = env_pdf(env, awo);
                    const float sa = apdf + pb_a;
                    const float wa = sa > 0.0f ? apdf / sa : 0.0f;
                    const f3 enva = env_lookup(env, awo);
                    const f3 ka = (base3 * enva) * wa;
                    put_pending(w, i, npend++, x, cur, nn, awo, ka, rr_bits(shadow_rr(max3(ka), J.shadow_rr, key, b, 0u)));

                    float bpdf;
                    const f3 bwo = env_sample(env, rng, &bpdf);
                    const float ph_b = phase_eval(J.hg, d, bwo);
                    const float sb = bpdf + ph_b;
                    const float wb = sb > 0.0f ? bpdf / sb : 0.0f;
                    const f3 envb = env_lookup(env, bwo);
                    const f3 kb = (((base3 * envb) * wb) * ph_b) * rcpf(bpdf);
                    put_pending(w, i, npend++, x, cur, nn, bwo, kb, rr_bits(shadow_rr(max3(kb), J.shadow_rr, key, b, 1u)));
                }
                beta = beta * albedo;

                // russian_roulette (path_stages.cuh:25-37) and guards (raygen.cuh:191-204)
                if (b >= J.rr_depth) {
                    const float ps = minf(J.rr_max_survival, max3(beta));
                    if (next_uniform(rng) > ps || ps <= 0.0f) finish = true;
                    else beta = beta * rcpf(ps);
                }
                if (!finish && max3(beta) < kMinThroughput) finish = true;
                if (!finish && !isfinitef_(sum3(beta))) finish = true;
                if (!finish) {
                    w.ob[i] = make_float4(x.x, x.y, x.z, __uint_as_float(b + 1));
                    w.dj[i] = make_float4(wo.x, wo.y, wo.z, dj.w);
                    if (b + 1 >= J.max_depth) finish = true;
                }
            }
            w.bt[i] = make_float4(beta.x, beta.y, beta.z, ts);
            w.lr[i].w = __uint_as_float(nn);
            w.rng[i] = make_uint4(static_cast<uint32_t>(rng.state), static_cast<uint32_t>(rng.state >> 32),
                                  static_cast<uint32_t>(rng.inc), static_cast<uint32_t>(rng.inc >> 32));
            m.w = f_pack(cur, finish ? kSlotDone : kSlotAlive, npend);
            w.misc[i].w = m.w;
        }
        // shadow-queue appends, one atomic per wavefront and term index
        for (uint32_t k = 0; k < 3; ++k) {
            const bool has = active && npend > k;
            const uint32_t q = wave_append(&qcs(w, par)[1], has);
            if (has) w.shadow[q] = (i << 2) | k;
        }
    }
    if (overflow) atomicAdd(&counters[kCntActiveOverflow], static_cast<unsigned long long>(overflow));
}

// Fixed-point transmittance (sampling.cuh:548-583, anyhit.cuh:53-66) of every
// queued shadow ray: origin-set chords first, then one walk over entered
// primitives; lanes refill from the shadow queue as their rays end.
__global__ void VP_WF_LB k_wf_shadow(const JobDesc* __restrict__ jobs, WfPool w, uint32_t par, uint32_t fetch_min) {
    extern __shared__ __align__(16) uint32_t s_leafq[];
    const WaveQueue wq = wave_queue(s_leafq);
    const uint32_t lane = __lane_id();
    unsigned int* qc = qcs(w, par);
    const uint32_t n = qc[1];
    const uint32_t P = w.P;
    bool have = false;
    uint32_t e = 0;
    Walk4 w4w = walk4_begin();
    f3 inv = mk(0.0f), oi = mk(0.0f);
    SceneView sc{};
    ShadowRr rr{};
    while (true) {
        const uint32_t idx = wave_fetch(&qc[5], !have, n, fetch_min);
        if (!have && idx < n) {
            const uint32_t ent = w.shadow[idx];
            const uint32_t i = ent >> 2, k = ent & 3u;
            e = k * P + i;
            const float4 po = w.po[e], pd = w.pd[e];
            const f3 o = f3{po.x, po.y, po.z};
            const f3 d = f3{pd.x, pd.y, pd.z};
            const uint32_t bits = __float_as_uint(po.w);
            const uint32_t set = bits & 1u, count = bits >> 1;
            sc = jobs[__float_as_uint(w.dj[i].w)].scene;
            const uint32_t* act = w.slab + (static_cast<size_t>(i) * kSlabBuffers + set) * w.cap;
            rr = rr_unpack(__float_as_uint(pd.w));
            uint64_t acc = 0;
            const uint64_t lim = rr.limit();
            bool full = acc >= lim;
            for (uint32_t c = 0; c < count && !full; ++c) {
                const float tc = inside_chord(load_prim(sc.prims, act[c]), o, d).tau;
                if (tc > 0.0f) {
                    if (tau_add(acc, tc)) acc = kTauLimitFixed;
                    full = acc >= lim;
                }
            }
            lds_store3(wq.ro + 3 * lane, o);
            lds_store3(wq.rd + 3 * lane, d);
            wq.acc[lane] = acc;
            wq.prims[lane] = sc.prims;
            inv = safe_inv(d);
            oi = o * inv;
            w4w = walk4_begin();
            if (full) w4w.n = kWalkEnd;
            have = true;
        }
        wave_sync();
        if (__ballot(have) == 0) break;
        // node phase: crossed leaves go to the wave queue
        uint32_t count = 0;
        while (true) {
            const bool step = have && w4w.n != kWalkEnd;
            const unsigned long long sm = __builtin_amdgcn_ballot_w64(step);
            if (sm == 0) break;
            if (count >= kWaveQ / VP_SHARED_END_DIV && static_cast<uint32_t>(__popcll(sm)) < VP_SHARED_STEP_MIN) break;
            uint32_t lm = 0, first = 0;
            Walk4 nw = w4w;
            if (step)
                lm = walk4_advance(nw, walk4_block_levels(sc.w4_levels),
                                   ray_mask_nf(sc.w4, sc.w4_half, w4w.n, inv, oi, kTMax), first);
            const bool fits = wave_push_blocks(wq, count, lm, first, lane);
            if (step && fits) w4w = nw;
            if (__builtin_amdgcn_ballot_w64(!fits) != 0) break;
        }
        wave_sync();
        // leaf phase, shared by all lanes: each queued leaf's crossing (entry
        // test and whole-chord depth from one transform) goes into its ray's
        // fixed-point sum (order free)
        const unsigned long long am = __builtin_amdgcn_ballot_w64(true);
        const uint32_t nact = static_cast<uint32_t>(__popcll(am));
        for (uint32_t g = lanes_below(am); g < count; g += nact) {
            const uint32_t own = wq.owner[g];
            const PrimRecord pr = load_prim(wq.prims[own], wq.prim[g]);
            const Crossing x = crossing(pr, lds_load3(wq.ro + 3 * own), lds_load3(wq.rd + 3 * own));
            if (x.tau > 0.0f) atomicAdd(&wq.acc[own], tau_fixed(x.tau));
        }
        wave_sync();
        if (have) {
            const unsigned long long acc = wq.acc[lane];
            if (acc >= rr.limit() || w4w.n == kWalkEnd) {
                w.pd[e].w = shadow_rr_T(acc, rr, depth_T(acc));
                have = false;
            }
        }
    }
}

inline uint32_t cdiv(uint64_t a, uint32_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

}  // namespace
}  // namespace vp

using namespace vp;

hipError_t vp_wf_launch_init(const WfPool& w, hipStream_t s) {
    hipLaunchKernelGGL(k_wf_init, dim3(cdiv(w.P, 256)), dim3(256), 0, s, w);
    return hipGetLastError();
}

// Two iterations (parities 0 and 1) of the pipeline.
hipError_t vp_wf_launch_iterations(vp_context* ctx, const WfPool& w, const WfLaunch& L, hipStream_t s) {
    const size_t lds = leaf_queue_bytes(256);
    for (uint32_t par = 0; par < 2; ++par) {
        hipLaunchKernelGGL(k_wf_shade, dim3(cdiv(w.P, 256)), dim3(256), 0, s, ctx->jobs_dev, w, par);
        hipLaunchKernelGGL(k_wf_claim, dim3(1), dim3(64), 0, s, ctx->launch_desc, ctx->path_counter, w, par);
        hipLaunchKernelGGL(k_wf_regen, dim3(L.grid_regen), dim3(256), 0, s, ctx->jobs_dev, ctx->job_begin_dev,
                           ctx->launch_desc, w, par, ctx->counters);
        hipLaunchKernelGGL(k_wf_extend, dim3(L.grid_extend), dim3(256), lds, s, ctx->jobs_dev, w, par, L.fetch_min);
        hipLaunchKernelGGL(k_wf_scatter, dim3(L.grid_scatter), dim3(256), lds, s, ctx->jobs_dev, w, par,
                           ctx->counters);
        hipLaunchKernelGGL(k_wf_shadow, dim3(L.grid_shadow), dim3(256), lds, s, ctx->jobs_dev, w, par, L.fetch_min);
    }
    return hipGetLastError();
}

int vp_wf_occupancy(int which, int* blocks_per_cu) {
    const size_t lds = which == 3 ? 0 : leaf_queue_bytes(256);
    const void* f = which == 0   ? reinterpret_cast<const void*>(k_wf_extend)
                    : which == 1 ? reinterpret_cast<const void*>(k_wf_scatter)
                    : which == 2 ? reinterpret_cast<const void*>(k_wf_shadow)
                                 : reinterpret_cast<const void*>(k_wf_regen);
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 256, lds);
    *blocks_per_cu = nb > 0 ? nb : 1;
    return e == hipSuccess ? 0 : -1;
}
