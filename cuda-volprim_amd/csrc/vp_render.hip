// vp_render.hip — the render path on gfx950.
//
// One persistent megakernel traces camera paths; each 64-lane wavefront
// claims new paths with one atomic per wave (ballot + popcount rank) once
// enough of its lanes are idle, so lanes stay busy across uneven path
// lengths.  Per bounce a lane
//   1. takes the argmin of closed-form per-primitive free flights over the
//      primitives containing its origin and, in one stackless walk of the
//      BVH, over every primitive the ray enters (no hit buffer: each
//      primitive's uniform is a counter hash, so the minimum can be taken in
//      any order and the walk prunes boxes beyond the running minimum); at
//      bounce 0 the same walk integrates the analytic direct term;
//   2. rebuilds the active set at the scatter point from the primitives the
//      argmin walk entered, applying the reference's per-primitive exit
//      tests;
//   3. evaluates albedo, samples the HG phase, and runs NEE (balance-MIS of
//      phase and environment sampling, or product RIS) with shadow rays whose
//      optical depth is accumulated in fixed point;
//   4. applies Russian roulette and the throughput guards.
// The argmin and MIS shadow walks share one leaf queue per wavefront: every
// lane walks a ray, and the leaves all rays cross are evaluated by whichever
// lanes are free (see WaveQueue).
// Each path writes its radiance to a per-sample slot; a second kernel folds
// slots into the per-pixel Welford mean in sample order, so the image does
// not depend on which lane ran which path or on what else shared the launch.
//
// Reference (kfernandez31/cuda-volprim): device/entry/raygen.cuh:21-243,
// device/entry/path_stages.cuh:25-226, device/core/sampling.cuh:257-583,
// device/entry/anyhit.cuh:53-86, device/kernels/normalize_accumulator.cu:16-35.
// The megakernel holds a whole path and a walk per lane, so it is tuned for
// registers over occupancy: 4 waves per SIMD (128 VGPRs), a 512-entry shared
// leaf queue in the LDS this frees (16 waves x 9856 bytes of the CU's
// 160 KiB), and the near/far node test.  On the scattering workload (4096
// Gaussians, 512x512, 64 spp): 26.5 Msamples/s with 448 entries, against 23.6
// at 6 waves with a 224-entry queue and the min/max test, 26.0 at 5 waves
// (320 entries) and 24.5 at 3 waves (640 entries); with the leaf-block walks,
// which fill the queue four leaves at a time, 512 entries gained 0.6-1.2 %
// over 448.
#ifndef VP_WAVE_Q
#define VP_WAVE_Q 512
#endif
#ifndef VP_MEGA_NEAR_FAR
#define VP_MEGA_NEAR_FAR 1
#endif
#include "vp_device.h"

namespace vp {
namespace {

// The argmin's origin-set tables (64 set pointers, 64 exit pointers, 64
// prefix sums) borrow wq.tj between leaf phases.
static_assert(kWaveQ * sizeof(float) >= 64 * (2 * sizeof(void*) + sizeof(uint32_t)),
              "the shared queue must hold the argmin's origin-set tables");

struct PathState {
    f3 o, d, beta, L;
    Pcg32 rng;
    uint64_t key;
    uint32_t bounce, job, na, cur, pix, sl;
};

// Minimum waves per SIMD the register allocator targets for the megakernel.
#ifndef VP_RENDER_WAVES_PER_EU
#define VP_RENDER_WAVES_PER_EU 4
#endif

// Argmin free flight over the entered primitives (sampling.cuh:372-412) with
// the wave-shared leaf queue: every active lane of the wavefront walks its
// own ray, and the queued leaves of all rays are shared out among the lanes.
// Each ray's running minimum is an LDS atomic minimum over float bits (all
// candidates are >= 0), so the result does not depend on which lane handles
// which leaf.  Candidates for the rebuild are recorded while t_entry is at
// most the running minimum at the time, which covers every primitive entered
// before the final one; their order is fixed by the sort in the rebuild.  At
// bounce 0 (direct) the walk also sums the direct term's optical depth into
// dacc, and runs over the whole ray until that sum reaches dlimit (the
// tau = 88 saturation or the direct term's roulette level, where T becomes
// 0); from there on it prunes at the running minimum like any other ray.
// ts_origin is the minimum over the origin set; the result keeps it unless a
// walk candidate is strictly smaller.
__device__ __forceinline__ void argmin_walk_shared(const SceneView& sc, f3 o, f3 d, uint64_t key, uint32_t b,
                                                   bool direct, const uint32_t* act, uint32_t na, float* act_exit,
                                                   uint32_t* __restrict__ slab, uint32_t cap, const WaveQueue& wq,
                                                   float& ts, uint32_t& nc, uint64_t dlimit, uint64_t& dacc) {
    const uint32_t lane = __lane_id();
    lds_store3(wq.ro + 3 * lane, o);
    lds_store3(wq.rd + 3 * lane, d);
    wq.key[lane] = key;
    wq.prims[lane] = sc.prims;
    wq.meta[lane] = b | (direct ? 0x80000000u : 0u);
    wq.acc[lane] = 0ull;
    wq.ts[lane] = __float_as_uint(kTMax);
    wq.nc[lane] = 0u;
    wave_sync();
    // The origin set first (sampling.cuh:356-370), for all lanes' rays
    // together: the wavefront's (member, lane) pairs are numbered through a
    // prefix sum of the set sizes and shared out over the lanes.  Per member:
    // its exit (kept in the owner's slab for the rebuild), its free flight
    // into the owner's running minimum, and at bounce 0 its chord into the
    // owner's direct-term optical depth; minima and fixed-point sums do not
    // depend on which lane adds what.  The owners' pointers and prefix sums
    // sit in wq.tj and their lane ids in wq.list, both free here.  (Active
    // sets of the megakernel are contiguous.)
    {
        const unsigned long long m0 = __builtin_amdgcn_ballot_w64(true);
        const uint32_t rk = lanes_below(m0), nl = static_cast<uint32_t>(__popcll(m0));
        uint32_t pre = 0, total = 0;
        for (uint32_t bit = 0; bit < 16u; ++bit) {
            const unsigned long long bm = __builtin_amdgcn_ballot_w64(((na >> bit) & 1u) != 0u);
            pre += lanes_below(bm) << bit;
            total += static_cast<uint32_t>(__popcll(bm)) << bit;
        }
        const uint32_t** set_of = reinterpret_cast<const uint32_t**>(wq.tj);  // [64] by rank
        float** exit_of = reinterpret_cast<float**>(wq.tj) + 64;               // [64] by rank
        uint32_t* pre_of = reinterpret_cast<uint32_t*>(wq.tj) + 4 * 64;       // [64] by rank
        set_of[rk] = act;
        exit_of[rk] = act_exit;
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
            const uint32_t ol = wq.list[lo], k = g - pre_of[lo];
            const uint32_t i = set_of[lo][k];
            const PrimRecord pr = load_prim(wq.prims[ol], i);
            const f3 ro = lds_load3(wq.ro + 3 * ol), rd = lds_load3(wq.rd + 3 * ol);
            const InsideChord x = inside_chord(pr, ro, rd);
            exit_of[lo][k] = x.t1;
            const uint32_t mo = wq.meta[ol];
            const float tau = free_flight_tau(prim_uniform(wq.key[ol], mo & 0x7fffffffu, 0u, i));
            if (tau <= x.tau) {
                const float t = origin_flight(pr, ro, rd, x, tau);
                atomicMin(&wq.ts[ol], t == 0.0f ? 0u : __float_as_uint(t));
            }
            if ((mo >> 31) != 0u && x.tau > 0.0f) atomicAdd(&wq.acc[ol], tau_fixed(x.tau));
        }
    }
    wave_sync();
    const float ts_origin = __uint_as_float(wq.ts[lane]);
    const f3 inv = safe_inv(d), oi = o * inv;
    float tmax = direct && wq.acc[lane] < dlimit ? kTMax : ts_origin;
    Walk4 w = walk4_begin();
    // the first thread of this wavefront, for the owners' candidate buffers
    const uint32_t gbase = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    while (__ballot(w.n != kWalkEnd) != 0) {
        uint32_t count = 0;
        while (true) {
            const bool step = w.n != kWalkEnd;
            const unsigned long long sm = __builtin_amdgcn_ballot_w64(step);
            if (sm == 0) break;
            if (count >= kWaveQ / VP_SHARED_END_DIV && static_cast<uint32_t>(__popcll(sm)) < VP_SHARED_STEP_MIN) break;
            uint32_t lm = 0, first = 0;
            Walk4 nw = w;
            if (step)
                lm = walk4_advance(nw, walk4_block_levels(sc.w4_levels),
                                   mega_ray_mask(sc.w4, sc.w4_half, w.n, inv, oi, tmax), first);
            const bool fits = wave_push_blocks(wq, count, lm, first, lane);
            if (step && fits) w = nw;
            if (__builtin_amdgcn_ballot_w64(!fits) != 0) break;
        }
        wave_sync();
        const unsigned long long am = __builtin_amdgcn_ballot_w64(true);
        const uint32_t nact = static_cast<uint32_t>(__popcll(am));
        // leaf phase in three passes, each at full width.  First the entry
        // tests' early rejects (crossing_setup): the leaves a ray may enter
        // move down to the front of the queue (an entry is only overwritten
        // after it has been read).  Then those leaves' crossings (entry, exit
        // and whole-chord depth from one transform): on a direct ray the
        // depth joins the ray's sum; if the leaf is entered before the
        // running minimum it is a rebuild candidate and its free flight is
        // sampled, which ends inside the chord when the sampled depth is at
        // most the chord's.  Those leaves move down again with their sampled
        // depth, and the third pass inverts their flights into distances.
        const uint32_t rk = lanes_below(am);
        uint32_t ne = 0;
        for (uint32_t base = 0; base < count; base += nact) {
            const uint32_t g = base + rk;
            bool may = false;
            uint32_t own = 0, prim = 0;
            if (g < count) {
                own = wq.owner[g];
                prim = wq.prim[g];
                CrossingSetup cs;
                may = crossing_setup(load_prim(wq.prims[own], prim), lds_load3(wq.ro + 3 * own),
                                     lds_load3(wq.rd + 3 * own), cs);
            }
            const unsigned long long mm = __builtin_amdgcn_ballot_w64(may);
            wave_sync();
            if (may) {
                const uint32_t h = ne + lanes_below(mm);
                wq.prim[h] = prim;
                wq.owner[h] = static_cast<uint8_t>(own);
            }
            ne += static_cast<uint32_t>(__popcll(mm));
        }
        wave_sync();
        uint32_t nk = 0;
        for (uint32_t base = 0; base < ne; base += nact) {
            const uint32_t g = base + rk;
            bool keep = false;
            uint32_t own = 0, prim = 0;
            float tau = 0.0f;
            if (g < ne) {
                own = wq.owner[g];
                prim = wq.prim[g];
                const PrimRecord pr = load_prim(wq.prims[own], prim);
                const Crossing x = crossing(pr, lds_load3(wq.ro + 3 * own), lds_load3(wq.rd + 3 * own));
                if (x.t0 >= 0.0f) {
                    const uint32_t meta = wq.meta[own];
                    if ((meta >> 31) != 0u && x.tau > 0.0f) atomicAdd(&wq.acc[own], tau_fixed(x.tau));
                    // entered before the running minimum: a rebuild candidate.
                    // Beyond it nothing can change (t >= entry > minimum).
                    if (x.t0 <= __uint_as_float(wq.ts[own])) {
                        const uint32_t pos = atomicAdd(&wq.nc[own], 1u);
                        if (pos < cap) {
                            const uint32_t bs = mega_slab_stride(cap);
                            uint32_t* mine = slab + static_cast<size_t>(gbase + own) * kMegaSlabBuffers * bs;
                            mine[2 * bs + pos] = prim;
                            mine[3 * bs + pos] = __float_as_uint(x.t0);
                            mine[4 * bs + pos] = __float_as_uint(x.t1);
                        }
                        tau = free_flight_tau(prim_uniform(wq.key[own], meta & 0x7fffffffu, 1u, prim));
                        keep = tau <= x.tau;
                    }
                }
            }
            const unsigned long long km = __builtin_amdgcn_ballot_w64(keep);
            wave_sync();
            if (keep) {
                const uint32_t h = nk + lanes_below(km);
                wq.prim[h] = prim;
                wq.owner[h] = static_cast<uint8_t>(own);
                wq.tj[h] = tau;
            }
            nk += static_cast<uint32_t>(__popcll(km));
        }
        wave_sync();
        for (uint32_t g = rk; g < nk; g += nact) {
            const uint32_t own = wq.owner[g];
            const f3 ro = lds_load3(wq.ro + 3 * own), rd = lds_load3(wq.rd + 3 * own);
            const PrimRecord pr = load_prim(wq.prims[own], wq.prim[g]);
            const float t = entered_flight(pr, ro, rd, crossing(pr, ro, rd), wq.tj[g]);
            atomicMin(&wq.ts[own], t == 0.0f ? 0u : __float_as_uint(t));
        }
        wave_sync();
        if (!direct || wq.acc[lane] >= dlimit) tmax = __uint_as_float(wq.ts[lane]);
    }
    wave_sync();
    const float wmin = __uint_as_float(wq.ts[lane]);
    ts = wmin < ts_origin ? wmin : ts_origin;
    nc = wq.nc[lane];
    dacc = wq.acc[lane];
}

// launch_desc[0] = number of job slices, launch_desc[1] = end of the
// launch's path range, launch_desc[2] = its start; read from device memory so
// one captured graph serves every step of a job mix.  A wavefront claims new
// paths once at least regen_min of its lanes are idle (or none is busy), with
// one fetch-and-add on the counter of a part of the launch's range (its XCD's
// part first): part start + counter.  A counter may run past its part (later
// claims there get nothing), and each launch has its own counters and range,
// so launches that run at once (the server starts a launch over the paths a
// longer job table adds while the earlier ones drain) hand out every path
// exactly once, without the retries a compare-and-swap on one shared counter
// costs when paths are short.
template <bool kCount>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VP_RENDER_WAVES_PER_EU)))
k_render(const JobDesc* __restrict__ jobs, const uint64_t* __restrict__ job_begin,
         const uint64_t* __restrict__ launch_desc, unsigned long long* __restrict__ path_counter,
         uint32_t* __restrict__ slab, uint32_t cap, uint32_t regen_min,
         unsigned long long* __restrict__ counters, int measure) {
    extern __shared__ __align__(16) uint32_t s_leafq[];  // [blockDim.x][kLeafRow]
    const uint32_t njobs = static_cast<uint32_t>(launch_desc[0]);
    const uint64_t total_paths = launch_desc[1];
    const uint64_t path_lo = launch_desc[2];
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    // the lane's two active-set buffers, contiguous (lanes read and append
    // to their own sets at different times, so each set stays in few lines)
    constexpr uint32_t stride = 1;
    // buffers 0 and 1: active sets (ping-pong); buffer 2: primitives the
    // argmin walk entered before the running minimum (rebuild candidates)
    const uint32_t bs = mega_slab_stride(cap);
    uint32_t* my = slab + static_cast<size_t>(gl) * kMegaSlabBuffers * bs;  // entry k of buffer b at my[b * bs + k]
    const LeafQueue lq = leaf_queue(s_leafq);
    const WaveQueue wq = wave_queue(s_leafq);  // the same LDS, used by the MIS shadow walks
    PathState st;
    bool alive = false;
    bool exhausted = false;
    Ev<kCount> ev;
    ev.reset();
    uint32_t overflow = 0;
    // The launch's paths in kClaimParts parts, part k claimed through
    // path_counter[k]; a wavefront starts with the part of the XCD it runs on
    // (the XCC_ID hardware register, bits 3:0), so each XCD's waves trace a
    // narrower band of the image and its L2 holds more of what they read.
    // Which lane traces which path changes no result.  parts_done: the parts
    // this wavefront has found used up.
    const uint32_t xcd = __builtin_amdgcn_s_getreg((3u << 11) | (0u << 6) | 20u) % kClaimParts;
    const uint64_t part = (total_paths - path_lo + kClaimParts - 1u) / kClaimParts;
    uint32_t parts_done = 0;

    while (true) {
        // ---- claim new paths for idle lanes: one atomic per wavefront and part tried ----
        const bool want = !alive && !exhausted;
        unsigned long long wmask = __ballot(want);
        if (wmask && __popcll(wmask) < regen_min && __ballot(alive) != 0) wmask = 0;
        if (wmask) {
            const uint32_t leader = __ffsll(static_cast<long long>(wmask)) - 1;
            unsigned long long base = 0, got = 0;
            if (lane == leader) {
                const unsigned long long n_want = static_cast<unsigned long long>(__popcll(wmask));
                for (uint32_t t = 0; t < kClaimParts; ++t) {
                    const uint32_t k = (xcd + t) % kClaimParts;
                    if ((parts_done >> k) & 1u) continue;
                    const uint64_t lo_k = path_lo + k * part;
                    const uint64_t hi_k = lo_k + part < total_paths ? lo_k + part : total_paths;
                    if (lo_k < hi_k) {
                        const unsigned long long c =
                            __hip_atomic_fetch_add(path_counter + k, n_want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        const uint64_t b = lo_k + c;
                        if (b < hi_k) {
                            base = b;
                            got = hi_k - b < n_want ? hi_k - b : n_want;
                            break;
                        }
                    }
                    parts_done |= 1u << k;
                }
            }
            base = __shfl(base, leader);
            got = __shfl(got, leader);
            parts_done = __builtin_amdgcn_readfirstlane(__shfl(parts_done, leader));
            if (want) {
                const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
                const uint64_t rank = static_cast<uint64_t>(__popcll(wmask & lt));
                const uint64_t p = base + rank;
                if (rank >= got) {
                    // every part used up; a lane beyond a part's last paths
                    // claims again next round
                    if (got == 0) exhausted = true;
                } else {
                    // ---- path setup (raygen.cuh:72-113) ----
                    const uint32_t j = find_job(job_begin, njobs, p);
                    const JobDesc& J = jobs[j];
                    if (path_begin(J, p - J.path_begin, st.o, st.d, st.rng, st.key, st.pix, st.sl)) {
                        st.beta = mk(1.0f);
                        st.L = mk(0.0f);
                        st.bounce = 0;
                        st.job = j;
                        st.cur = 0;
                        // origin-inside set (sampling.cuh:322-341) by point query
                        uint32_t na = 0, ov = 0, visits = 0, tests = 0;
                        const SceneView sc = J.scene;
                        const f3 o = st.o;
                        walk_point(sc, o, visits, [&](uint32_t prim) {
                            ++tests;
                            if (point_inside(load_prim(sc.prims, prim), o)) {
                                // counters updated without a branch keep them in registers
                                const bool fits = na < cap;
                                if (fits) my[na * stride] = prim;
                                na += fits ? 1u : 0u;
                                overflow += fits ? 0u : 1u;
                                ++ov;
                            }
                        });
                        st.na = na;
                        ev.add(kCntNodeVisits, visits);
                        ev.add(kCntVisitsPoint, visits);
                        ev.add(kCntContain, tests);
                        ev.add(kCntPaths);
                        if (measure) atomicMax(&counters[kCntMaxOverlap], static_cast<unsigned long long>(ov));
                        alive = true;
                    }
                }
            }
        }
        if (__ballot(alive) == 0) {
            if (__ballot(!exhausted) == 0) break;
            continue;
        }
        if (!alive) continue;

        // ---- one bounce (raygen.cuh:115-208) ----
        const JobDesc& J = jobs[st.job];
        const SceneView sc = J.scene;
        const f3 o = st.o, d = st.d;
        const uint32_t b = st.bounce;
        uint32_t* act = my + (st.cur * bs) * stride;
        uint32_t* nxt = my + ((st.cur ^ 1u) * bs) * stride;
        const uint32_t na = st.na;
        bool finish = false;

        // At bounce 0 the analytic direct term needs the transmittance of the
        // camera ray itself, so the timed kernel integrates it during the
        // argmin loops below (the walk then runs over the whole ray instead
        // of stopping at the running minimum): the same chords, summed in the
        // same fixed point as transmittance() does.  The counting build
        // keeps the separate shadow walk, so its counts are per-ray work.
        const bool direct = !kCount && b == 0;
        uint64_t dacc = 0;
        bool dfull = false;

        // argmin over the origin set (sampling.cuh:356-370); the members'
        // exits are kept for the rebuild.  The timed kernel does this for the
        // whole wavefront inside argmin_walk_shared.
        float ts = kTMax;
        float* act_exit = reinterpret_cast<float*>(my + 5 * bs * stride);
        if (kCount || measure) {
            for (uint32_t k = 0; k < na; ++k) {
                const uint32_t i = act[k * stride];
                const PrimRecord pr = load_prim(sc.prims, i);
                const InsideChord x = inside_chord(pr, o, d);
                act_exit[k * stride] = x.t1;
                const float tau = free_flight_tau(prim_uniform(st.key, b, 0u, i));
                if (tau <= x.tau) ts = minf(ts, origin_flight(pr, o, d, x, tau));
                if (direct && !dfull) dfull = tau_add(dacc, x.tau);
            }
        }
        ev.add(kCntArgActive, na);
        ev.add(kCntTrace);
        // argmin over entered primitives (sampling.cuh:372-412) in one walk;
        // primitives entered before the running minimum are kept as rebuild
        // candidates (every primitive entered before the final minimum is
        // among them, because the walk only prunes beyond the running one)
        uint32_t* cnd = my + 2 * bs * stride;
        uint32_t nc = 0;
        // the analytic direct term's factor and its roulette (a no-op unless
        // J.shadow_rr > 0): the walk that integrates its depth needs only the
        // depth where the term becomes 0; both are formed again after the
        // walk (the same values), so they hold no registers during it
        auto direct_rr = [&](f3& kd) {
            kd = st.beta * env_lookup(env_view(J.env), d);
            return shadow_rr(max3(kd), J.shadow_rr, st.key, 0u, 2u);
        };
        uint64_t dwalk = 0;
        if (!kCount && !measure) {
            uint64_t dlimit = kTauLimitFixed;
            if (direct) {
                f3 kd;
                dlimit = direct_rr(kd).limit();
            }
            argmin_walk_shared(sc, o, d, st.key, b, direct, act, na, act_exit, slab, cap, wq, ts, nc, dlimit, dwalk);
        } else {
            const f3 inv = safe_inv(d);
            const bool prune = !measure && !direct;
            float tmax = prune ? ts : kTMax;
            uint32_t hits = 0, visits = 0, leaves = 0, accept = 0;
            walk_ray_q(sc, o, inv, tmax, visits, lq, [&](uint32_t prim) -> bool {
                const PrimRecord pr = load_prim(sc.prims, prim);
                ++leaves;
                const Crossing x = crossing(pr, o, d);
                if (x.t0 < 0.0f) return true;
                ++hits;
                if (direct && !dfull) dfull = tau_add(dacc, x.tau);
                // beyond the running minimum nothing below can change
                // (t >= entry > ts)
                if (x.t0 > ts) return true;
                if (nc < cap) {
                    cnd[nc * stride] = prim;
                    my[(3 * bs + nc) * stride] = __float_as_uint(x.t0);
                    my[(4 * bs + nc) * stride] = __float_as_uint(x.t1);
                }
                ++nc;
                const float tau = free_flight_tau(prim_uniform(st.key, b, 1u, prim));
                if (tau <= x.tau) {
                    ++accept;
                    const float t = entered_flight(pr, o, d, x, tau);
                    if (t < ts) {
                        ts = t;
                        if (prune) tmax = ts;
                    }
                }
                return true;
            });
            ev.add(kCntArgHit, hits);
            ev.add(kCntArgAccept, accept);
            ev.add(kCntNodeVisits, visits);
            ev.add(kCntLeafTests, leaves);
            if (measure) atomicMax(&counters[kCntMaxHits], static_cast<unsigned long long>(hits));
        }

        const EnvView env = env_view(J.env);
        // analytic direct term at bounce 0 (raygen.cuh:148-160)
        if (b == 0) {
            ev.add(kCntShadowRays);
            f3 kd;
            const ShadowRr drr = direct_rr(kd);
            uint64_t a;
            if constexpr (kCount) a = shadow_depth(sc, o, d, act, na, stride, lq, ev, drr.limit());
            else a = measure ? (dfull ? kTauLimitFixed : dacc) : dwalk;
            const float T = shadow_rr_T(a, drr, depth_T(a));
            st.L = st.L + kd * T;
        }

        if (ts >= kTMax) {
            finish = true;  // escape: with NEE + analytic direct nothing more is added
        } else {
            // rebuild the active set at the scatter point (sampling.cuh:422-458)
            const f3 x = o + ts * d;
            uint32_t nn = 0, ia = 0, visits = 0, rb_a = 0, rb_h = 0;
            auto push = [&](uint32_t v) {
                const bool fits = nn < cap;
                if (fits) nxt[nn * stride] = v;
                nn += fits ? 1u : 0u;
                overflow += fits ? 0u : 1u;
            };
            // old member ia stays while the vertex is before its exit (kept
            // from the argmin's origin loop)
            auto keep_member = [&]() {
                ++rb_a;
                const bool k = ts <= act_exit[ia * stride];
                ++ia;
                return k;
            };
            auto drain_below = [&](uint32_t c) {
                while (ia < na) {
                    const uint32_t a = act[ia * stride];
                    if (a >= c) break;
                    if (keep_member()) push(a);
                }
            };
            // a candidate: an old member kept while the vertex is inside it,
            // else an entered primitive kept if the vertex lies on its chord
            // (entry and exit as the argmin recorded them, or computed here
            // for the point query)
            auto visit = [&](uint32_t c, bool stored, float tj, float te) {
                bool keep = false;
                if (ia < na && act[ia * stride] == c) keep = keep_member();
                if (!keep) {
                    ++rb_h;
                    if (!stored) {
                        const Crossing cx = crossing(load_prim(sc.prims, c), o, d);
                        tj = cx.t0;
                        te = cx.t1;
                    }
                    keep = tj >= 0.0f && tj <= ts && ts <= te;
                }
                if (keep) push(c);
            };
            if (nc <= cap && !measure) {
                // candidates from the argmin walk, in ascending index order
                float* ctj = reinterpret_cast<float*>(my + 3 * bs * stride);
                float* cte = reinterpret_cast<float*>(my + 4 * bs * stride);
                for (uint32_t i = 1; i < nc; ++i) {
                    const uint32_t v = cnd[i * stride];
                    const float vt = ctj[i * stride], ve = cte[i * stride];
                    uint32_t j = i;
                    while (j > 0 && cnd[(j - 1) * stride] > v) {
                        cnd[j * stride] = cnd[(j - 1) * stride];
                        ctj[j * stride] = ctj[(j - 1) * stride];
                        cte[j * stride] = cte[(j - 1) * stride];
                        --j;
                    }
                    cnd[j * stride] = v;
                    ctj[j * stride] = vt;
                    cte[j * stride] = ve;
                }
                // Merge with the old members, read four at a time (ids and
                // exits: the buffers are 16-byte aligned and the reads stay
                // inside them) so the loop does not wait on a load per
                // member.  Same tests and the same ascending order as
                // drain_below/visit: an old member stays while the vertex is
                // before its exit; a candidate that is also an old member is
                // tested as a candidate only if the member test fails.
                static_assert(stride == 1, "the merge reads the slab buffers four entries at a time");
                // the next candidate is held in registers (id ~0 once none is left)
                uint32_t ic = 0, c_id = 0xffffffffu;
                float c_tj = 0.0f, c_te = 0.0f;
                auto next_candidate = [&]() {
                    c_id = ic < nc ? cnd[ic] : 0xffffffffu;
                    if (ic < nc) {
                        c_tj = ctj[ic];
                        c_te = cte[ic];
                    }
                    ++ic;
                };
                auto candidate = [&]() {
                    ++rb_h;
                    if (c_tj >= 0.0f && c_tj <= ts && ts <= c_te) push(c_id);
                };
                next_candidate();
                for (uint32_t base = 0; base < na; base += 4u) {
                    const uint4 id4 = *reinterpret_cast<const uint4*>(act + base);
                    const float4 ex4 = *reinterpret_cast<const float4*>(act_exit + base);
                    const uint32_t id[4] = {id4.x, id4.y, id4.z, id4.w};
                    const float ex[4] = {ex4.x, ex4.y, ex4.z, ex4.w};
                    const uint32_t nk = na - base < 4u ? na - base : 4u;
                    for (uint32_t k = 0; k < nk; ++k) {
                        const uint32_t a = id[k];
                        while (c_id < a) {
                            candidate();
                            next_candidate();
                        }
                        ++rb_a;
                        const bool stay = ts <= ex[k];
                        if (c_id == a) {
                            if (stay) push(a);
                            else candidate();
                            next_candidate();
                        } else if (stay) {
                            push(a);
                        }
                    }
                }
                ia = na;
                while (c_id != 0xffffffffu) {
                    candidate();
                    next_candidate();
                }
            } else {
                // more candidates than fit: one BVH point query at the vertex
                walk_point_q(sc, x, visits, lq, [&](uint32_t c) {
                    drain_below(c);
                    visit(c, false, 0.0f, 0.0f);
                });
            }
            while (ia < na) {
                const uint32_t a = act[ia * stride];
                if (keep_member()) push(a);
            }
            ev.add(kCntNodeVisits, visits);
            ev.add(kCntVisitsPoint, visits);
            ev.add(kCntRbActive, rb_a);
            ev.add(kCntRbHit, rb_h);
            st.cur ^= 1u;
            st.na = nn;
            act = nxt;
            if (measure) {
                uint32_t ovl = 0, v2 = 0;
                walk_point(sc, x, v2, [&](uint32_t c) { ovl += point_inside(load_prim(sc.prims, c), x) ? 1u : 0u; });
                atomicMax(&counters[kCntMaxOverlap], static_cast<unsigned long long>(ovl));
            }

            float phase_pdf;
            const f3 wo = phase_sample(J.hg, d, st.rng, &phase_pdf);
            ev.add(kCntScatter);

            // evaluate_albedo (sampling.cuh:257-283)
            f3 aa = mk(0.0f);
            float aw = 0.0f;
            for (uint32_t k = 0; k < nn; ++k) {
                const PrimRecord pr = load_prim(sc.prims, act[k * stride]);
                const float wgt = pr.mass * prim_pdf(pr, x);
                aa = aa + f3{pr.ar, pr.ag, pr.ab} * wgt;
                aw += wgt;
            }
            ev.add(kCntAlb, nn);
            const f3 albedo = aw <= 0.0f ? mk(0.0f) : aa * rcpf(aw);

            // accumulate_nee_direct (path_stages.cuh:148-226); base is also
            // the next bounce's throughput (raygen.cuh:189), set here so the
            // albedo holds no registers during the shadow walks
            const f3 base = st.beta * albedo;
            st.beta = base;
            const f3 luma = f3{kLumR, kLumG, kLumB};
            if (J.use_ris) {
                float wsum = 0.0f;
                f3 ydir = mk(0.0f), yenv = mk(0.0f);
                for (uint32_t c = 0; c < J.ris_candidates; ++c) {
                    float cpdf;
                    const f3 cwo = env_sample(env, st.rng, &cpdf);
                    if (cpdf <= 0.0f) continue;
                    const f3 envc = env_lookup(env, cwo);
                    const float lumc = dot(envc, luma);
                    const float wgt = phase_eval(J.hg, d, cwo) * lumc * rcpf(cpdf);
                    if (!(wgt > 0.0f)) continue;
                    wsum += wgt;
                    if (next_uniform(st.rng) * wsum < wgt) {
                        ydir = cwo;
                        yenv = envc;
                    }
                }
                if (wsum > 0.0f) {
                    const float lumy = dot(yenv, luma);
                    if (lumy > 0.0f) {
                        ev.add(kCntShadowRays);
                        const float Wr = wsum * rcpf(static_cast<float>(J.ris_candidates));
                        // the factor T multiplies; the roulette is drawn again
                        // after the walk (the same values), so it holds no
                        // registers during it
                        const f3 ky = (base * (yenv * rcpf(lumy))) * Wr;
                        uint64_t acc, acc_unused;
                        {
                            const uint64_t lim = shadow_rr(max3(ky), J.shadow_rr, st.key, b, 0u).limit();
                            if constexpr (kCount) acc = shadow_depth(sc, x, ydir, act, nn, stride, lq, ev, lim);
                            else transmittance_pool(sc, x, ydir, ydir, act, nn, stride, wq, lim, 0ull, acc, acc_unused,
                                                    false);
                        }
                        const ShadowRr rr = shadow_rr(max3(ky), J.shadow_rr, st.key, b, 0u);
                        st.L = st.L + ky * shadow_rr_T(acc, rr, depth_T(acc));
                    }
                }
            } else {
                float apdf;
                const f3 awo = phase_sample(J.hg, d, st.rng, &apdf);
                const float pb_a = env_pdf(env, awo);
                const float sa = apdf + pb_a;
                const float wa = sa > 0.0f ? apdf / sa : 0.0f;
                const f3 enva = env_lookup(env, awo);

                float bpdf;
                const f3 bwo = env_sample(env, st.rng, &bpdf);
                const float ph_b = phase_eval(J.hg, d, bwo);
                const float sb = bpdf + ph_b;
                const float wb = sb > 0.0f ? bpdf / sb : 0.0f;
                const f3 envb = env_lookup(env, bwo);
                ev.add(kCntShadowRays, 2);
                // the factors the two transmittances multiply, and each ray's
                // roulette (a no-op unless J.shadow_rr > 0); the roulettes
                // are drawn again after the walks (the same values), so they
                // hold no registers during them
                const f3 ka = (base * enva) * wa;
                const f3 kb = (((base * envb) * wb) * ph_b) * rcpf(bpdf);
                // the counting build walks the two rays with separate calls,
                // so its counts are per-ray work
                uint64_t acc_a, acc_b;
                {
                    const uint64_t lim_a = shadow_rr(max3(ka), J.shadow_rr, st.key, b, 0u).limit();
                    const uint64_t lim_b = shadow_rr(max3(kb), J.shadow_rr, st.key, b, 1u).limit();
                    if constexpr (kCount) {
                        acc_a = shadow_depth(sc, x, awo, act, nn, stride, lq, ev, lim_a);
                        acc_b = shadow_depth(sc, x, bwo, act, nn, stride, lq, ev, lim_b);
                    } else {
                        transmittance_pool(sc, x, awo, bwo, act, nn, stride, wq, lim_a, lim_b, acc_a, acc_b);
                    }
                }
                const float Ta = shadow_rr_T(acc_a, shadow_rr(max3(ka), J.shadow_rr, st.key, b, 0u), depth_T(acc_a));
                const float Tb = shadow_rr_T(acc_b, shadow_rr(max3(kb), J.shadow_rr, st.key, b, 1u), depth_T(acc_b));
                st.L = st.L + ka * Ta;
                st.L = st.L + kb * Tb;
            }

            // russian_roulette (path_stages.cuh:25-37) and guards (raygen.cuh:191-204)
            if (b >= J.rr_depth) {
                const float ps = minf(J.rr_max_survival, max3(st.beta));
                if (next_uniform(st.rng) > ps || ps <= 0.0f) finish = true;
                else st.beta = st.beta * rcpf(ps);
            }
            if (!finish && max3(st.beta) < kMinThroughput) finish = true;
            if (!finish && !isfinitef_(sum3(st.beta))) finish = true;
            if (!finish) {
                st.o = x;
                st.d = wo;
                st.bounce = b + 1;
                if (st.bounce >= J.max_depth) finish = true;
            }
        }

        if (finish) {
            path_finish(J, st.pix, st.sl, st.L);
            alive = false;
            ev.flush(counters);
        }
    }
    if (overflow) atomicAdd(&counters[kCntActiveOverflow], static_cast<unsigned long long>(overflow));
}

// Adaptive sampling, after a batch (the slice of J): the batch's samples of
// pixel i, in sample order, are folded into its Welford mean and M2 as the
// reference's raygen does per sample (raygen.cuh:214-233: n_inv =
// 1/(prev + s + 1), delta1 = x - mean, mean += delta1 n_inv, M2 += delta1
// (x - mean)), and the convergence test the reference runs at the start of
// the next batch (raygen.cuh:53-69: n >= min samples and max over channels
// of sqrt(M2/(n-1)) / max(mean, 1e-6) below the threshold) sets bit 31 of
// the state, so the next batch skips the pixel.  A converged pixel traced
// nothing in this batch and keeps its state.
__device__ __forceinline__ void fold_pixel(const JobDesc& J, uint32_t i) {
    const uint32_t st = J.adapt_state[i];
    if ((st >> 31) != 0u) return;
    const size_t npix = static_cast<size_t>(J.width) * J.height;
    float4 m = J.mean[i], m2 = J.adapt_m2[i];
    float mean[3] = {m.x, m.y, m.z}, M2[3] = {m2.x, m2.y, m2.z};
    for (uint32_t k = 0; k < J.s_count; ++k) {
        const float4 v = J.samples[k * npix + i];
        const float x[3] = {v.x, v.y, v.z};
        const float n_inv = 1.0f / static_cast<float>(st + k + 1u);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float d1 = x[c] - mean[c];
            mean[c] = mean[c] + d1 * n_inv;
            M2[c] = M2[c] + d1 * (x[c] - mean[c]);
        }
    }
    const uint32_t n = st + J.s_count;
    bool conv = false;
    if (n >= J.adapt_min) {
        const float r = 1.0f / static_cast<float>(n - 1u);
        float err = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float e = sqrtf_(M2[c] * r) / maxf(mean[c], 1e-6f);
            err = c == 0 ? e : maxf(err, e);
        }
        conv = err < J.adapt_threshold;
    }
    J.mean[i] = make_float4(mean[0], mean[1], mean[2], 1.0f);
    J.adapt_m2[i] = make_float4(M2[0], M2[1], M2[2], 0.0f);
    J.adapt_state[i] = n | (conv ? 0x80000000u : 0u);
}

// The mean of every pixel of job-table entry J from its accumulator
// (raygen.cuh:214-238 keeps a running mean; here the sums are exact and the
// division happens once).  Alpha is 1, as the reference writes it.  Under
// adaptive sampling the batch is folded instead.
__device__ __forceinline__ void resolve_pixel(const JobDesc& J, uint32_t i) {
    if (J.adapt_state != nullptr) {
        fold_pixel(J, i);
        return;
    }
    const unsigned long long* a = J.accum + static_cast<size_t>(i) * kAccumWords;
    const uint64_t n = a[3];
    J.mean[i] = make_float4(accum_mean(a[0], a[4], n), accum_mean(a[1], a[5], n), accum_mean(a[2], a[6], n), 1.0f);
}

// Resolve every entry of the launch's job table: blockIdx.y picks the entry;
// blocks beyond the table's length or the entry's pixels exit.  Entries that
// are slices of one job (the same accumulator) write the same values.
__global__ void k_resolve_all(const JobDesc* __restrict__ jobs, const uint64_t* __restrict__ launch_desc) {
    const uint32_t j = blockIdx.y;
    if (j >= static_cast<uint32_t>(launch_desc[0])) return;
    const JobDesc& J = jobs[j];
    const uint32_t npix = J.width * J.height;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) resolve_pixel(J, i);
}

__global__ void k_resolve_job(const JobDesc* __restrict__ jobs, uint32_t j) {
    const JobDesc& J = jobs[j];
    const uint32_t npix = J.width * J.height;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) resolve_pixel(J, i);
}

__global__ void k_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// normalize_accumulator_kernel (device/kernels/normalize_accumulator.cu:16-35)
__global__ void k_normalize(float4* __restrict__ out, const float4* __restrict__ in, size_t n, float scale) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const float4 v = in[i];
        out[i] = make_float4(v.x * scale, v.y * scale, v.z * scale, v.w * scale);
    }
}

// Transmittance along fixed rays, origin set = every primitive containing the
// origin (the bounce-0 scan), no capacity limit.
__global__ void k_transmittance_rays(SceneView sc, const float* __restrict__ rays, uint32_t m, float* __restrict__ T) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const f3 o = f3{rays[6 * r], rays[6 * r + 1], rays[6 * r + 2]};
    const f3 d = f3{rays[6 * r + 3], rays[6 * r + 4], rays[6 * r + 5]};
    uint64_t acc = 0;
    bool full = false;
    uint32_t visits = 0;
    walk_point(sc, o, visits, [&](uint32_t prim) {
        const PrimRecord pr = load_prim(sc.prims, prim);
        if (full || !point_inside(pr, o)) return;
        if (tau_add(acc, inside_chord(pr, o, d).tau)) full = true;
    });
    if (!full) {
        const f3 inv = safe_inv(d);
        const float tmax = kTMax;
        walk_ray(sc, o, inv, tmax, visits, [&](uint32_t prim) -> bool {
            const Crossing x = crossing(load_prim(sc.prims, prim), o, d);
            if (x.t0 < 0.0f) return true;
            if (tau_add(acc, x.tau)) {
                full = true;
                return false;
            }
            return true;
        });
    }
    T[r] = full ? 0.0f : tau_to_T(acc);
}

// Entry list of each ray (collect_hits, sampling.cuh:287-305) sorted by (t, index).
__global__ void k_overlaps(SceneView sc, const float* __restrict__ rays, uint32_t m, uint32_t cap,
                           uint32_t* __restrict__ counts, float* __restrict__ tout, uint32_t* __restrict__ iout) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const f3 o = f3{rays[6 * r], rays[6 * r + 1], rays[6 * r + 2]};
    const f3 d = f3{rays[6 * r + 3], rays[6 * r + 4], rays[6 * r + 5]};
    float* tr = tout + static_cast<size_t>(r) * cap;
    uint32_t* ir = iout + static_cast<size_t>(r) * cap;
    uint32_t c = 0, visits = 0;
    const f3 inv = safe_inv(d);
    const float tmax = kTMax;
    walk_ray(sc, o, inv, tmax, visits, [&](uint32_t prim) -> bool {
        const float tj = entry_t(load_prim(sc.prims, prim), o, d);
        if (tj < 0.0f) return true;
        if (c < cap) {
            uint32_t k = c;
            while (k > 0 && (tr[k - 1] > tj || (tr[k - 1] == tj && ir[k - 1] > prim))) {
                tr[k] = tr[k - 1];
                ir[k] = ir[k - 1];
                --k;
            }
            tr[k] = tj;
            ir[k] = prim;
        }
        ++c;
        return true;
    });
    counts[r] = c;
}

__global__ void k_math(int fn, const float* __restrict__ x, const float* __restrict__ y, uint32_t n,
                       float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = x[i], b = y[i];
    float r;
    switch (fn) {
        case 0: r = expf_(a); break;
        case 1: r = logf_(a); break;
        case 2: r = sinf_(a); break;
        case 3: r = cosf_(a); break;
        case 4: r = atan2f_(a, b); break;
        case 5: r = acosf_(a); break;
        case 6: r = erff_(a); break;
        case 7: r = erfcf_(a); break;
        case 8: r = erfinvf_(a); break;
        case 9: r = erf_diff_(a, b); break;
        case 10: r = sqrtf_(a); break;
        case 11: r = a / b; break;
        case 12: r = chord_poly_(a); break;
        default: r = __builtin_nanf("");
    }
    out[i] = r;
}

__global__ void k_prim(int fn, const PrimRecord* __restrict__ rec, const float* __restrict__ args, uint32_t m,
                       float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const PrimRecord p = *rec;
    const float* a = args + 9 * i;
    const f3 o = f3{a[0], a[1], a[2]}, d = f3{a[3], a[4], a[5]};
    float r;
    switch (fn) {
        case 0: r = inv_cdf(p, o, d, a[8]); break;
        case 1: r = inv_cdf_span(p, o, d, a[6], a[8]); break;
        case 2: r = optical_depth(p, o, d, a[6], a[7]); break;
        case 3: r = exit_from_inside(p, o, d); break;
        case 4: r = crossing(p, o, d).t0; break;
        case 5: {
            const Crossing x = crossing(p, o, d);
            r = x.t0 >= 0.0f ? x.t1 : -1.0f;
        } break;
        case 6: r = point_inside(p, o) ? 1.0f : 0.0f; break;
        case 7: r = prim_pdf(p, o); break;
        case 8: r = crossing(p, o, d).tau; break;
        case 9: r = entered_flight(p, o, d, crossing(p, o, d), a[8]); break;
        case 10: r = inside_chord(p, o, d).tau; break;
        case 11: r = origin_flight(p, o, d, inside_chord(p, o, d), a[8]); break;
        default: r = __builtin_nanf("");
    }
    out[i] = r;
}

inline uint32_t cdiv(uint64_t a, uint32_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

}  // namespace
}  // namespace vp

using namespace vp;

// One persistent render launch over a job table (device pointers): the
// kernel claims the paths from launch_desc[2] below launch_desc[1], in
// kClaimParts parts with one counter each (counter[0 .. kClaimParts), zero at
// the start).
hipError_t vp_launch_render(const RenderLaunch& L, const JobDesc* jobs, const uint64_t* job_begin,
                            const uint64_t* launch_desc, unsigned long long* counter, uint32_t* slab,
                            unsigned long long* counters, hipStream_t s) {
    const size_t lds = leaf_queue_bytes(L.block);
    if (L.count_events)
        hipLaunchKernelGGL(k_render<true>, dim3(L.grid), dim3(L.block), lds, s, jobs, job_begin, launch_desc, counter,
                           slab, L.active_cap, L.regen_min, counters, L.measure);
    else
        hipLaunchKernelGGL(k_render<false>, dim3(L.grid), dim3(L.block), lds, s, jobs, job_begin, launch_desc, counter,
                           slab, L.active_cap, L.regen_min, counters, L.measure);
    return hipGetLastError();
}

// One step of the persistent renderer over the job table in ctx->jobs_dev:
// reset the path counter (a fill kernel, so the sequence can be captured in a
// graph), trace every path of every job slice, resolve the means.  Job count
// and path total come from ctx->launch_desc.
hipError_t vp_launch_step(vp_context* ctx, const RenderLaunch& L, uint32_t max_jobs, uint32_t max_npix,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, ctx->path_counter, kClaimParts, 0ull);
    const hipError_t e = vp_launch_render(L, ctx->jobs_dev, ctx->job_begin_dev, ctx->launch_desc, ctx->path_counter,
                                          ctx->slab, ctx->counters, s);
    if (e != hipSuccess) return e;
    return vp_launch_resolve(ctx, max_jobs, max_npix, s);
}

hipError_t vp_launch_resolve(vp_context* ctx, uint32_t max_jobs, uint32_t max_npix, hipStream_t s) {
    const uint32_t gx = std::max<uint32_t>(1u, std::min<uint32_t>(cdiv(max_npix, 256), 512u));
    hipLaunchKernelGGL(k_resolve_all, dim3(gx, max_jobs), dim3(256), 0, s, ctx->jobs_dev, ctx->launch_desc);
    return hipGetLastError();
}

hipError_t vp_launch_resolve_job(const JobDesc* jobs, uint32_t j, uint32_t npix, hipStream_t s) {
    const uint32_t gx = std::max<uint32_t>(1u, std::min<uint32_t>(cdiv(npix, 256), 1024u));
    hipLaunchKernelGGL(k_resolve_job, dim3(gx), dim3(256), 0, s, jobs, j);
    return hipGetLastError();
}

hipError_t vp_launch_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(n, 64)), dim3(64), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t vp_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

int vp_render_occupancy(uint32_t block, uint32_t, int count_events, int* blocks_per_cu) {
    int nb = 0;
    hipError_t e;
    const size_t lds = leaf_queue_bytes(block);
    if (count_events)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(k_render<true>), block, lds);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(k_render<false>), block, lds);
    *blocks_per_cu = nb > 0 ? nb : 1;
    return e == hipSuccess ? 0 : -1;
}

hipError_t vp_launch_normalize(float4* out, const float4* in, size_t n, float scale, hipStream_t s) {
    const uint32_t grid = std::min<uint32_t>(cdiv(n, 256), 1024u);
    hipLaunchKernelGGL(k_normalize, dim3(grid ? grid : 1), dim3(256), 0, s, out, in, n, scale);
    return hipGetLastError();
}

hipError_t vp_launch_transmittance(SceneView sc, const float* rays, uint32_t m, float* T, hipStream_t s) {
    hipLaunchKernelGGL(k_transmittance_rays, dim3(cdiv(m, 128)), dim3(128), 0, s, sc, rays, m, T);
    return hipGetLastError();
}

hipError_t vp_launch_overlaps(SceneView sc, const float* rays, uint32_t m, uint32_t cap, uint32_t* counts, float* t,
                              uint32_t* idx, hipStream_t s) {
    hipLaunchKernelGGL(k_overlaps, dim3(cdiv(m, 128)), dim3(128), 0, s, sc, rays, m, cap, counts, t, idx);
    return hipGetLastError();
}

hipError_t vp_launch_math(int fn, const float* x, const float* y, uint32_t n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_math, dim3(cdiv(n, 256)), dim3(256), 0, s, fn, x, y, n, out);
    return hipGetLastError();
}

hipError_t vp_launch_prim(int fn, const PrimRecord* rec, const float* args, uint32_t m, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_prim, dim3(cdiv(m, 256)), dim3(256), 0, s, fn, rec, args, m, out);
    return hipGetLastError();
}
