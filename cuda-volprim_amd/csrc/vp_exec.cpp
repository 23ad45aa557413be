// vp_exec.cpp — execution of job lists on the persistent renderer.
//
// A step renders one slice of samples of every job in a table: the table,
// the slice start offsets and a two-word launch descriptor are staged in
// pinned host memory, copied to the device, and a fixed sequence (counter
// reset, k_render, k_resolve_all) runs over them.  Finished paths add into
// per-pixel accumulators that live across the slices of a job, so slicing
// (the reference's 16-spp batches) does not change the result.
#include <algorithm>
#include <string>

#include "vp_internal.h"

namespace {
vp_status fail(vp_status code, const std::string& msg) {
    vp_set_error(msg);
    return code;
}
}  // namespace

vp_status vp_ensure_slab(vp_context* ctx, size_t words) {
    if (ctx->slab_words >= words) return VP_OK;
    if (ctx->slab) (void)hipFree(ctx->slab);
    ctx->slab = nullptr;
    ctx->slab_words = 0;
    if (hipMalloc(&ctx->slab, words * sizeof(uint32_t)) != hipSuccess)
        return fail(VP_ERR_OOM, "cannot allocate the active-set slab (" + std::to_string(words * 4) + " bytes)");
    ctx->slab_words = words;
    vp_drop_graph(ctx);
    return VP_OK;
}

void vp_drop_graph(vp_context* ctx) {
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    ctx->graph_exec = nullptr;
    for (auto& k : ctx->graph_key) k = 0;
    if (ctx->wf_graph) (void)hipGraphExecDestroy(ctx->wf_graph);
    ctx->wf_graph = nullptr;
    for (auto& k : ctx->wf_graph_key) k = 0;
}

void vp_free_wf(vp_context* ctx) {
    if (ctx->wf_graph) (void)hipGraphExecDestroy(ctx->wf_graph);
    ctx->wf_graph = nullptr;
    for (auto& k : ctx->wf_graph_key) k = 0;
    if (ctx->wf_mem) (void)hipFree(ctx->wf_mem);
    ctx->wf_mem = nullptr;
    ctx->wf = vp::WfPool{};
    if (ctx->wf_done_pinned) (void)hipHostFree(ctx->wf_done_pinned);
    ctx->wf_done_pinned = nullptr;
    for (auto& e : ctx->wf_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
}

// Path pool of P slots with active-set capacity cap (one allocation, carved).
vp_status vp_ensure_wf(vp_context* ctx, uint32_t P, uint32_t cap) {
    if (ctx->wf_mem && ctx->wf.P == P && ctx->wf.cap == cap) return VP_OK;
    if (static_cast<uint64_t>(P) * cap * vp::kSlabBuffers >= (uint64_t(1) << 32))
        return fail(VP_ERR_INVALID, "wavefront pool: pool_paths x active_cap too large");
    vp_free_wf(ctx);
    const size_t p = P;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_ob = carve(16 * p), o_dj = carve(16 * p), o_bt = carve(16 * p), o_lr = carve(16 * p);
    const size_t o_rng = carve(16 * p), o_misc = carve(16 * p);
    const size_t o_pw = carve(48 * p), o_po = carve(48 * p), o_pd = carve(48 * p);
    const size_t o_live = carve(4 * p), o_sh = carve(12 * p), o_free = carve(4 * p), o_qc = carve(64);
    const size_t o_cb = carve(16), o_fin = carve(8);
    const size_t o_slab = carve(static_cast<size_t>(vp::kSlabBuffers) * cap * p * 4);
    const size_t o_ncand = carve(4 * p);
    if (hipMalloc(&ctx->wf_mem, off) != hipSuccess) {
        ctx->wf_mem = nullptr;
        return fail(VP_ERR_OOM, "cannot allocate the wavefront path pool (" + std::to_string(off) + " bytes)");
    }
    char* b = static_cast<char*>(ctx->wf_mem);
    vp::WfPool& w = ctx->wf;
    w.ob = reinterpret_cast<float4*>(b + o_ob);
    w.dj = reinterpret_cast<float4*>(b + o_dj);
    w.bt = reinterpret_cast<float4*>(b + o_bt);
    w.lr = reinterpret_cast<float4*>(b + o_lr);
    w.rng = reinterpret_cast<uint4*>(b + o_rng);
    w.misc = reinterpret_cast<uint4*>(b + o_misc);
    w.pw = reinterpret_cast<float4*>(b + o_pw);
    w.po = reinterpret_cast<float4*>(b + o_po);
    w.pd = reinterpret_cast<float4*>(b + o_pd);
    w.live = reinterpret_cast<uint32_t*>(b + o_live);
    w.shadow = reinterpret_cast<uint32_t*>(b + o_sh);
    w.free_q = reinterpret_cast<uint32_t*>(b + o_free);
    w.qc = reinterpret_cast<unsigned int*>(b + o_qc);
    w.claim_base = reinterpret_cast<unsigned long long*>(b + o_cb);
    w.finished = reinterpret_cast<unsigned long long*>(b + o_fin);
    w.slab = reinterpret_cast<uint32_t*>(b + o_slab);
    w.ncand = reinterpret_cast<uint32_t*>(b + o_ncand);
    w.P = P;
    w.cap = cap;
    // on any failure below, drop the partial pool so the next call starts over
    // instead of finding wf_mem set with the pinned counters or events missing
    if (hipHostMalloc(&ctx->wf_done_pinned, 2 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        ctx->wf_done_pinned = nullptr;
        vp_free_wf(ctx);
        return fail(VP_ERR_HIP, "wavefront pool: cannot allocate the pinned completion counters");
    }
    for (auto& e : ctx->wf_ev) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            e = nullptr;
            vp_free_wf(ctx);
            return fail(VP_ERR_HIP, "wavefront pool: cannot create its events");
        }
    }
    return VP_OK;
}

vp_status vp_ensure_jobs(vp_context* ctx, uint32_t n) {
    if (ctx->jobs_cap >= n && ctx->pinned_cap >= n) return VP_OK;
    n = std::max<uint32_t>(n, 64u);
    vp_drop_graph(ctx);
    if (ctx->jobs_dev) (void)hipFree(ctx->jobs_dev);
    if (ctx->job_begin_dev) (void)hipFree(ctx->job_begin_dev);
    if (ctx->job_done_dev) (void)hipFree(ctx->job_done_dev);
    if (ctx->jobs_pinned) (void)hipHostFree(ctx->jobs_pinned);
    if (ctx->begin_pinned) (void)hipHostFree(ctx->begin_pinned);
    ctx->jobs_dev = nullptr;
    ctx->job_begin_dev = nullptr;
    ctx->job_done_dev = nullptr;
    ctx->jobs_pinned = nullptr;
    ctx->begin_pinned = nullptr;
    ctx->jobs_cap = ctx->pinned_cap = 0;
    VP_HIP_TRY(hipMalloc(&ctx->jobs_dev, sizeof(vp::JobDesc) * n));
    VP_HIP_TRY(hipMalloc(&ctx->job_begin_dev, sizeof(uint64_t) * n));
    VP_HIP_TRY(hipMalloc(&ctx->job_done_dev, sizeof(unsigned long long) * n));
    VP_HIP_TRY(hipHostMalloc(&ctx->jobs_pinned, sizeof(vp::JobDesc) * n, hipHostMallocDefault));
    VP_HIP_TRY(hipHostMalloc(&ctx->begin_pinned, sizeof(uint64_t) * n, hipHostMallocDefault));
    ctx->jobs_cap = ctx->pinned_cap = n;
    return VP_OK;
}

RenderLaunch vp_resolve_launch(vp_context* ctx, const vp_launch_config* cfg, uint32_t active_cap, int count,
                               int measure) {
    vp_launch_config c{};
    if (cfg) c = *cfg;
    else if (ctx->have_tuned) c = ctx->tuned;
    RenderLaunch L{};
    L.block = c.block ? c.block : 256u;
    L.block = std::max<uint32_t>(64u, std::min<uint32_t>(256u, L.block / 64u * 64u));
    int occ = 1;
    vp_render_occupancy(L.block, active_cap, count, &occ);
    uint32_t bpc = c.blocks_per_cu ? std::min<uint32_t>(c.blocks_per_cu, static_cast<uint32_t>(occ))
                                   : static_cast<uint32_t>(occ);
    if (bpc == 0) bpc = 1;
    L.grid = static_cast<uint32_t>(ctx->prop.multiProcessorCount) * bpc;
    L.active_cap = active_cap;
    // a wavefront claims paths once 8 of its lanes are idle (measured best on
    // the scattering workload: 15.9 against 15.5 Msamples/s at 1)
    L.regen_min = std::max<uint32_t>(1u, std::min<uint32_t>(64u, c.regen_min ? c.regen_min : 8u));
    L.count_events = count;
    L.measure = measure;
    // the wavefront pipeline has no event-counting or measuring variant
    L.pipeline = (c.pipeline == VP_PIPELINE_WAVEFRONT && !count && !measure) ? VP_PIPELINE_WAVEFRONT
                                                                              : VP_PIPELINE_MEGAKERNEL;
    L.pool = c.pool_paths ? std::max<uint32_t>(4096u, c.pool_paths / 64u * 64u) : (1u << 20);
    if (L.pipeline == VP_PIPELINE_WAVEFRONT) {
        const uint32_t cus = static_cast<uint32_t>(ctx->prop.multiProcessorCount);
        int oe = 1, os = 1, osh = 1, org = 1;
        vp_wf_occupancy(0, &oe);
        vp_wf_occupancy(1, &os);
        vp_wf_occupancy(2, &osh);
        vp_wf_occupancy(3, &org);
        L.wf.grid_regen = cus * static_cast<uint32_t>(org);
        L.wf.fetch_min = c.regen_min ? L.regen_min : 48u;
        L.wf.grid_extend = cus * static_cast<uint32_t>(oe);
        L.wf.grid_scatter = cus * static_cast<uint32_t>(os);
        L.wf.grid_shadow = cus * static_cast<uint32_t>(osh);
    }
    return L;
}

// Wavefront pipeline resources for launch L: the path pool and the captured
// two-iteration graph (recaptured when the shape, stream or job table moves).
vp_status vp_wf_prepare(vp_context* ctx, const RenderLaunch& L, hipStream_t s) {
    vp_status st = vp_ensure_wf(ctx, L.pool, L.active_cap);
    if (st != VP_OK) return st;
    ctx->wf.job_done = ctx->job_done_dev;
    const uint64_t key[6] = {(uint64_t(L.wf.fetch_min) << 32) | L.pool, L.active_cap,
                             (uint64_t(L.wf.grid_regen) << 32) | L.wf.grid_extend,
                             (uint64_t(L.wf.grid_scatter) << 32) | L.wf.grid_shadow,
                             reinterpret_cast<uint64_t>(ctx->job_done_dev), reinterpret_cast<uint64_t>(s)};
    if (ctx->wf_graph && std::equal(key, key + 6, ctx->wf_graph_key)) return VP_OK;
    if (ctx->wf_graph) (void)hipGraphExecDestroy(ctx->wf_graph);
    ctx->wf_graph = nullptr;
    VP_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const hipError_t e = vp_wf_launch_iterations(ctx, ctx->wf, L.wf, s);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(VP_ERR_HIP,
                    std::string("wavefront graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    const hipError_t e3 = hipGraphInstantiate(&ctx->wf_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) {
        ctx->wf_graph = nullptr;
        return fail(VP_ERR_HIP, std::string("wavefront graph instantiation failed: ") + hipGetErrorString(e3));
    }
    std::copy(key, key + 6, ctx->wf_graph_key);
    return VP_OK;
}

// Wavefront step: stage the tables, reset the pool, replay the captured
// two-iteration graph until every path of the step has retired (checked one
// replay behind, so the device never waits on the host), resolve the means.
static vp_status vp_run_step_wf(vp_context* ctx, const RenderLaunch& L, uint64_t total_paths, uint32_t max_jobs,
                                uint32_t max_npix, hipStream_t s) {
    vp_status st = vp_wf_prepare(ctx, L, s);
    if (st != VP_OK) return st;
    VP_HIP_TRY(hipMemcpyAsync(ctx->jobs_dev, ctx->jobs_pinned, sizeof(vp::JobDesc) * max_jobs, hipMemcpyHostToDevice,
                              s));
    VP_HIP_TRY(hipMemcpyAsync(ctx->job_begin_dev, ctx->begin_pinned, sizeof(uint64_t) * max_jobs,
                              hipMemcpyHostToDevice, s));
    VP_HIP_TRY(hipMemcpyAsync(ctx->launch_desc, ctx->desc_pinned, sizeof(uint64_t) * 3, hipMemcpyHostToDevice, s));
    VP_HIP_TRY(vp_launch_fill_u64(ctx->path_counter, vp::kClaimParts, 0ull, s));
    VP_HIP_TRY(vp_launch_fill_u64(ctx->job_done_dev, max_jobs, 0ull, s));
    VP_HIP_TRY(vp_wf_launch_init(ctx->wf, s));
    // Every replay advances each live path by two bounces or retires it, and
    // refills free slots, so the loop ends; the bound below only guards
    // against a broken invariant.
    const uint64_t max_replays = (total_paths / L.pool + 2) * 1024 + 1024;
    for (uint64_t k = 0;; ++k) {
        if (k > max_replays) return fail(VP_ERR_HIP, "wavefront pipeline did not drain");
        VP_HIP_TRY(hipGraphLaunch(ctx->wf_graph, s));
        VP_HIP_TRY(hipMemcpyAsync(&ctx->wf_done_pinned[k & 1], ctx->wf.finished, sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
        VP_HIP_TRY(hipEventRecord(ctx->wf_ev[k & 1], s));
        if (k > 0) {
            VP_HIP_TRY(hipEventSynchronize(ctx->wf_ev[(k - 1) & 1]));
            if (ctx->wf_done_pinned[(k - 1) & 1] >= total_paths) break;
        }
    }
    VP_HIP_TRY(vp_launch_resolve(ctx, max_jobs, max_npix, s));
    return VP_OK;
}

vp_status vp_run_step(vp_context* ctx, const RenderLaunch& L, uint32_t nslices, uint64_t total_paths,
                      uint32_t max_jobs, uint32_t max_npix, bool use_graph, hipStream_t s) {
    ctx->desc_pinned[0] = nslices;
    ctx->desc_pinned[1] = total_paths;
    ctx->desc_pinned[2] = 0;  // the launch's paths start at 0
    if (L.pipeline == VP_PIPELINE_WAVEFRONT) return vp_run_step_wf(ctx, L, total_paths, max_jobs, max_npix, s);
    auto enqueue = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(ctx->jobs_dev, ctx->jobs_pinned, sizeof(vp::JobDesc) * max_jobs,
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->job_begin_dev, ctx->begin_pinned, sizeof(uint64_t) * max_jobs,
                               hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->launch_desc, ctx->desc_pinned, sizeof(uint64_t) * 3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = vp_launch_step(ctx, L, max_jobs, max_npix, s);
        return e;
    };
    if (!use_graph) {
        VP_HIP_TRY(enqueue());
        return VP_OK;
    }
    const uint64_t key[6] = {L.block, L.grid, (uint64_t(L.active_cap) << 8) | L.regen_min,
                             uint64_t(L.count_events) << 1 | uint64_t(L.measure), (uint64_t(max_jobs) << 32) | max_npix,
                             reinterpret_cast<uint64_t>(s)};
    if (!ctx->graph_exec || !std::equal(key, key + 6, ctx->graph_key)) {
        vp_drop_graph(ctx);
        VP_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const hipError_t e = enqueue();
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(VP_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        }
        const hipError_t e3 = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e3 != hipSuccess) {
            ctx->graph_exec = nullptr;
            return fail(VP_ERR_HIP, std::string("graph instantiation failed: ") + hipGetErrorString(e3));
        }
        std::copy(key, key + 6, ctx->graph_key);
    }
    VP_HIP_TRY(hipGraphLaunch(ctx->graph_exec, s));
    return VP_OK;
}

vp_status vp_run_jobs(vp_context* ctx, std::vector<JobRun>& runs, const vp_launch_config* cfg, uint32_t active_cap,
                      int count, int measure, hipStream_t s, double* kernel_ms, uint64_t* paths) {
    RenderLaunch L = vp_resolve_launch(ctx, cfg, active_cap, count, measure);
    // per-sample output (and adaptive sampling, which folds the per-sample
    // radiance of each batch) is written by the megakernel only
    bool per_sample = false;
    for (auto& r : runs) {
        const bool adaptive = r.desc.adapt_threshold > 0.0f;
        if (adaptive && r.samples_all != nullptr)
            return fail(VP_ERR_INVALID, "adaptive sampling and per-sample output exclude each other");
        per_sample |= r.samples_all != nullptr || adaptive;
    }
    if (per_sample && L.pipeline == VP_PIPELINE_WAVEFRONT) {
        vp_launch_config c{};
        if (cfg) c = *cfg;
        else if (ctx->have_tuned) c = ctx->tuned;
        c.pipeline = VP_PIPELINE_MEGAKERNEL;
        L = vp_resolve_launch(ctx, &c, active_cap, count, measure);
    }
    const size_t lanes = static_cast<size_t>(L.grid) * L.block;
    vp_status st = VP_OK;
    if (L.pipeline == VP_PIPELINE_MEGAKERNEL)
        st = vp_ensure_slab(ctx, vp::kMegaSlabBuffers * static_cast<size_t>(vp::mega_slab_stride(active_cap)) * lanes);
    if (st != VP_OK) return st;
    st = vp_ensure_jobs(ctx, static_cast<uint32_t>(runs.size()));
    if (st != VP_OK) return st;

    // per-pixel accumulators of every job, zeroed once: the slices of a job
    // add into the same sums
    size_t acc_words = 0;
    uint32_t max_npix = 1;
    for (auto& r : runs) {
        const size_t npix = static_cast<size_t>(r.desc.width) * r.desc.height;
        acc_words += npix * vp::kAccumWords;
        max_npix = std::max(max_npix, static_cast<uint32_t>(npix));
    }
    void* acc_buf = nullptr;
    if (hipMalloc(&acc_buf, acc_words * sizeof(unsigned long long)) != hipSuccess)
        return fail(VP_ERR_OOM, "cannot allocate " + std::to_string(acc_words * 8) + " bytes of pixel accumulators");
    struct Free {
        void* p;
        ~Free() {
            if (p) (void)hipFree(p);
        }
    } free_acc{acc_buf};
    // adaptive sampling state of every adaptive job, in one allocation: per
    // pixel the Welford M2 (float4) and the sample count / converged flag,
    // and the per-sample radiance of one batch (the reference's batch of 16
    // unless the launch sets batch_spp); the Welford mean lives in the
    // job's output image
    size_t adapt_bytes = 0;
    for (auto& r : runs) {
        if (!(r.desc.adapt_threshold > 0.0f)) continue;
        if (r.slice_spp == 0) r.slice_spp = kAdaptiveBatch;
        const size_t npix = static_cast<size_t>(r.desc.width) * r.desc.height;
        adapt_bytes += npix * (sizeof(float4) * (1 + std::min(r.slice_spp, r.desc.spp)) + sizeof(uint32_t));
        adapt_bytes = (adapt_bytes + 255) & ~size_t(255);
    }
    void* adapt_buf = nullptr;
    if (adapt_bytes && hipMalloc(&adapt_buf, adapt_bytes) != hipSuccess)
        return fail(VP_ERR_OOM, "cannot allocate " + std::to_string(adapt_bytes) + " bytes of adaptive-sampling state");
    Free free_adapt{adapt_buf};
    {
        size_t off = 0, aoff = 0;
        for (auto& r : runs) {
            const size_t npix = static_cast<size_t>(r.desc.width) * r.desc.height;
            r.desc.accum = static_cast<unsigned long long*>(acc_buf) + off;
            r.desc.mean = r.mean;
            r.desc.samples = nullptr;
            r.desc.done = nullptr;
            r.desc.adapt_m2 = nullptr;
            r.desc.adapt_state = nullptr;
            r.batch = nullptr;
            r.issued = 0;
            off += npix * vp::kAccumWords;
            if (r.desc.adapt_threshold > 0.0f) {
                char* b = static_cast<char*>(adapt_buf) + aoff;
                r.desc.adapt_m2 = reinterpret_cast<float4*>(b);
                r.batch = r.desc.adapt_m2 + npix;
                r.desc.adapt_state =
                    reinterpret_cast<uint32_t*>(r.batch + npix * std::min(r.slice_spp, r.desc.spp));
                aoff += npix * (sizeof(float4) * (1 + std::min(r.slice_spp, r.desc.spp)) + sizeof(uint32_t));
                aoff = (aoff + 255) & ~size_t(255);
                // Welford state from zero (raygen.cuh:30-36 reads zeroed buffers)
                VP_HIP_TRY(vp_launch_fill_u64(reinterpret_cast<unsigned long long*>(r.desc.adapt_m2),
                                              static_cast<uint32_t>(2 * npix), 0ull, s));
                VP_HIP_TRY(vp_launch_fill_u64(reinterpret_cast<unsigned long long*>(r.mean),
                                              static_cast<uint32_t>(2 * npix), 0ull, s));
                VP_HIP_TRY(vp_launch_fill_u32(r.desc.adapt_state, static_cast<uint32_t>(npix), 0u, s));
            }
        }
        // fill kernels (not hipMemsetAsync): the launch sequences below may
        // be captured into graphs by callers
        for (size_t o = 0; o < acc_words; o += (size_t(1) << 30)) {
            const uint32_t n = static_cast<uint32_t>(std::min<size_t>(acc_words - o, size_t(1) << 30));
            VP_HIP_TRY(vp_launch_fill_u64(static_cast<unsigned long long*>(acc_buf) + o, n, 0ull, s));
        }
    }

    uint64_t total_paths_all = 0;
    double render_ms = 0.0;
    uint32_t launches = 0;
    const uint32_t max_jobs = static_cast<uint32_t>(runs.size());
    while (true) {
        uint64_t total = 0;
        uint32_t k = 0;
        for (auto& r : runs) {
            if (r.issued >= r.desc.spp) continue;
            vp::JobDesc J = r.desc;
            const uint32_t slice = r.slice_spp ? r.slice_spp : r.desc.spp;
            J.s_begin = r.issued;
            J.s_count = std::min(slice, r.desc.spp - r.issued);
            const size_t npix = static_cast<size_t>(J.width) * J.height;
            if (r.samples_all) J.samples = r.samples_all + static_cast<size_t>(J.s_begin) * npix;
            if (r.batch) J.samples = r.batch;
            J.path_begin = total;
            J.path_count = static_cast<uint64_t>(J.npad) * J.s_count;
            total += J.path_count;
            ctx->jobs_pinned[k] = J;
            ctx->begin_pinned[k] = J.path_begin;
            r.issued += J.s_count;
            ++k;
        }
        if (k == 0) break;
        VP_HIP_TRY(hipEventRecord(ctx->ev0, s));
        st = vp_run_step(ctx, L, k, total, max_jobs, max_npix, false, s);
        if (st != VP_OK) return st;
        VP_HIP_TRY(hipEventRecord(ctx->ev1, s));
        // the pinned staging is rewritten next round (and the accumulators
        // are freed on return): wait for this step
        VP_HIP_TRY(hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        VP_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        render_ms += ms;
        total_paths_all += total;
        ++launches;
    }
    if (kernel_ms) *kernel_ms = render_ms;
    ctx->last_launches = launches;
    if (paths) *paths = total_paths_all;
    // samples traced per job: every sample, or under adaptive sampling the
    // per-pixel counts
    for (auto& r : runs) {
        const size_t npix = static_cast<size_t>(r.desc.width) * r.desc.height;
        r.traced = static_cast<uint64_t>(npix) * r.desc.spp;
        if (r.desc.adapt_state == nullptr) continue;
        std::vector<uint32_t> state(npix);
        VP_HIP_TRY(hipMemcpyAsync(state.data(), r.desc.adapt_state, npix * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        VP_HIP_TRY(hipStreamSynchronize(s));
        uint64_t t = 0;
        for (uint32_t v : state) t += v & 0x7fffffffu;
        r.traced = t;
    }
    return VP_OK;
}

vp_status vp_read_stats(vp_context* ctx, hipStream_t s, vp_render_stats* stats) {
    unsigned long long c[vp::kNumCounters];
    VP_HIP_TRY(hipMemcpyAsync(c, ctx->counters, sizeof(c), hipMemcpyDeviceToHost, s));
    VP_HIP_TRY(hipStreamSynchronize(s));
    if (stats) {
        stats->active_overflow = c[vp::kCntActiveOverflow];
        stats->hit_overflow = 0;
        stats->max_hits_per_ray = static_cast<uint32_t>(c[vp::kCntMaxHits]);
        stats->max_point_overlap = static_cast<uint32_t>(c[vp::kCntMaxOverlap]);
    }
    return VP_OK;
}

vp_status vp_reset_counters(vp_context* ctx, hipStream_t s) {
    VP_HIP_TRY(vp_launch_fill_u64(ctx->counters, static_cast<uint32_t>(vp::kNumCounters), 0ull, s));
    return VP_OK;
}
