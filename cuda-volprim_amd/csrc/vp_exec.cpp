// vp_exec.cpp — execution of job lists on the persistent renderer.
//
// A step renders one slice of samples of every job in a table: the table,
// the slice start offsets and a two-word launch descriptor are staged in
// pinned host memory, copied to the device, and a fixed sequence (counter
// reset, k_render, k_welford_all) runs over them.  Because the sequence only
// reads sizes from device memory, it can be captured once as a HIP graph and
// replayed for every later step with the same launch shape (the serving loop
// does this); the plain renderer launches it directly.
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
}

vp_status vp_ensure_jobs(vp_context* ctx, uint32_t n) {
    if (ctx->jobs_cap >= n && ctx->pinned_cap >= n) return VP_OK;
    n = std::max<uint32_t>(n, 64u);
    vp_drop_graph(ctx);
    if (ctx->jobs_dev) (void)hipFree(ctx->jobs_dev);
    if (ctx->job_begin_dev) (void)hipFree(ctx->job_begin_dev);
    if (ctx->jobs_pinned) (void)hipHostFree(ctx->jobs_pinned);
    if (ctx->begin_pinned) (void)hipHostFree(ctx->begin_pinned);
    ctx->jobs_dev = nullptr;
    ctx->job_begin_dev = nullptr;
    ctx->jobs_pinned = nullptr;
    ctx->begin_pinned = nullptr;
    ctx->jobs_cap = ctx->pinned_cap = 0;
    VP_HIP_TRY(hipMalloc(&ctx->jobs_dev, sizeof(vp::JobDesc) * n));
    VP_HIP_TRY(hipMalloc(&ctx->job_begin_dev, sizeof(uint64_t) * n));
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
    L.regen_min = std::max<uint32_t>(1u, std::min<uint32_t>(64u, c.regen_min ? c.regen_min : 1u));
    L.count_events = count;
    L.measure = measure;
    return L;
}

vp_status vp_run_step(vp_context* ctx, const RenderLaunch& L, uint32_t nslices, uint64_t total_paths,
                      uint32_t max_jobs, uint32_t max_npix, bool use_graph, hipStream_t s) {
    ctx->desc_pinned[0] = nslices;
    ctx->desc_pinned[1] = total_paths;
    auto enqueue = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(ctx->jobs_dev, ctx->jobs_pinned, sizeof(vp::JobDesc) * max_jobs,
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->job_begin_dev, ctx->begin_pinned, sizeof(uint64_t) * max_jobs,
                               hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->launch_desc, ctx->desc_pinned, sizeof(uint64_t) * 2, hipMemcpyHostToDevice, s);
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
    const RenderLaunch L = vp_resolve_launch(ctx, cfg, active_cap, count, measure);
    const size_t lanes = static_cast<size_t>(L.grid) * L.block;
    vp_status st = vp_ensure_slab(ctx, 2 * static_cast<size_t>(active_cap) * lanes);
    if (st != VP_OK) return st;
    st = vp_ensure_jobs(ctx, static_cast<uint32_t>(runs.size()));
    if (st != VP_OK) return st;

    // slice sizes: spread the per-launch sample budget over the jobs
    size_t total_pix = 0;
    uint32_t max_npix = 1;
    for (auto& r : runs) {
        total_pix += static_cast<size_t>(r.desc.width) * r.desc.height;
        max_npix = std::max(max_npix, r.desc.width * r.desc.height);
    }
    const size_t budget_samples = std::max<size_t>(1, kSliceBudgetBytes / (16 * std::max<size_t>(total_pix, 1)));
    size_t slice_bytes = 0;
    for (auto& r : runs) {
        if (r.slice_spp == 0) r.slice_spp = static_cast<uint32_t>(std::min<size_t>(r.desc.spp, budget_samples));
        if (!r.samples_all)
            slice_bytes += static_cast<size_t>(r.slice_spp) * r.desc.width * r.desc.height * sizeof(float4);
    }
    void* slice_buf = nullptr;
    if (slice_bytes && hipMalloc(&slice_buf, slice_bytes) != hipSuccess)
        return fail(VP_ERR_OOM, "cannot allocate " + std::to_string(slice_bytes) + " bytes of sample slots");
    struct Free {
        void* p;
        ~Free() {
            if (p) (void)hipFree(p);
        }
    } free_slice{slice_buf};

    uint64_t total_paths_all = 0;
    double render_ms = 0.0;
    uint32_t launches = 0;
    const uint32_t max_jobs = static_cast<uint32_t>(runs.size());
    while (true) {
        uint64_t total = 0;
        size_t off = 0;
        uint32_t k = 0;
        for (auto& r : runs) {
            if (r.done >= r.desc.spp) continue;
            vp::JobDesc J = r.desc;
            J.s_begin = r.done;
            J.s_count = std::min(r.slice_spp, r.desc.spp - r.done);
            J.path_begin = total;
            J.path_count = static_cast<uint64_t>(J.npad) * J.s_count;
            J.mean = r.mean;
            const size_t npix = static_cast<size_t>(J.width) * J.height;
            if (r.samples_all) {
                J.samples = r.samples_all + static_cast<size_t>(J.s_begin) * npix;
            } else {
                J.samples = reinterpret_cast<float4*>(static_cast<char*>(slice_buf) + off);
                off += static_cast<size_t>(J.s_count) * npix * sizeof(float4);
            }
            total += J.path_count;
            ctx->jobs_pinned[k] = J;
            ctx->begin_pinned[k] = J.path_begin;
            r.done += J.s_count;
            ++k;
        }
        if (k == 0) break;
        VP_HIP_TRY(hipEventRecord(ctx->ev0, s));
        st = vp_run_step(ctx, L, k, total, max_jobs, max_npix, false, s);
        if (st != VP_OK) return st;
        VP_HIP_TRY(hipEventRecord(ctx->ev1, s));
        // the pinned staging is rewritten next round: wait for this step
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
