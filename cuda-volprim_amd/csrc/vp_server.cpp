// vp_server.cpp — the serving form and the launch autotuner.
//
// Serving: a request loop on its own thread keeps scenes and environment
// tables resident (requests reference vp_scene / vp_env handles created once)
// and batches continuously: before every step it admits queued requests while
// the bytes they need on the device (Welford image + one slice of per-sample
// radiance) fit the budget and a slot is free; the step then advances every
// admitted request by one slice of samples in a single persistent launch
// (replayed from a captured HIP graph), and requests whose last slice landed
// are copied out and retired, freeing their slot and bytes for the queue.
// Each request's image depends only on its own (seed, pixel, sample) paths
// and on folding its slices in sample order, so it is the same alone as in
// any mix.
//
// Autotuning: vp_autotune times a reduced copy of a job mix for a grid of
// launch shapes (megakernel: workgroup size, persistent workgroups per CU,
// regeneration threshold; wavefront pipeline: path-pool size) and sizes the active-set capacity from the measured maximum
// point overlap (ceil(1.125 v) rounded up to 16, the rule of the reference's
// --measure-caps, src/thesis/host/app/renderer.cpp:405-418, and at least the
// largest capacity the jobs request); the best shape
// is kept in the context and used by every later launch that passes no
// explicit configuration.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vp_internal.h"

namespace {
using clk = std::chrono::steady_clock;

vp_status fail(vp_status code, const std::string& msg) {
    vp_set_error(msg);
    return code;
}

// Job-table entries one server epoch can hold before it drains and restarts.
constexpr uint32_t kServerTable = 4096;

struct Request {
    uint64_t ticket = 0;
    vp_job job{};
    vp::JobDesc desc{};
    float4* mean = nullptr;
    float4* samples = nullptr;
    size_t bytes = 0;
    uint32_t entry = 0;     // job-table entry in the current epoch
    bool completing = false;
    hipEvent_t ev = nullptr;  // recorded after the request's Welford fold
    clk::time_point t_submit;
    double latency_ms = 0.0;
    bool done = false;
    vp_status status = VP_OK;
};
}  // namespace

struct vp_server {
    vp_context* ctx = nullptr;
    vp_server_config cfg{};
    RenderLaunch L{};
    uint32_t max_npix = 0;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<Request*> pending;
    std::vector<Request*> inflight;
    std::unordered_map<uint64_t, Request*> by_ticket;
    uint64_t next_ticket = 1;
    size_t bytes_inflight = 0;
    bool stop = false;
    std::vector<double> latencies;
    uint64_t completed = 0, samples = 0, steps = 0;
    double busy_ms = 0.0;
    std::string error;
    // epoch state, server thread only
    uint32_t njobs = 0;
    uint64_t total = 0;
    unsigned long long* done_pinned = nullptr;  // [2][kServerTable] per-entry retired paths
    uint64_t* desc_pinned = nullptr;            // [2][2] launch descriptors
    uint32_t njobs_copied[2] = {0, 0};
    hipEvent_t rep_ev[2] = {nullptr, nullptr};
    hipStream_t copy_stream = nullptr;
};

namespace {

void release_request(vp_server* S, Request* r, hipStream_t s) {
    if (r->mean) (void)hipFreeAsync(r->mean, s);
    if (r->samples) (void)hipFreeAsync(r->samples, s);
    r->mean = r->samples = nullptr;
    if (r->ev) (void)hipEventDestroy(r->ev);
    r->ev = nullptr;
    (void)S;
}

// Finish request r (under S->mu): record its latency and wake pollers.
void finish_request(vp_server* S, Request* r, vp_status st) {
    r->status = st;
    r->latency_ms = std::chrono::duration<double, std::milli>(clk::now() - r->t_submit).count();
    r->done = true;
    S->bytes_inflight -= r->bytes;
    S->latencies.push_back(r->latency_ms);
    ++S->completed;
    S->samples += static_cast<uint64_t>(r->job.params.width) * r->job.params.height * r->job.params.spp;
    S->inflight.erase(std::find(S->inflight.begin(), S->inflight.end(), r));
    S->cv_done.notify_all();
}

// Continuous batching over the wavefront pipeline.  Each admitted request
// becomes one job-table entry (all its samples) appended between two graph
// replays; the pipeline's free path slots pick its paths up at the next
// iteration, next to whatever else is in flight.  A request is complete
// when its entry's retired-path count reaches its path count; its samples
// are then folded and copied out while the pipeline keeps running.
void server_loop(vp_server* S) {
    vp_context* ctx = S->ctx;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    uint64_t k = 0;
    bool fresh = true;  // the pool must be reset before the next replay
    clk::time_point busy_since{};
    std::vector<Request*> admit;
    while (true) {
        admit.clear();
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->inflight.empty() && (S->pending.empty() || S->njobs >= kServerTable)) {
                // nothing on the GPU: start a new epoch
                if (!fresh) S->busy_ms += std::chrono::duration<double, std::milli>(clk::now() - busy_since).count();
                S->njobs = 0;
                S->total = 0;
                fresh = true;
                k = 0;
                if (S->pending.empty()) {
                    if (S->stop) return;
                    S->cv.wait(lk);
                    continue;
                }
            }
            // admission: FIFO, by device bytes in flight and free slots
            while (!S->pending.empty() && S->inflight.size() < S->cfg.max_inflight && S->njobs < kServerTable) {
                Request* r = S->pending.front();
                if (!S->inflight.empty() && S->bytes_inflight + r->bytes > S->cfg.mem_budget_bytes) break;
                S->pending.pop_front();
                S->inflight.push_back(r);
                S->bytes_inflight += r->bytes;
                r->entry = S->njobs++;
                admit.push_back(r);
            }
        }
        vp_status st = VP_OK;
        if (fresh) {
            // the previous epoch's staging copies must have run before its
            // pinned entries are rewritten
            if (hipStreamSynchronize(s) != hipSuccess) st = VP_ERR_HIP;
            busy_since = clk::now();
            if (st == VP_OK) st = vp_wf_prepare(ctx, S->L, s);
            if (st == VP_OK && (vp_launch_fill_u64(ctx->path_counter, 1u, 0ull, s) != hipSuccess ||
                                vp_wf_launch_init(ctx->wf, s) != hipSuccess))
                st = VP_ERR_HIP;
            fresh = false;
        }
        // install the admitted requests
        for (Request* r : admit) {
            const size_t npix = static_cast<size_t>(r->job.params.width) * r->job.params.height;
            vp::JobDesc J = r->desc;
            J.s_begin = 0;
            J.s_count = J.spp;
            J.path_begin = S->total;
            J.path_count = static_cast<uint64_t>(J.npad) * J.spp;
            if (st == VP_OK && (hipMallocAsync(reinterpret_cast<void**>(&r->mean), npix * sizeof(float4), s) != hipSuccess ||
                                hipMallocAsync(reinterpret_cast<void**>(&r->samples),
                                               npix * J.spp * sizeof(float4), s) != hipSuccess))
                st = VP_ERR_OOM;
            J.mean = r->mean;
            J.samples = r->samples;
            r->desc = J;
            ctx->jobs_pinned[r->entry] = J;
            ctx->begin_pinned[r->entry] = J.path_begin;
            if (st == VP_OK &&
                (hipMemcpyAsync(ctx->jobs_dev + r->entry, ctx->jobs_pinned + r->entry, sizeof(vp::JobDesc),
                                hipMemcpyHostToDevice, s) != hipSuccess ||
                 hipMemcpyAsync(ctx->job_begin_dev + r->entry, ctx->begin_pinned + r->entry, sizeof(uint64_t),
                                hipMemcpyHostToDevice, s) != hipSuccess ||
                 vp_launch_fill_u64(ctx->job_done_dev + r->entry, 1u, 0ull, s) != hipSuccess))
                st = VP_ERR_HIP;
            S->total += J.path_count;
        }
        if (!admit.empty() && st == VP_OK) {
            uint64_t* d = S->desc_pinned + 2 * (k & 1);
            d[0] = S->njobs;
            d[1] = S->total;
            if (hipMemcpyAsync(ctx->launch_desc, d, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, s) != hipSuccess)
                st = VP_ERR_HIP;
        }
        // one replay (two iterations), then the per-entry progress snapshot
        if (st == VP_OK && hipGraphLaunch(ctx->wf_graph, s) != hipSuccess) st = VP_ERR_HIP;
        S->njobs_copied[k & 1] = S->njobs;
        if (st == VP_OK &&
            (hipMemcpyAsync(S->done_pinned + static_cast<size_t>(k & 1) * kServerTable, ctx->job_done_dev,
                            sizeof(unsigned long long) * S->njobs, hipMemcpyDeviceToHost, s) != hipSuccess ||
             hipEventRecord(S->rep_ev[k & 1], s) != hipSuccess))
            st = VP_ERR_HIP;
        if (st != VP_OK) {
            std::lock_guard<std::mutex> lk(S->mu);
            S->error = "server step failed";
            while (!S->inflight.empty()) {
                Request* r = S->inflight.back();
                release_request(S, r, s);
                finish_request(S, r, st);
            }
            continue;
        }
        ++S->steps;
        if (k > 0) {
            // progress as of the previous replay (the current one keeps running)
            const uint64_t pk = k - 1;
            (void)hipEventSynchronize(S->rep_ev[pk & 1]);
            const unsigned long long* dn = S->done_pinned + static_cast<size_t>(pk & 1) * kServerTable;
            const uint32_t seen = S->njobs_copied[pk & 1];
            std::vector<Request*> snapshot;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                snapshot = S->inflight;
            }
            for (Request* r : snapshot) {
                if (!r->completing && r->entry < seen && dn[r->entry] >= r->desc.path_count) {
                    const uint32_t npix = r->desc.width * r->desc.height;
                    vp_status rs = VP_OK;
                    if (vp_launch_welford_job(ctx, r->entry, npix, s) != hipSuccess ||
                        hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess ||
                        hipEventRecord(r->ev, s) != hipSuccess)
                        rs = VP_ERR_HIP;
                    r->completing = true;
                    if (rs != VP_OK) {
                        release_request(S, r, s);
                        std::lock_guard<std::mutex> lk(S->mu);
                        finish_request(S, r, rs);
                    }
                } else if (r->completing && hipEventQuery(r->ev) == hipSuccess) {
                    const size_t bytes = static_cast<size_t>(r->desc.width) * r->desc.height * sizeof(float4);
                    vp_status rs = VP_OK;
                    if (hipMemcpyAsync(r->job.out_rgba, r->mean, bytes, hipMemcpyDeviceToHost, S->copy_stream) !=
                            hipSuccess ||
                        hipStreamSynchronize(S->copy_stream) != hipSuccess)
                        rs = VP_ERR_HIP;
                    release_request(S, r, s);
                    std::lock_guard<std::mutex> lk(S->mu);
                    finish_request(S, r, rs);
                }
            }
        }
        ++k;
    }
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * static_cast<double>(v.size() - 1);
    const size_t i = static_cast<size_t>(pos);
    const double f = pos - static_cast<double>(i);
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

vp_status vp_server_create(vp_context* ctx, const vp_server_config* cfg, vp_server** out) {
    if (!ctx || !out) return fail(VP_ERR_INVALID, "null argument");
    auto S = std::make_unique<vp_server>();
    S->ctx = ctx;
    if (cfg) S->cfg = *cfg;
    if (S->cfg.mem_budget_bytes == 0) S->cfg.mem_budget_bytes = size_t(8) << 30;
    if (S->cfg.slice_spp == 0) S->cfg.slice_spp = 4;
    if (S->cfg.max_inflight == 0) S->cfg.max_inflight = 64;
    S->max_npix = 1024u * 1024u;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(VP_ERR_HIP, "hipSetDevice failed");
    // the tuned (or default) shape, on the wavefront pipeline
    vp_launch_config c{};
    if (ctx->have_tuned) c = ctx->tuned;
    c.pipeline = VP_PIPELINE_WAVEFRONT;
    const uint32_t cap = c.active_cap ? c.active_cap : 128u;
    S->L = vp_resolve_launch(ctx, &c, cap, 0, 0);
    vp_status st = vp_ensure_jobs(ctx, kServerTable);
    if (st == VP_OK) st = vp_wf_prepare(ctx, S->L, ctx->stream);
    if (st == VP_OK) st = vp_reset_counters(ctx, ctx->stream);
    if (st != VP_OK) return st;
    if (hipHostMalloc(&S->done_pinned, 2 * kServerTable * sizeof(unsigned long long), hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(&S->desc_pinned, 4 * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&S->rep_ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&S->rep_ev[1], hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking) != hipSuccess) {
        if (S->done_pinned) (void)hipHostFree(S->done_pinned);
        if (S->desc_pinned) (void)hipHostFree(S->desc_pinned);
        for (auto& e : S->rep_ev)
            if (e) (void)hipEventDestroy(e);
        if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
        return fail(VP_ERR_HIP, "server resources");
    }
    S->th = std::thread(server_loop, S.get());
    *out = S.release();
    return VP_OK;
}

vp_status vp_server_submit(vp_server* S, const vp_job* job, uint64_t* ticket) {
    if (!S || !job || !ticket) return fail(VP_ERR_INVALID, "null argument");
    if (!job->scene || !job->env || !job->out_rgba) return fail(VP_ERR_INVALID, "job has a null field");
    vp_status st = vp_check_params(&job->params);
    if (st != VP_OK) return st;
    if (job->params.width * job->params.height > S->max_npix)
        return fail(VP_ERR_INVALID, "request larger than the server's 1024x1024 pixel limit");
    if (job->params.measure_caps) return fail(VP_ERR_INVALID, "measure_caps is not available through the server");
    if (job->params.active_cap > S->L.active_cap)
        return fail(VP_ERR_INVALID, "request asks for an active-set capacity above the server's (" +
                                        std::to_string(S->L.active_cap) + "): autotune with it first");
    auto* r = new Request();
    r->job = *job;
    r->desc = vp_make_job(job->scene, job->env, &job->camera, &job->params);
    const size_t npix = static_cast<size_t>(job->params.width) * job->params.height;
    r->bytes = npix * sizeof(float4) * (1 + static_cast<size_t>(job->params.spp));
    r->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->ticket = S->next_ticket++;
        S->by_ticket[r->ticket] = r;
        S->pending.push_back(r);
    }
    S->cv.notify_one();
    *ticket = r->ticket;
    return VP_OK;
}

vp_status vp_server_poll(vp_server* S, uint64_t ticket, double* latency_ms) {
    if (!S) return fail(VP_ERR_INVALID, "server is null");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->by_ticket.find(ticket);
    if (it == S->by_ticket.end()) return fail(VP_ERR_INVALID, "unknown ticket");
    Request* r = it->second;
    if (!r->done) return VP_ERR_NOT_READY;
    if (latency_ms) *latency_ms = r->latency_ms;
    const vp_status st = r->status;
    S->by_ticket.erase(it);
    delete r;
    if (st != VP_OK) return fail(st, "request failed");
    return VP_OK;
}

vp_status vp_server_drain(vp_server* S) {
    if (!S) return fail(VP_ERR_INVALID, "server is null");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->pending.empty() && S->inflight.empty(); });
    return S->error.empty() ? VP_OK : fail(VP_ERR_HIP, S->error);
}

vp_status vp_server_stats_get(vp_server* S, vp_server_stats* out) {
    if (!S || !out) return fail(VP_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    out->completed = S->completed;
    out->samples = S->samples;
    out->p50_ms = percentile(S->latencies, 0.50);
    out->p99_ms = percentile(S->latencies, 0.99);
    double sum = 0.0;
    for (double v : S->latencies) sum += v;
    out->mean_ms = S->latencies.empty() ? 0.0 : sum / static_cast<double>(S->latencies.size());
    out->busy_ms = S->busy_ms;
    out->steps = S->steps;
    return VP_OK;
}

vp_status vp_server_destroy(vp_server* S) {
    if (!S) return VP_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        for (Request* r : S->pending) {
            r->status = VP_ERR_STATE;
            r->done = true;
        }
        S->pending.clear();
    }
    S->cv.notify_all();
    // the loop runs until its in-flight requests are done, then exits
    if (S->th.joinable()) S->th.join();
    (void)hipSetDevice(S->ctx->device);
    (void)hipStreamSynchronize(S->ctx->stream);
    for (auto& kv : S->by_ticket) delete kv.second;
    if (S->done_pinned) (void)hipHostFree(S->done_pinned);
    if (S->desc_pinned) (void)hipHostFree(S->desc_pinned);
    for (auto& e : S->rep_ev)
        if (e) (void)hipEventDestroy(e);
    if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
    delete S;
    return VP_OK;
}

vp_status vp_autotune(vp_context* ctx, vp_job* jobs, uint32_t njobs, vp_launch_config* best) {
    if (!ctx || !jobs || njobs == 0) return fail(VP_ERR_INVALID, "no jobs");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(VP_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    // reduced copy of the mix: every view downscaled by an integer factor k
    // to at most kTuneRes pixels a side (same field of view, k-times wider
    // pixels) at one sample per pixel and depth at most 16
    constexpr uint32_t kTuneRes = 64;
    std::vector<vp_render_params> params(njobs);
    std::vector<vp_camera> cams(njobs);
    for (uint32_t i = 0; i < njobs; ++i) {
        vp_status st = vp_check_params(&jobs[i].params);
        if (st != VP_OK) return st;
        params[i] = jobs[i].params;
        params[i].spp = 1;
        params[i].measure_caps = 0;
        // long paths only lengthen each timing run's tail
        params[i].max_depth = std::min<uint32_t>(params[i].max_depth, 16u);
        const uint32_t side = std::max(params[i].width, params[i].height);
        const uint32_t k = (side + kTuneRes - 1) / kTuneRes;
        params[i].width = (params[i].width + k - 1) / k;
        params[i].height = (params[i].height + k - 1) / k;
        vp_camera c = jobs[i].camera;
        const float kf = static_cast<float>(k), h = 0.5f * (kf - 1.0f);
        for (int a = 0; a < 3; ++a) {
            c.pixel00_rel[a] += h * (c.pixel_du[a] + c.pixel_dv[a]);
            c.pixel_du[a] *= kf;
            c.pixel_dv[a] *= kf;
        }
        cams[i] = c;
    }
    std::vector<void*> means(njobs, nullptr);
    struct Free {
        std::vector<void*>& v;
        ~Free() {
            for (void* p : v)
                if (p) (void)hipFree(p);
        }
    } fr{means};
    for (uint32_t i = 0; i < njobs; ++i)
        if (hipMalloc(&means[i], sizeof(float4) * params[i].width * params[i].height) != hipSuccess)
            return fail(VP_ERR_OOM, "autotune: cannot allocate images");
    auto make_runs = [&](std::vector<JobRun>& runs) {
        runs.assign(njobs, JobRun{});
        for (uint32_t i = 0; i < njobs; ++i) {
            runs[i].desc = vp_make_job(jobs[i].scene, jobs[i].env, &cams[i], &params[i]);
            runs[i].mean = static_cast<float4*>(means[i]);
        }
    };
    // active-set capacity from the measured maximum overlap
    uint32_t cap = 128;
    {
        std::vector<JobRun> runs;
        make_runs(runs);
        vp_status st = vp_reset_counters(ctx, s);
        if (st == VP_OK) st = vp_run_jobs(ctx, runs, nullptr, 1024, 0, 1, s, nullptr, nullptr);
        vp_render_stats rs{};
        if (st == VP_OK) st = vp_read_stats(ctx, s, &rs);
        if (st != VP_OK) return st;
        const uint32_t v = rs.max_point_overlap;
        cap = std::max<uint32_t>(16u, ((v * 9 + 7) / 8 + 15) / 16 * 16);
        // the reduced mix sees fewer vertices than the full one: never go
        // below the capacities the jobs ask for
        for (uint32_t i = 0; i < njobs; ++i) cap = std::max(cap, jobs[i].params.active_cap);
        cap = std::min(cap, vp::kMaxActiveCap);  // beyond it sets overflow and are counted
    }
    // candidate shapes: megakernel workgroup size x regeneration threshold
    // at the occupancy limit, one at half occupancy, and the wavefront
    // pipeline at two pool sizes
    std::vector<vp_launch_config> cands;
    for (uint32_t b : {128u, 256u}) {
        int occ = 1;
        vp_render_occupancy(b, cap, 0, &occ);
        for (uint32_t rg : {1u, 8u, 16u})
            cands.push_back(vp_launch_config{b, static_cast<uint32_t>(occ), cap, 0, rg, VP_PIPELINE_MEGAKERNEL, 0});
        if (b == 256u && occ > 1)
            cands.push_back(
                vp_launch_config{b, static_cast<uint32_t>(occ) / 2u, cap, 0, 8u, VP_PIPELINE_MEGAKERNEL, 0});
    }
    for (uint32_t pool : {1u << 20, 1u << 21})
        if (static_cast<uint64_t>(pool) * cap * vp::kSlabBuffers < (uint64_t(1) << 32))
            cands.push_back(vp_launch_config{256u, 0u, cap, 0u, 48u, VP_PIPELINE_WAVEFRONT, pool});
    vp_launch_config best_cfg = cands.front();
    double best_rate = -1.0;
    for (const vp_launch_config& c : cands) {
        std::vector<JobRun> runs;
        make_runs(runs);
        uint64_t paths = 0;
        const auto t0 = clk::now();
        vp_status st = vp_run_jobs(ctx, runs, &c, cap, 0, 0, s, nullptr, &paths);
        if (st != VP_OK) return st;
        const double wall = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        const double rate = static_cast<double>(paths) / std::max(wall, 1e-3);
        if (rate > best_rate) {
            best_rate = rate;
            best_cfg = c;
        }
    }
    ctx->tuned = best_cfg;
    ctx->have_tuned = true;
    if (best) *best = best_cfg;
    return VP_OK;
}

}  // extern "C"
