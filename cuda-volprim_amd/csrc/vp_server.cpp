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

struct Request {
    uint64_t ticket = 0;
    vp_job job{};
    JobRun run;
    float4* mean = nullptr;
    float4* slice = nullptr;
    size_t bytes = 0;
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
};

namespace {

void free_request_device(Request* r) {
    if (r->mean) (void)hipFree(r->mean);
    if (r->slice) (void)hipFree(r->slice);
    r->mean = r->slice = nullptr;
}

void server_loop(vp_server* S) {
    vp_context* ctx = S->ctx;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    std::vector<Request*> batch;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // admission: FIFO, by device bytes in flight and free slots
            while (!S->pending.empty() && S->inflight.size() < S->cfg.max_inflight) {
                Request* r = S->pending.front();
                if (!S->inflight.empty() && S->bytes_inflight + r->bytes > S->cfg.mem_budget_bytes) break;
                const size_t npix = static_cast<size_t>(r->job.params.width) * r->job.params.height;
                if (hipMalloc(&r->mean, npix * sizeof(float4)) != hipSuccess ||
                    hipMalloc(&r->slice, static_cast<size_t>(r->run.slice_spp) * npix * sizeof(float4)) != hipSuccess) {
                    free_request_device(r);
                    if (!S->inflight.empty()) break;  // wait for memory held by running requests
                    r->status = VP_ERR_OOM;
                    r->done = true;
                    S->pending.pop_front();
                    S->cv_done.notify_all();
                    continue;
                }
                r->run.mean = r->mean;
                S->pending.pop_front();
                S->inflight.push_back(r);
                S->bytes_inflight += r->bytes;
            }
            if (S->inflight.empty()) {
                if (S->stop) return;
                S->cv.wait(lk);
                continue;
            }
            batch = S->inflight;
        }
        // stage one slice of every admitted request
        uint64_t total = 0;
        uint32_t k = 0;
        for (Request* r : batch) {
            JobRun& run = r->run;
            vp::JobDesc J = run.desc;
            J.s_begin = run.done;
            J.s_count = std::min(run.slice_spp, run.desc.spp - run.done);
            J.path_begin = total;
            J.path_count = static_cast<uint64_t>(J.npad) * J.s_count;
            J.mean = r->mean;
            J.samples = r->slice;
            ctx->jobs_pinned[k] = J;
            ctx->begin_pinned[k] = J.path_begin;
            total += J.path_count;
            ++k;
        }
        const auto t0 = clk::now();
        vp_status st = vp_run_step(ctx, S->L, k, total, S->cfg.max_inflight, S->max_npix, true, s);
        if (st == VP_OK && hipStreamSynchronize(s) != hipSuccess) st = VP_ERR_HIP;
        const double ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        std::lock_guard<std::mutex> lk(S->mu);
        S->busy_ms += ms;
        ++S->steps;
        std::vector<Request*> still;
        for (uint32_t i = 0; i < batch.size(); ++i) {
            Request* r = batch[i];
            r->run.done += ctx->jobs_pinned[i].s_count;
            if (st != VP_OK || r->run.done >= r->run.desc.spp) {
                const size_t npix = static_cast<size_t>(r->job.params.width) * r->job.params.height;
                if (st == VP_OK &&
                    hipMemcpy(r->job.out_rgba, r->mean, npix * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess)
                    st = VP_ERR_HIP;
                r->status = st;
                r->latency_ms = std::chrono::duration<double, std::milli>(clk::now() - r->t_submit).count();
                r->done = true;
                free_request_device(r);
                S->bytes_inflight -= r->bytes;
                S->latencies.push_back(r->latency_ms);
                ++S->completed;
                S->samples += static_cast<uint64_t>(npix) * r->run.desc.spp;
            } else {
                still.push_back(r);
            }
        }
        S->inflight.swap(still);
        if (st != VP_OK) S->error = "step failed";
        S->cv_done.notify_all();
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
    auto* S = new vp_server();
    S->ctx = ctx;
    if (cfg) S->cfg = *cfg;
    if (S->cfg.mem_budget_bytes == 0) S->cfg.mem_budget_bytes = size_t(8) << 30;
    if (S->cfg.slice_spp == 0) S->cfg.slice_spp = 4;
    if (S->cfg.max_inflight == 0) S->cfg.max_inflight = 64;
    S->max_npix = 1024u * 1024u;
    if (hipSetDevice(ctx->device) != hipSuccess) {
        delete S;
        return fail(VP_ERR_HIP, "hipSetDevice failed");
    }
    const uint32_t cap = ctx->have_tuned && ctx->tuned.active_cap ? ctx->tuned.active_cap : 128u;
    S->L = vp_resolve_launch(ctx, nullptr, cap, 0, 0);
    vp_status st = S->L.pipeline == VP_PIPELINE_MEGAKERNEL
                       ? vp_ensure_slab(ctx, 2 * static_cast<size_t>(cap) * S->L.grid * S->L.block)
                       : VP_OK;
    if (st == VP_OK) st = vp_ensure_jobs(ctx, S->cfg.max_inflight);
    if (st == VP_OK) st = vp_reset_counters(ctx, ctx->stream);
    if (st != VP_OK) {
        delete S;
        return st;
    }
    S->th = std::thread(server_loop, S);
    *out = S;
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
    auto* r = new Request();
    r->job = *job;
    r->run.desc = vp_make_job(job->scene, job->env, &job->camera, &job->params);
    r->run.slice_spp = std::min(S->cfg.slice_spp, job->params.spp);
    const size_t npix = static_cast<size_t>(job->params.width) * job->params.height;
    r->bytes = npix * sizeof(float4) * (1 + r->run.slice_spp);
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
    if (S->th.joinable()) S->th.join();
    for (auto& kv : S->by_ticket) {
        free_request_device(kv.second);
        delete kv.second;
    }
    vp_drop_graph(S->ctx);
    delete S;
    return VP_OK;
}

vp_status vp_autotune(vp_context* ctx, vp_job* jobs, uint32_t njobs, vp_launch_config* best) {
    if (!ctx || !jobs || njobs == 0) return fail(VP_ERR_INVALID, "no jobs");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(VP_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    // reduced copy of the mix: every view downscaled by an integer factor k
    // to at most kTuneRes pixels a side (same field of view, k-times wider
    // pixels) at one sample per pixel
    constexpr uint32_t kTuneRes = 64;
    std::vector<vp_render_params> params(njobs);
    std::vector<vp_camera> cams(njobs);
    for (uint32_t i = 0; i < njobs; ++i) {
        vp_status st = vp_check_params(&jobs[i].params);
        if (st != VP_OK) return st;
        params[i] = jobs[i].params;
        params[i].spp = 1;
        params[i].measure_caps = 0;
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
    }
    vp_launch_config best_cfg{};
    double best_rate = -1.0;
    const uint32_t blocks[] = {64, 128, 256};
    const uint32_t regens[] = {1, 8, 32};
    for (uint32_t b : blocks) {
        int occ = 1;
        vp_render_occupancy(b, cap, 0, &occ);
        const uint32_t bpcs[] = {static_cast<uint32_t>(occ), std::max<uint32_t>(1u, static_cast<uint32_t>(occ) / 2u)};
        for (uint32_t bi = 0; bi < 2; ++bi) {
            if (bi == 1 && bpcs[1] == bpcs[0]) continue;
            for (uint32_t rg : regens) {
                vp_launch_config c{b, bpcs[bi], cap, 0, rg, VP_PIPELINE_MEGAKERNEL, 0};
                std::vector<JobRun> runs;
                make_runs(runs);
                double ms = 0.0;
                uint64_t paths = 0;
                const auto t0 = clk::now();
                vp_status st = vp_run_jobs(ctx, runs, &c, cap, 0, 0, s, &ms, &paths);
                if (st != VP_OK) return st;
                const double wall = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
                const double rate = static_cast<double>(paths) / std::max(wall, 1e-3);
                if (rate > best_rate) {
                    best_rate = rate;
                    best_cfg = c;
                }
            }
        }
    }
    // the wavefront pipeline at three pool sizes
    for (uint32_t pool : {1u << 19, 1u << 20, 1u << 21}) {
        if (static_cast<uint64_t>(pool) * cap * 2 >= (uint64_t(1) << 32)) continue;
        vp_launch_config c{256u, 0u, cap, 0u, 1u, VP_PIPELINE_WAVEFRONT, pool};
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
