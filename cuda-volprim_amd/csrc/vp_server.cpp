// vp_server.cpp — the serving form and the launch autotuner.
//
// Serving: a request loop on its own thread keeps scenes and environment
// tables resident (requests reference vp_scene / vp_env handles created once)
// and batches continuously on the persistent megakernel.  An admitted request
// becomes one entry of the server's job table (all its samples), appended
// while earlier launches are still running; a new launch over the longer
// table then starts on the second of two render streams.  Each launch
// claims, with its own counter, the range of paths its table adds to the
// previous launch's, so the new launch takes the new request's paths while
// the older one finishes its own: as the older launch's wavefronts run out of
// paths and exit, the new launch's workgroups take their place, and no
// request waits for a step barrier or another request's tail.  Each launch has its own
// active-set slab, so two can run at once.  Finished paths add into their
// request's per-pixel accumulator (exact fixed-point sums, so the image does
// not depend on which launch or lane ran which path) and count against its
// table entry; once the count reaches the request's paths, its accumulator
// is copied to the host and the mean resolved there while rendering
// continues.  The side stream installs requests and copies results; its
// copies run as blit kernels on this runtime, so the render streams are
// masked to every CU but one and their launches sized for that, which keeps
// one CU free for them while a persistent launch holds the rest.  Request memory is the
// accumulator (64 bytes per pixel), whatever the sample count; admission is
// by those bytes in flight and by a limit on requests in flight.
//
// Autotuning: vp_autotune times a reduced copy of a job mix (about four
// GPU-fills of paths) for a grid of launch shapes (megakernel: workgroup
// size, regeneration threshold; wavefront pipeline: path-pool size) and
// sizes the active-set capacity from the measured maximum
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
constexpr uint32_t kServerTable = 16384;
// Device bytes per pixel of a request: its accumulator.
constexpr size_t kRequestBytesPerPixel = vp::kAccumWords * sizeof(unsigned long long);

enum class ReqState { Queued, Running, Resolving, Done };

struct Request {
    uint64_t ticket = 0;
    vp_job job{};
    vp::JobDesc desc{};
    unsigned long long* accum = nullptr;
    size_t bytes = 0;
    uint64_t paths = 0;     // camera paths of the request (W * H * spp)
    uint32_t entry = 0;     // job-table entry in the current epoch
    ReqState state = ReqState::Queued;
    hipEvent_t ev = nullptr;  // recorded after the copy of the request's accumulator
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
    uint64_t completed = 0, samples = 0, launches = 0;
    double busy_ms = 0.0;
    std::string error;
    // device resources, owned by the server (the context's tables stay free
    // for other render calls)
    vp::JobDesc* jobs_dev = nullptr;
    uint64_t* begin_dev = nullptr;
    unsigned long long* done_dev = nullptr;  // [kServerTable] finished paths per entry
    unsigned long long* counter = nullptr;   // [2] path counter of the launch on each render stream
    uint64_t* desc_dev = nullptr;            // [2][4] launch descriptors, one per render stream
    uint32_t* slab[2] = {nullptr, nullptr};
    vp::JobDesc* jobs_pinned = nullptr;
    uint64_t* begin_pinned = nullptr;
    unsigned long long* done_pinned = nullptr;
    uint64_t* desc_pinned = nullptr;  // [2][4]
    hipStream_t ks[2] = {nullptr, nullptr};  // render streams
    hipStream_t side = nullptr;              // installs, progress copies, result copies (copies only)
    unsigned long long* stage = nullptr;     // pinned: accumulator of the request being resolved
    unsigned long long* zeros = nullptr;     // pinned zeros: accumulator and progress initialisation
    Request* resolving = nullptr;            // request whose accumulator is being copied
    hipEvent_t kev[2] = {nullptr, nullptr};  // end of the last launch on each render stream
    hipEvent_t install_ev = nullptr, poll_ev = nullptr;
    // epoch state, server thread only
    bool kbusy[2] = {false, false};
    uint32_t njobs = 0;
    uint64_t total = 0;  // paths installed in the epoch
    uint64_t snap = 0;   // paths visible to the newest launch
    bool poll_pending = false;
    uint32_t poll_n = 0;
};

namespace {

void release_request(vp_server* S, Request* r) {
    if (r->accum) (void)hipFreeAsync(r->accum, S->side);
    r->accum = nullptr;
    if (r->ev) (void)hipEventDestroy(r->ev);
    r->ev = nullptr;
}

// Finish request r (under S->mu): record its latency and wake pollers.
void finish_request(vp_server* S, Request* r, vp_status st) {
    r->status = st;
    r->state = ReqState::Done;
    r->latency_ms = std::chrono::duration<double, std::milli>(clk::now() - r->t_submit).count();
    r->done = true;
    S->bytes_inflight -= r->bytes;
    if (st == VP_OK) {
        S->latencies.push_back(r->latency_ms);
        ++S->completed;
        S->samples += r->paths;
    }
    S->inflight.erase(std::find(S->inflight.begin(), S->inflight.end(), r));
    S->cv_done.notify_all();
}

bool streams_idle(vp_server* S) {
    for (int i = 0; i < 2; ++i)
        if (S->kbusy[i] && hipEventQuery(S->kev[i]) == hipSuccess) S->kbusy[i] = false;
    return !S->kbusy[0] && !S->kbusy[1];
}

// Install request r as the next table entry: accumulator (stream-ordered
// on the side stream, zeroed by a copy), table entry and progress counter.
// On failure nothing of the epoch changes and only r fails.
vp_status install_request(vp_server* S, Request* r) {
    const size_t npix = static_cast<size_t>(r->job.params.width) * r->job.params.height;
    const size_t acc_bytes = npix * vp::kAccumWords * sizeof(unsigned long long);
    if (hipMallocAsync(reinterpret_cast<void**>(&r->accum), acc_bytes, S->side) != hipSuccess) {
        (void)hipGetLastError();
        r->accum = nullptr;
        return fail(VP_ERR_OOM, "server: cannot allocate a request's accumulator");
    }
    const uint32_t e = S->njobs;
    vp::JobDesc J = r->desc;
    J.s_begin = 0;
    J.s_count = J.spp;
    J.path_begin = S->total;
    J.path_count = static_cast<uint64_t>(J.npad) * J.spp;
    J.accum = r->accum;
    J.mean = nullptr;  // resolved on the host
    J.samples = nullptr;
    J.done = S->done_dev + e;
    S->jobs_pinned[e] = J;
    S->begin_pinned[e] = J.path_begin;
    if (hipMemcpyAsync(r->accum, S->zeros, acc_bytes, hipMemcpyHostToDevice, S->side) != hipSuccess ||
        hipMemcpyAsync(S->done_dev + e, S->zeros, sizeof(unsigned long long), hipMemcpyHostToDevice, S->side) !=
            hipSuccess ||
        hipMemcpyAsync(S->jobs_dev + e, S->jobs_pinned + e, sizeof(vp::JobDesc), hipMemcpyHostToDevice, S->side) !=
            hipSuccess ||
        hipMemcpyAsync(S->begin_dev + e, S->begin_pinned + e, sizeof(uint64_t), hipMemcpyHostToDevice, S->side) !=
            hipSuccess)
        return fail(VP_ERR_HIP, "server: cannot install a request");
    r->desc = J;
    r->entry = e;
    r->state = ReqState::Running;
    ++S->njobs;
    S->total += J.path_count;
    return VP_OK;
}

// Fail every request in flight after a device error; the next epoch starts
// from an empty table once the render streams have drained.
void fail_all(vp_server* S, const std::string& what) {
    for (auto& k : S->ks) (void)hipStreamSynchronize(k);
    (void)hipStreamSynchronize(S->side);
    std::lock_guard<std::mutex> lk(S->mu);
    S->error = what;
    while (!S->inflight.empty()) {
        Request* r = S->inflight.back();
        release_request(S, r);
        finish_request(S, r, VP_ERR_HIP);
    }
    S->kbusy[0] = S->kbusy[1] = false;
    S->poll_pending = false;
    S->resolving = nullptr;
}

void server_loop(vp_server* S) {
    vp_context* ctx = S->ctx;
    (void)hipSetDevice(ctx->device);
    bool fresh = true;  // no launch in this epoch yet
    clk::time_point busy_since{};
    std::vector<Request*> admit;
    while (true) {
        admit.clear();
        bool progress = false;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->inflight.empty() && (S->pending.empty() || S->njobs >= kServerTable)) {
                // nothing on the GPU: start a new epoch, with the table refilled
                // from entry 0, once the last launches have exited
                lk.unlock();
                for (auto& k : S->ks) (void)hipStreamSynchronize(k);
                (void)hipStreamSynchronize(S->side);
                lk.lock();
                S->kbusy[0] = S->kbusy[1] = false;
                S->poll_pending = false;
                if (!fresh) S->busy_ms += std::chrono::duration<double, std::milli>(clk::now() - busy_since).count();
                S->njobs = 0;
                S->total = S->snap = 0;
                fresh = true;
                if (S->pending.empty()) {
                    if (S->stop) return;
                    S->cv.wait(lk);
                    continue;
                }
            }
            // admission: FIFO, by device bytes in flight and requests in flight
            while (!S->pending.empty() && S->inflight.size() < S->cfg.max_inflight &&
                   S->njobs + admit.size() < kServerTable) {
                Request* r = S->pending.front();
                if (!S->inflight.empty() && S->bytes_inflight + r->bytes > S->cfg.mem_budget_bytes) break;
                S->pending.pop_front();
                S->inflight.push_back(r);
                S->bytes_inflight += r->bytes;
                admit.push_back(r);
            }
        }
        if (fresh && !admit.empty()) {
            busy_since = clk::now();
            fresh = false;
        }
        // install the admitted requests (a request that cannot be installed
        // fails alone; the table and path total are untouched by it)
        bool installed = false;
        for (Request* r : admit) {
            const vp_status st = install_request(S, r);
            if (st != VP_OK) {
                release_request(S, r);
                std::lock_guard<std::mutex> lk(S->mu);
                finish_request(S, r, st);
                continue;
            }
            installed = true;
        }
        if (installed) {
            progress = true;
            if (hipEventRecord(S->install_ev, S->side) != hipSuccess) {
                fail_all(S, "server: event record failed");
                continue;
            }
        }
        // a launch over the paths the grown table adds, [snap, total), on a
        // render stream that is free, with that stream's counter reset
        streams_idle(S);
        if (S->total > S->snap) {
            const int i = !S->kbusy[0] ? 0 : (!S->kbusy[1] ? 1 : -1);
            if (i >= 0) {
                uint64_t* d = S->desc_pinned + 4 * i;
                d[0] = S->njobs;
                d[1] = S->total;
                d[2] = S->snap;
                d[3] = 0;
                if (hipStreamWaitEvent(S->ks[i], S->install_ev, 0) != hipSuccess ||
                    hipMemcpyAsync(S->desc_dev + 4 * i, d, 4 * sizeof(uint64_t), hipMemcpyHostToDevice, S->ks[i]) !=
                        hipSuccess ||
                    hipMemcpyAsync(S->counter + vp::kClaimParts * i, S->zeros,
                                   vp::kClaimParts * sizeof(unsigned long long), hipMemcpyHostToDevice,
                                   S->ks[i]) != hipSuccess ||
                    vp_launch_render(S->L, S->jobs_dev, S->begin_dev, S->desc_dev + 4 * i, S->counter + vp::kClaimParts * i,
                                     S->slab[i],
                                     ctx->counters, S->ks[i]) != hipSuccess ||
                    hipEventRecord(S->kev[i], S->ks[i]) != hipSuccess) {
                    fail_all(S, "server: render launch failed");
                    continue;
                }
                S->kbusy[i] = true;
                S->snap = S->total;
                ++S->launches;
                progress = true;
            }
        }
        // progress of the entries, one copy in flight at a time
        if (!S->poll_pending && S->njobs > 0) {
            S->poll_n = S->njobs;
            if (hipMemcpyAsync(S->done_pinned, S->done_dev, sizeof(unsigned long long) * S->poll_n,
                               hipMemcpyDeviceToHost, S->side) != hipSuccess ||
                hipEventRecord(S->poll_ev, S->side) != hipSuccess) {
                fail_all(S, "server: progress copy failed");
                continue;
            }
            S->poll_pending = true;
        } else if (S->poll_pending && hipEventQuery(S->poll_ev) == hipSuccess) {
            S->poll_pending = false;
            std::vector<Request*> snapshot;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                snapshot = S->inflight;
            }
            bool err = false;
            for (Request* r : snapshot) {
                if (r->state == ReqState::Running && S->resolving == nullptr && r->entry < S->poll_n &&
                    S->done_pinned[r->entry] >= r->paths) {
                    // every path of the request has finished: copy its
                    // accumulator out, behind the rendering (one at a time
                    // through the pinned stage)
                    const size_t npix = static_cast<size_t>(r->desc.width) * r->desc.height;
                    if (hipMemcpyAsync(S->stage, r->accum, npix * vp::kAccumWords * sizeof(unsigned long long),
                                       hipMemcpyDeviceToHost, S->side) != hipSuccess ||
                        hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess ||
                        hipEventRecord(r->ev, S->side) != hipSuccess) {
                        err = true;
                        break;
                    }
                    r->state = ReqState::Resolving;
                    S->resolving = r;
                    progress = true;
                }
            }
            Request* r = S->resolving;
            if (!err && r != nullptr && hipEventQuery(r->ev) == hipSuccess) {
                // the mean of every pixel, as k_resolve_* forms it on the device
                const size_t npix = static_cast<size_t>(r->desc.width) * r->desc.height;
                float* out = r->job.out_rgba;
                for (size_t i = 0; i < npix; ++i) {
                    const unsigned long long* a = S->stage + i * vp::kAccumWords;
                    out[4 * i + 0] = vp::accum_mean(a[0], a[4], a[3]);
                    out[4 * i + 1] = vp::accum_mean(a[1], a[5], a[3]);
                    out[4 * i + 2] = vp::accum_mean(a[2], a[6], a[3]);
                    out[4 * i + 3] = 1.0f;
                }
                S->resolving = nullptr;
                release_request(S, r);
                std::lock_guard<std::mutex> lk(S->mu);
                finish_request(S, r, VP_OK);
                progress = true;
            }
            if (err) {
                fail_all(S, "server: result copy failed");
                continue;
            }
        }
        if (!progress) {
            // nothing to do until a launch or copy advances or a request
            // arrives: wait briefly for a submission
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait_for(lk, std::chrono::microseconds(200));
        }
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

void free_server(vp_server* S) {
    if (S->jobs_dev) (void)hipFree(S->jobs_dev);
    if (S->begin_dev) (void)hipFree(S->begin_dev);
    if (S->done_dev) (void)hipFree(S->done_dev);
    if (S->counter) (void)hipFree(S->counter);
    if (S->desc_dev) (void)hipFree(S->desc_dev);
    for (auto& p : S->slab)
        if (p) (void)hipFree(p);
    if (S->jobs_pinned) (void)hipHostFree(S->jobs_pinned);
    if (S->begin_pinned) (void)hipHostFree(S->begin_pinned);
    if (S->done_pinned) (void)hipHostFree(S->done_pinned);
    if (S->desc_pinned) (void)hipHostFree(S->desc_pinned);
    if (S->stage) (void)hipHostFree(S->stage);
    if (S->zeros) (void)hipHostFree(S->zeros);
    for (auto& k : S->ks)
        if (k) (void)hipStreamDestroy(k);
    if (S->side) (void)hipStreamDestroy(S->side);
    for (auto& e : S->kev)
        if (e) (void)hipEventDestroy(e);
    if (S->install_ev) (void)hipEventDestroy(S->install_ev);
    if (S->poll_ev) (void)hipEventDestroy(S->poll_ev);
    for (auto& kv : S->by_ticket) delete kv.second;
    delete S;
}

}  // namespace

extern "C" {

vp_status vp_server_create(vp_context* ctx, const vp_server_config* cfg, vp_server** out) {
    if (!ctx || !out) return fail(VP_ERR_INVALID, "null argument");
    std::unique_ptr<vp_server, void (*)(vp_server*)> S(new vp_server(), free_server);
    S->ctx = ctx;
    if (cfg) S->cfg = *cfg;
    if (S->cfg.mem_budget_bytes == 0) S->cfg.mem_budget_bytes = size_t(8) << 30;
    if (S->cfg.max_inflight == 0) S->cfg.max_inflight = 64;
    S->max_npix = 1024u * 1024u;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(VP_ERR_HIP, "hipSetDevice failed");
    // the tuned (or default) megakernel shape; a tuned wavefront shape keeps
    // its capacity and takes the megakernel's default shape
    vp_launch_config c{};
    if (ctx->have_tuned) c = ctx->tuned;
    if (c.pipeline != VP_PIPELINE_MEGAKERNEL) c = vp_launch_config{0, 0, c.active_cap, 0, 0, VP_PIPELINE_MEGAKERNEL, 0};
    uint32_t cap = c.active_cap ? c.active_cap : 128u;
    if (S->cfg.active_cap > cap) cap = S->cfg.active_cap;
    c.active_cap = cap;
    S->L = vp_resolve_launch(ctx, &c, cap, 0, 0);
    // With cfg.reserve_cus the render streams leave one CU of each of the 8
    // XCDs free, and their launches are sized for the rest: the side stream's
    // copies are blit kernels on this runtime, and a persistent launch holding
    // every CU keeps them (progress polls, result copies) waiting until
    // workgroups exit, so a finished request can wait for the rest of its
    // launch.  Workgroups are dealt out to the XCDs in turn, so a free CU in
    // one XCD alone would not be found by a copy dealt to another.  The free
    // CUs are 33 k (k = 0..7), one in each block of 32 CUs and one of each
    // residue mod 8, whichever of the two numberings maps CUs to XCDs.  The
    // masked launches measured 11 % below unmasked ones on the request-stream
    // workload (3.78 against 4.21 Msamples/s saturated), far more than the
    // 3 % of CUs they give up, so the default renders on every CU.
    const uint32_t cus = static_cast<uint32_t>(ctx->prop.multiProcessorCount);
    std::vector<uint32_t> cu_mask((cus + 31) / 32, 0u);
    uint32_t free_cus = 0;
    for (uint32_t cu = 0; cu < cus; ++cu) {
        const bool spare = S->cfg.reserve_cus != 0 && cus >= 64 && cus % 8 == 0 && cu % (cus / 8 + 1) == 0 &&
                           cu / (cus / 8 + 1) < 8;
        if (spare) ++free_cus;
        else cu_mask[cu / 32] |= 1u << (cu % 32);
    }
    if (free_cus > 0 && S->L.grid >= cus) S->L.grid = S->L.grid / cus * (cus - free_cus);
    const size_t slab_words = vp::kMegaSlabBuffers * static_cast<size_t>(vp::mega_slab_stride(cap)) * S->L.grid * S->L.block;
    bool ok = hipMalloc(&S->jobs_dev, sizeof(vp::JobDesc) * kServerTable) == hipSuccess &&
              hipMalloc(&S->begin_dev, sizeof(uint64_t) * kServerTable) == hipSuccess &&
              hipMalloc(&S->done_dev, sizeof(unsigned long long) * kServerTable) == hipSuccess &&
              hipMalloc(&S->counter, 2 * vp::kClaimParts * sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&S->desc_dev, 8 * sizeof(uint64_t)) == hipSuccess &&
              hipMalloc(&S->slab[0], slab_words * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&S->slab[1], slab_words * sizeof(uint32_t)) == hipSuccess;
    if (!ok) return fail(VP_ERR_OOM, "server: cannot allocate its tables and active-set slabs");
    ok = hipHostMalloc(&S->jobs_pinned, sizeof(vp::JobDesc) * kServerTable, hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc(&S->begin_pinned, sizeof(uint64_t) * kServerTable, hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc(&S->done_pinned, sizeof(unsigned long long) * kServerTable, hipHostMallocDefault) ==
             hipSuccess &&
         hipHostMalloc(&S->desc_pinned, 8 * sizeof(uint64_t), hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc(&S->stage, S->max_npix * kRequestBytesPerPixel, hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc(&S->zeros, S->max_npix * kRequestBytesPerPixel, hipHostMallocDefault) == hipSuccess;
    if (ok) std::memset(S->zeros, 0, S->max_npix * kRequestBytesPerPixel);
    {
        for (auto& k : S->ks) {
            if (!ok) break;
            if (free_cus == 0 ||
                hipExtStreamCreateWithCUMask(&k, static_cast<uint32_t>(cu_mask.size()), cu_mask.data()) !=
                    hipSuccess) {
                (void)hipGetLastError();
                k = nullptr;
                ok = hipStreamCreateWithFlags(&k, hipStreamNonBlocking) == hipSuccess;
            }
        }
    }
    ok = ok && hipStreamCreateWithFlags(&S->side, hipStreamNonBlocking) == hipSuccess;
    for (auto& e : S->kev) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&S->install_ev, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&S->poll_ev, hipEventDisableTiming) == hipSuccess;
    if (!ok) return fail(VP_ERR_HIP, "server: cannot create its streams, events and pinned staging");
    S->th = std::thread(server_loop, S.get());
    *out = S.release();
    return VP_OK;
}

vp_status vp_server_submit(vp_server* S, const vp_job* job, uint64_t* ticket) {
    if (!S || !job || !ticket) return fail(VP_ERR_INVALID, "null argument");
    if (!job->scene || !job->env || !job->out_rgba) return fail(VP_ERR_INVALID, "job has a null field");
    vp_status st = vp_check_params(&job->params);
    if (st != VP_OK) return st;
    const uint64_t npix = static_cast<uint64_t>(job->params.width) * job->params.height;
    if (npix > S->max_npix) return fail(VP_ERR_INVALID, "request larger than the server's 1024x1024 pixel limit");
    if (job->params.measure_caps) return fail(VP_ERR_INVALID, "measure_caps is not available through the server");
    if (job->params.adaptive_threshold > 0.0f)
        return fail(VP_ERR_INVALID, "adaptive sampling is not available through the server (a request runs as one "
                                    "slice; adaptive sampling needs its 16-sample batches)");
    if (job->params.active_cap > S->L.active_cap)
        return fail(VP_ERR_INVALID, "request asks for an active-set capacity above the server's (" +
                                        std::to_string(S->L.active_cap) + "): autotune with it first");
    const size_t bytes = static_cast<size_t>(npix) * kRequestBytesPerPixel;
    if (bytes > S->cfg.mem_budget_bytes)
        return fail(VP_ERR_INVALID, "request needs " + std::to_string(bytes) +
                                        " device bytes, more than the server's memory budget");
    auto* r = new Request();
    r->job = *job;
    r->desc = vp_make_job(job->scene, job->env, &job->camera, &job->params);
    r->bytes = bytes;
    r->paths = npix * job->params.spp;
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
    // an error fails the requests it hit and is reported once
    std::string err;
    std::swap(err, S->error);
    return err.empty() ? VP_OK : fail(VP_ERR_HIP, err);
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
    out->launches = S->launches;
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
    for (auto& k : S->ks) (void)hipStreamSynchronize(k);
    (void)hipStreamSynchronize(S->side);
    free_server(S);
    return VP_OK;
}

vp_status vp_autotune(vp_context* ctx, vp_job* jobs, uint32_t njobs, vp_launch_config* best) {
    if (!ctx || !jobs || njobs == 0) return fail(VP_ERR_INVALID, "no jobs");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(VP_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = ctx->stream;
    // Reduced copy of the mix, sized to keep the GPU full several times
    // over (about four waves of the megakernel's lanes): every job's samples
    // are scaled by the same factor f, half through fewer samples per pixel
    // and half through a view downscaled by an integer factor k (same field
    // of view, k-times wider pixels), at depth at most 16.  A mix much smaller
    // than the GPU would time launch tails only, where the wavefront pool
    // (all paths in flight at once) wins, though it loses on full-size work.
    int occ_max = 1;
    vp_render_occupancy(256u, 128u, 0, &occ_max);
    const double target = 4.0 * static_cast<double>(ctx->prop.multiProcessorCount) * occ_max * 256.0;
    double total = 0.0;
    for (uint32_t i = 0; i < njobs; ++i)
        total += static_cast<double>(jobs[i].params.width) * jobs[i].params.height * jobs[i].params.spp;
    const double f = std::min(1.0, target / std::max(total, 1.0));
    std::vector<vp_render_params> params(njobs);
    std::vector<vp_camera> cams(njobs);
    for (uint32_t i = 0; i < njobs; ++i) {
        vp_status st = vp_check_params(&jobs[i].params);
        if (st != VP_OK) return st;
        params[i] = jobs[i].params;
        params[i].measure_caps = 0;
        params[i].adaptive_threshold = 0.0f;
        // long paths only lengthen each timing run's tail
        params[i].max_depth = std::min<uint32_t>(params[i].max_depth, 16u);
        const double want = static_cast<double>(params[i].spp) * f;  // samples per full-size pixel
        const uint32_t spp = static_cast<uint32_t>(std::max(1.0, std::floor(params[i].spp * std::sqrt(f) + 0.5)));
        const double area = want / spp;  // fraction of the pixels kept
        const uint32_t k = std::max<uint32_t>(1u, static_cast<uint32_t>(std::floor(1.0 / std::sqrt(area) + 0.5)));
        params[i].spp = spp;
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
    // at the occupancy limit, and the wavefront pipeline at two pool sizes
    std::vector<vp_launch_config> cands;
    for (uint32_t b : {256u, 128u}) {
        int occ = 1;
        vp_render_occupancy(b, cap, 0, &occ);
        for (uint32_t rg : {8u, 16u})
            cands.push_back(vp_launch_config{b, static_cast<uint32_t>(occ), cap, 0, rg, VP_PIPELINE_MEGAKERNEL, 0});
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
