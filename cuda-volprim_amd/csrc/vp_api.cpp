// vp_api.cpp — the C ABI (include/vp_api.h): context, scene, environment,
// camera, the plain and batched renderers, and test entry points.  Host code
// only; the kernels live in vp_bvh.hip and vp_render.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "vp_internal.h"

namespace {
thread_local std::string g_error;

struct Dev {
    void* p = nullptr;
    Dev() = default;
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};

vp_status fail(vp_status code, const std::string& msg) {
    g_error = msg;
    return code;
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

vp_status check_params_impl(const vp_render_params* p) {
    if (!p) return fail(VP_ERR_INVALID, "params is null");
    if (p->width == 0 || p->height == 0) return fail(VP_ERR_INVALID, "width and height must be positive");
    if (static_cast<uint64_t>(p->width) * p->height > (1ull << 28))
        return fail(VP_ERR_INVALID, "image larger than 2^28 pixels");
    if (p->spp == 0) return fail(VP_ERR_INVALID, "spp must be positive");
    if (p->max_depth == 0) return fail(VP_ERR_INVALID, "max_depth must be positive");
    if (!(p->phase_g >= -0.999f && p->phase_g <= 0.999f)) return fail(VP_ERR_INVALID, "phase_g must be in [-0.999, 0.999]");
    if (!(p->rr_max_survival >= 0.0f && p->rr_max_survival <= 1.0f))
        return fail(VP_ERR_INVALID, "rr_max_survival must be in [0, 1]");
    if (!(p->firefly_clamp >= 0.0f)) return fail(VP_ERR_INVALID, "firefly_clamp must be >= 0");
    if (!(p->filter_stddev >= 0.0f)) return fail(VP_ERR_INVALID, "filter_stddev must be >= 0");
    if (p->use_ris && p->ris_candidates == 0) return fail(VP_ERR_INVALID, "ris_candidates must be positive");
    if (p->active_cap > vp::kMaxActiveCap) return fail(VP_ERR_INVALID, "active_cap must be at most 65535");
    if (!(p->shadow_rr >= 0.0f && p->shadow_rr < 1e30f)) return fail(VP_ERR_INVALID, "shadow_rr must be >= 0");
    if (!(p->adaptive_threshold >= 0.0f && p->adaptive_threshold <= 3.4e38f))
        return fail(VP_ERR_INVALID, "adaptive_threshold must be finite and >= 0");
    if (p->adaptive_threshold > 0.0f && p->adaptive_min_samples == 0)
        return fail(VP_ERR_INVALID, "adaptive_min_samples must be positive");
    if (p->adaptive_threshold > 0.0f && p->spp >= (1u << 31))
        return fail(VP_ERR_INVALID, "adaptive sampling counts samples in 31 bits: spp must be below 2^31");
    return VP_OK;
}

vp::JobDesc make_job_impl(const vp_scene* sc, const vp_env* env, const vp_camera* cam, const vp_render_params* p) {
    vp::JobDesc J{};
    J.scene = sc->view();
    J.env = env->view();
    for (int k = 0; k < 3; ++k) {
        J.cam.eye[k] = cam->eye[k];
        J.cam.p00[k] = cam->pixel00_rel[k];
        J.cam.du[k] = cam->pixel_du[k];
        J.cam.dv[k] = cam->pixel_dv[k];
        J.cam.view_dir[k] = cam->view_dir[k];
    }
    J.cam.orthographic = cam->orthographic != 0.0f ? 1 : 0;
    J.hg = vp::make_phase(p->phase_g);
    J.width = p->width;
    J.height = p->height;
    J.spp = p->spp;
    J.seed = p->seed;
    J.max_depth = p->max_depth;
    J.rr_depth = p->rr_depth;
    J.rr_max_survival = p->rr_max_survival;
    J.firefly_clamp = p->firefly_clamp;
    J.filter_stddev = p->filter_stddev;
    J.use_ris = p->use_ris ? 1u : 0u;
    J.ris_candidates = p->ris_candidates ? p->ris_candidates : 6u;
    J.shadow_rr = p->shadow_rr;
    J.adapt_threshold = p->adaptive_threshold;
    J.adapt_min = p->adaptive_min_samples;
    J.tiles_x = (p->width + 7) / 8;
    const uint32_t tiles_y = (p->height + 7) / 8;
    J.npad = J.tiles_x * tiles_y * 64u;
    return J;
}

// Vose alias table over p_i = joint_i / sum(joint), double precision,
// stacks popped from the back.
void build_alias(const std::vector<float>& joint, std::vector<vp::AliasEntry>& out) {
    const size_t n = joint.size();
    out.assign(n, vp::AliasEntry{1.0f, 0u});
    double sum = 0.0;
    for (float v : joint) sum += static_cast<double>(v);
    std::vector<double> q(n);
    std::vector<uint32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        q[i] = sum > 0.0 ? static_cast<double>(joint[i]) / sum * static_cast<double>(n) : 1.0;
        if (q[i] < 1.0) small.push_back(static_cast<uint32_t>(i));
        else large.push_back(static_cast<uint32_t>(i));
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t s = small.back();
        small.pop_back();
        const uint32_t l = large.back();
        large.pop_back();
        out[s].prob = static_cast<float>(q[s]);
        out[s].alias = l;
        q[l] = (q[l] + q[s]) - 1.0;
        if (q[l] < 1.0) small.push_back(l);
        else large.push_back(l);
    }
    while (!large.empty()) {
        const uint32_t l = large.back();
        large.pop_back();
        out[l] = vp::AliasEntry{1.0f, l};
    }
    while (!small.empty()) {
        const uint32_t s = small.back();
        small.pop_back();
        out[s] = vp::AliasEntry{1.0f, s};
    }
}

}  // namespace

void vp_set_error(const std::string& msg) { g_error = msg; }
vp_status vp_check_params(const vp_render_params* p) { return check_params_impl(p); }
vp::JobDesc vp_make_job(const vp_scene* sc, const vp_env* env, const vp_camera* cam, const vp_render_params* p) {
    return make_job_impl(sc, env, cam, p);
}

extern "C" {

const char* vp_last_error(void) { return g_error.c_str(); }

vp_status vp_context_create(int device, vp_context** out) {
    if (!out) return fail(VP_ERR_INVALID, "out is null");
    int count = 0;
    VP_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(VP_ERR_INVALID, "no such device");
    VP_HIP_TRY(hipSetDevice(device));
    auto* ctx = new vp_context();
    ctx->device = device;
    if (hipGetDeviceProperties(&ctx->prop, device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
        hipMalloc(&ctx->counters, sizeof(unsigned long long) * vp::kNumCounters) != hipSuccess ||
        hipMalloc(&ctx->path_counter, sizeof(unsigned long long) * vp::kClaimParts) != hipSuccess ||
        hipMalloc(&ctx->launch_desc, sizeof(uint64_t) * 4) != hipSuccess ||
        hipHostMalloc(&ctx->desc_pinned, sizeof(uint64_t) * 4, hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        vp_context_destroy(ctx);
        return fail(VP_ERR_HIP, "context setup failed");
    }
    const vp_status st = vp_noalloc_attach(ctx);
    if (st != VP_OK) {
        const std::string msg = vp_last_error();
        vp_context_destroy(ctx);
        return fail(st, msg);
    }
    *out = ctx;
    return VP_OK;
}

vp_status vp_context_destroy(vp_context* ctx) {
    if (!ctx) return VP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->slab) (void)hipFree(ctx->slab);
    if (ctx->counters) (void)hipFree(ctx->counters);
    if (ctx->path_counter) (void)hipFree(ctx->path_counter);
    vp_drop_graph(ctx);
    vp_free_wf(ctx);
    if (ctx->jobs_dev) (void)hipFree(ctx->jobs_dev);
    if (ctx->job_begin_dev) (void)hipFree(ctx->job_begin_dev);
    if (ctx->job_done_dev) (void)hipFree(ctx->job_done_dev);
    if (ctx->launch_desc) (void)hipFree(ctx->launch_desc);
    if (ctx->jobs_pinned) (void)hipHostFree(ctx->jobs_pinned);
    if (ctx->begin_pinned) (void)hipHostFree(ctx->begin_pinned);
    if (ctx->desc_pinned) (void)hipHostFree(ctx->desc_pinned);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VP_OK;
}

vp_status vp_device_info(vp_context* ctx, char* name, size_t name_len, uint32_t* num_cus, double* clock_mhz) {
    if (!ctx) return fail(VP_ERR_INVALID, "ctx is null");
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", ctx->prop.name, ctx->prop.gcnArchName);
    }
    if (num_cus) *num_cus = static_cast<uint32_t>(ctx->prop.multiProcessorCount);
    if (clock_mhz) *clock_mhz = ctx->prop.clockRate / 1000.0;
    return VP_OK;
}

vp_status vp_scene_create(vp_context* ctx, const vp_primitive* prims, uint32_t n, vp_scene** out) {
    return vp_scene_create_ex(ctx, prims, n, 0u, out);
}

vp_status vp_scene_create_ex(vp_context* ctx, const vp_primitive* prims, uint32_t n, uint32_t flags,
                             vp_scene** out) {
    if (!ctx || !prims || !out) return fail(VP_ERR_INVALID, "null argument");
    if (flags & ~(VP_SCENE_BVH_SAH | VP_SCENE_BVH_MEDIAN | VP_SCENE_WALK_F32 | VP_SCENE_WALK_F16))
        return fail(VP_ERR_INVALID, "unknown scene flags");
    if ((flags & VP_SCENE_BVH_SAH) && (flags & VP_SCENE_BVH_MEDIAN))
        return fail(VP_ERR_INVALID, "VP_SCENE_BVH_SAH and VP_SCENE_BVH_MEDIAN exclude each other");
    if ((flags & VP_SCENE_WALK_F32) && (flags & VP_SCENE_WALK_F16))
        return fail(VP_ERR_INVALID, "VP_SCENE_WALK_F32 and VP_SCENE_WALK_F16 exclude each other");
    if (n == 0) return fail(VP_ERR_INVALID, "scene has no primitives");
    if (n > (1u << 30)) return fail(VP_ERR_INVALID, "scene has more than 2^30 primitives");
    // the validation of loadPrimitivesFromPLY (ply.cpp:146-172)
    for (uint32_t i = 0; i < n; ++i) {
        const vp_primitive& p = prims[i];
        if (!(p.scale[0] > 0.0f && p.scale[1] > 0.0f && p.scale[2] > 0.0f) || !finite3(p.scale))
            return fail(VP_ERR_INVALID, "primitive " + std::to_string(i) + ": scale must be finite and > 0");
        if (!(p.mass > 0.0f) || !std::isfinite(p.mass))
            return fail(VP_ERR_INVALID, "primitive " + std::to_string(i) + ": mass must be finite and > 0");
        if (!finite3(p.center)) return fail(VP_ERR_INVALID, "primitive " + std::to_string(i) + ": non-finite centre");
        const float q2 = p.quat_wxyz[0] * p.quat_wxyz[0] + p.quat_wxyz[1] * p.quat_wxyz[1] +
                         p.quat_wxyz[2] * p.quat_wxyz[2] + p.quat_wxyz[3] * p.quat_wxyz[3];
        if (!(q2 > 0.0f) || !std::isfinite(q2))
            return fail(VP_ERR_INVALID, "primitive " + std::to_string(i) + ": degenerate quaternion");
    }
    VP_HIP_TRY(hipSetDevice(ctx->device));
    auto* sc = new vp_scene();
    sc->ctx = ctx;
    sc->n = n;
    vp_status st = vp_build_scene_device(ctx, prims, n, sc);
    if (st == VP_OK && (flags & (VP_SCENE_BVH_SAH | VP_SCENE_BVH_MEDIAN)))
        st = vp_build_scene_sah(ctx, prims, n, sc, (flags & VP_SCENE_BVH_MEDIAN) != 0u);
    if (st == VP_OK) st = vp_build_walk_tree(ctx, sc, flags & (VP_SCENE_WALK_F32 | VP_SCENE_WALK_F16));
    if (st != VP_OK) {
        vp_scene_destroy(sc);
        return st;
    }
    *out = sc;
    return VP_OK;
}

vp_status vp_scene_destroy(vp_scene* sc) {
    if (!sc) return VP_OK;
    if (sc->ctx && sc->ctx->stream) (void)hipStreamSynchronize(sc->ctx->stream);
    if (sc->prims) (void)hipFree(sc->prims);
    if (sc->nodes) (void)hipFree(sc->nodes);
    if (sc->w4) (void)hipFree(sc->w4);
    if (sc->order) (void)hipFree(sc->order);
    if (sc->morton) (void)hipFree(sc->morton);
    delete sc;
    return VP_OK;
}

vp_status vp_scene_export(vp_scene* sc, uint32_t* order, uint32_t* morton, float* records16, uint32_t* depth) {
    if (!sc) return fail(VP_ERR_INVALID, "scene is null");
    hipStream_t s = sc->ctx->stream;
    if (order) VP_HIP_TRY(hipMemcpyAsync(order, sc->order, sizeof(uint32_t) * sc->n, hipMemcpyDeviceToHost, s));
    if (morton) VP_HIP_TRY(hipMemcpyAsync(morton, sc->morton, sizeof(uint32_t) * sc->n, hipMemcpyDeviceToHost, s));
    if (records16)
        VP_HIP_TRY(hipMemcpyAsync(records16, sc->prims, sizeof(vp::PrimRecord) * sc->n, hipMemcpyDeviceToHost, s));
    VP_HIP_TRY(hipStreamSynchronize(s));
    if (depth) *depth = sc->depth;
    return VP_OK;
}

vp_status vp_scene_walk_info(vp_scene* sc, uint32_t* half_records, double* box_growth) {
    if (!sc) return fail(VP_ERR_INVALID, "scene is null");
    if (half_records) *half_records = sc->w4_half;
    if (box_growth) *box_growth = sc->w4_growth;
    return VP_OK;
}

vp_status vp_env_create(vp_context* ctx, const float* rgba, uint32_t w, uint32_t h, vp_env** out) {
    if (!ctx || !rgba || !out) return fail(VP_ERR_INVALID, "null argument");
    if (w == 0 || h == 0 || static_cast<uint64_t>(w) * h > (1ull << 28))
        return fail(VP_ERR_INVALID, "environment size out of range");
    VP_HIP_TRY(hipSetDevice(ctx->device));
    auto* env = new vp_env();
    env->ctx = ctx;
    env->w = w;
    env->h = h;
    const size_t n = static_cast<size_t>(w) * h;
    // joint = lum709 * sin(theta of the texel centre); total = sum of row sums
    // (host/params/environment_map.h:45-101)
    env->joint_host.assign(n, 0.0f);
    float total = 0.0f;
    for (uint32_t v = 0; v < h; ++v) {
        const float theta = (static_cast<float>(v) + 0.5f) / static_cast<float>(h) * vp::kPi;
        const float st = vp::sinf_(theta);
        float row = 0.0f;
        for (uint32_t u = 0; u < w; ++u) {
            const float* t = rgba + 4 * (static_cast<size_t>(v) * w + u);
            const float lum = 0.2126f * t[0] + 0.7152f * t[1] + 0.0722f * t[2];
            const float wt = lum * st;
            env->joint_host[static_cast<size_t>(v) * w + u] = wt;
            row += wt;
        }
        total += row;
    }
    env->total = total;
    build_alias(env->joint_host, env->alias_host);
    hipStream_t s = ctx->stream;
    if (hipMalloc(&env->tex, n * 4 * sizeof(float)) != hipSuccess ||
        hipMalloc(&env->joint, n * sizeof(float)) != hipSuccess ||
        hipMalloc(&env->alias, n * sizeof(vp::AliasEntry)) != hipSuccess) {
        vp_env_destroy(env);
        return fail(VP_ERR_OOM, "cannot allocate the environment");
    }
    VP_HIP_TRY(hipMemcpyAsync(env->tex, rgba, n * 4 * sizeof(float), hipMemcpyHostToDevice, s));
    VP_HIP_TRY(hipMemcpyAsync(env->joint, env->joint_host.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    VP_HIP_TRY(hipMemcpyAsync(env->alias, env->alias_host.data(), n * sizeof(vp::AliasEntry), hipMemcpyHostToDevice, s));
    VP_HIP_TRY(hipStreamSynchronize(s));
    *out = env;
    return VP_OK;
}

vp_status vp_env_destroy(vp_env* env) {
    if (!env) return VP_OK;
    if (env->ctx && env->ctx->stream) (void)hipStreamSynchronize(env->ctx->stream);
    if (env->tex) (void)hipFree(env->tex);
    if (env->joint) (void)hipFree(env->joint);
    if (env->alias) (void)hipFree(env->alias);
    delete env;
    return VP_OK;
}

vp_status vp_env_export(vp_env* env, float* joint, float* prob, uint32_t* alias, float* total) {
    if (!env) return fail(VP_ERR_INVALID, "env is null");
    const size_t n = env->joint_host.size();
    if (joint) std::memcpy(joint, env->joint_host.data(), n * sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        if (prob) prob[i] = env->alias_host[i].prob;
        if (alias) alias[i] = env->alias_host[i].alias;
    }
    if (total) *total = env->total;
    return VP_OK;
}

// host/params/camera.h:40-70
vp_status vp_camera_perspective(uint32_t width, uint32_t height, const float eye[3], const float target[3],
                                const float up[3], float vfov_deg, vp_camera* out) {
    using namespace vp;
    if (!eye || !target || !up || !out || width == 0 || height == 0)
        return fail(VP_ERR_INVALID, "bad camera arguments");
    const float theta = vfov_deg * kDegToRad;
    const float hh = std::tan(0.5f * theta);
    const f3 lf = f3{eye[0], eye[1], eye[2]}, la = f3{target[0], target[1], target[2]}, vu = f3{up[0], up[1], up[2]};
    const f3 W = normalize(lf - la);
    const f3 U = normalize(cross(vu, W));
    const f3 V = cross(W, U);
    const float focal = sqrtf_(length2(lf - la));
    const float aspect = static_cast<float>(width) * rcpf(static_cast<float>(height));
    const float vh = 2.0f * hh * focal;
    const float vw = vh * aspect;
    const f3 vpu = U * vw;
    const f3 vpv = (-V) * vh;
    const f3 pdu = vpu * rcpf(static_cast<float>(width));
    const f3 pdv = vpv * rcpf(static_cast<float>(height));
    const f3 ul = (lf - W * focal) - fma3(0.5f, vpv - vpu, vpu);
    const f3 p00 = ul + fma3(0.5f, pdv - pdu, pdu);
    const f3 rel = p00 - lf;
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < 3; ++k) out->eye[k] = eye[k];
    out->pixel00_rel[0] = rel.x, out->pixel00_rel[1] = rel.y, out->pixel00_rel[2] = rel.z;
    out->pixel_du[0] = pdu.x, out->pixel_du[1] = pdu.y, out->pixel_du[2] = pdu.z;
    out->pixel_dv[0] = pdv.x, out->pixel_dv[1] = pdv.y, out->pixel_dv[2] = pdv.z;
    out->view_dir[2] = 1.0f;
    out->orthographic = 0.0f;
    return VP_OK;
}

// host/params/camera.h:72-111
vp_status vp_camera_orthographic(uint32_t width, uint32_t height, const float eye[3], const float target[3],
                                 const float up[3], float ortho_height, vp_camera* out) {
    using namespace vp;
    if (!eye || !target || !up || !out || width == 0 || height == 0)
        return fail(VP_ERR_INVALID, "bad camera arguments");
    const f3 lf = f3{eye[0], eye[1], eye[2]}, la = f3{target[0], target[1], target[2]}, vu = f3{up[0], up[1], up[2]};
    const f3 vd = normalize(la - lf);
    const f3 right = normalize(cross(vd, vu));
    const f3 upv = cross(right, vd);
    const float aspect = static_cast<float>(width) * rcpf(static_cast<float>(height));
    const float vh = ortho_height;
    const float vw = vh * aspect;
    const f3 vpu = right * vw;
    const f3 vpv = upv * vh;
    const f3 pdu = vpu * rcpf(static_cast<float>(width));
    const f3 pdv = vpv * rcpf(static_cast<float>(height));
    const f3 ul = lf - fma3(0.5f, vpv - vpu, vpu);
    const f3 p00 = ul + fma3(0.5f, pdv - pdu, pdu);
    const f3 rel = p00 - lf;
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < 3; ++k) out->eye[k] = eye[k];
    out->pixel00_rel[0] = rel.x, out->pixel00_rel[1] = rel.y, out->pixel00_rel[2] = rel.z;
    out->pixel_du[0] = pdu.x, out->pixel_du[1] = pdu.y, out->pixel_du[2] = pdu.z;
    out->pixel_dv[0] = pdv.x, out->pixel_dv[1] = pdv.y, out->pixel_dv[2] = pdv.z;
    out->view_dir[0] = vd.x, out->view_dir[1] = vd.y, out->view_dir[2] = vd.z;
    out->orthographic = 1.0f;
    return VP_OK;
}

// test_runner defaults (test/test_runner.cpp:38-57)
void vp_default_params(vp_render_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->width = 1920;
    p->height = 1080;
    p->spp = 64;
    p->seed = 42;
    p->max_depth = 128;
    p->rr_depth = 12;
    p->rr_max_survival = 0.99f;
    p->firefly_clamp = 0.0f;
    p->filter_stddev = 0.0f;
    p->phase_g = 0.85f;
    p->use_ris = 0;
    p->ris_candidates = 6;
    p->measure_caps = 0;
    p->active_cap = 128;
    p->adaptive_threshold = 0.0f;   // ENABLE_ADAPTIVE_SAMPLING = false (constants.cuh:196-199)
    p->adaptive_min_samples = 32;  // ADAPTIVE_MIN_SAMPLES (constants.cuh:204)
}

vp_status vp_render_device(vp_context* ctx, vp_scene* sc, vp_env* env, const vp_camera* cam,
                           const vp_render_params* params, const vp_launch_config* launch, float* out_rgba_dev,
                           float* out_samples_dev, void* stream, vp_render_stats* stats) {
    if (!ctx || !sc || !env || !cam || !out_rgba_dev) return fail(VP_ERR_INVALID, "null argument");
    vp_status st = vp_check_params(params);
    if (st != VP_OK) return st;
    const auto t0 = std::chrono::steady_clock::now();
    VP_HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    const uint32_t cap = params->active_cap ? params->active_cap : 128u;
    std::vector<JobRun> runs(1);
    runs[0].desc = vp_make_job(sc, env, cam, params);
    runs[0].mean = reinterpret_cast<float4*>(out_rgba_dev);
    runs[0].samples_all = reinterpret_cast<float4*>(out_samples_dev);
    if (launch && launch->batch_spp) runs[0].slice_spp = launch->batch_spp;
    st = vp_reset_counters(ctx, s);
    if (st != VP_OK) return st;
    double kms = 0.0;
    uint64_t paths = 0;
    st = vp_run_jobs(ctx, runs, launch, cap, 0, params->measure_caps ? 1 : 0, s, &kms, &paths);
    if (st != VP_OK) return st;
    st = vp_read_stats(ctx, s, stats);
    if (st != VP_OK) return st;
    if (stats) {
        stats->kernel_ms = kms;
        stats->launches = ctx->last_launches;
        stats->paths = runs[0].traced;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return VP_OK;
}

vp_status vp_render_count_events(vp_context* ctx, vp_scene* sc, vp_env* env, const vp_camera* cam,
                                 const vp_render_params* params, uint64_t* events, uint32_t n_events) {
    if (!ctx || !sc || !env || !cam || !events) return fail(VP_ERR_INVALID, "null argument");
    vp_status st = vp_check_params(params);
    if (st != VP_OK) return st;
    VP_HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const size_t bytes = static_cast<size_t>(params->width) * params->height * sizeof(float4);
    Dev mean;
    if (hipMalloc(&mean.p, bytes) != hipSuccess) return fail(VP_ERR_OOM, "cannot allocate the image");
    std::vector<JobRun> runs(1);
    runs[0].desc = vp_make_job(sc, env, cam, params);
    runs[0].mean = static_cast<float4*>(mean.p);
    st = vp_reset_counters(ctx, s);
    if (st != VP_OK) return st;
    st = vp_run_jobs(ctx, runs, nullptr, params->active_cap ? params->active_cap : 128u, 1,
                  params->measure_caps ? 1 : 0, s, nullptr, nullptr);
    if (st != VP_OK) return st;
    unsigned long long c[vp::kNumCounters];
    VP_HIP_TRY(hipMemcpyAsync(c, ctx->counters, sizeof(c), hipMemcpyDeviceToHost, s));
    VP_HIP_TRY(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < n_events && i < static_cast<uint32_t>(vp::kNumCounters); ++i) events[i] = c[i];
    return VP_OK;
}

vp_status vp_render(vp_context* ctx, vp_scene* sc, vp_env* env, const vp_camera* cam, const vp_render_params* params,
                    float* out_rgba, vp_render_stats* stats) {
    if (!ctx || !out_rgba) return fail(VP_ERR_INVALID, "null argument");
    vp_status st = vp_check_params(params);
    if (st != VP_OK) return st;
    const auto t0 = std::chrono::steady_clock::now();
    VP_HIP_TRY(hipSetDevice(ctx->device));
    const size_t bytes = static_cast<size_t>(params->width) * params->height * sizeof(float4);
    Dev mean;
    if (hipMalloc(&mean.p, bytes) != hipSuccess) return fail(VP_ERR_OOM, "cannot allocate the image");
    st = vp_render_device(ctx, sc, env, cam, params, nullptr, static_cast<float*>(mean.p), nullptr, nullptr, stats);
    if (st != VP_OK) return st;
    VP_HIP_TRY(hipMemcpyAsync(out_rgba, mean.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (stats)
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return VP_OK;
}

vp_status vp_normalize_accumulator(float* out_rgba_dev, const float* mean_rgba_dev, size_t n, float scale,
                                   void* stream) {
    if (!out_rgba_dev || !mean_rgba_dev) return fail(VP_ERR_INVALID, "null argument");
    VP_HIP_TRY(vp_launch_normalize(reinterpret_cast<float4*>(out_rgba_dev),
                                   reinterpret_cast<const float4*>(mean_rgba_dev), n, scale,
                                   static_cast<hipStream_t>(stream)));
    return VP_OK;
}

vp_status vp_render_jobs(vp_context* ctx, vp_job* jobs, uint32_t njobs, const vp_launch_config* launch,
                         vp_render_stats* stats) {
    if (!ctx || !jobs || njobs == 0) return fail(VP_ERR_INVALID, "no jobs");
    const auto t0 = std::chrono::steady_clock::now();
    VP_HIP_TRY(hipSetDevice(ctx->device));
    uint32_t cap = 0;
    int measure = 0;
    uint64_t samples = 0;
    std::vector<JobRun> runs(njobs);
    std::vector<Dev> means(njobs);
    for (uint32_t i = 0; i < njobs; ++i) {
        const vp_job& j = jobs[i];
        if (!j.scene || !j.env || !j.out_rgba) return fail(VP_ERR_INVALID, "job " + std::to_string(i) + ": null field");
        vp_status st = vp_check_params(&j.params);
        if (st != VP_OK) return st;
        cap = std::max(cap, j.params.active_cap ? j.params.active_cap : 128u);
        measure |= j.params.measure_caps ? 1 : 0;
        runs[i].desc = vp_make_job(j.scene, j.env, &j.camera, &j.params);
        const size_t bytes = static_cast<size_t>(j.params.width) * j.params.height * sizeof(float4);
        if (hipMalloc(&means[i].p, bytes) != hipSuccess) return fail(VP_ERR_OOM, "cannot allocate job images");
        runs[i].mean = static_cast<float4*>(means[i].p);
        if (launch && launch->batch_spp) runs[i].slice_spp = launch->batch_spp;
        samples += static_cast<uint64_t>(j.params.width) * j.params.height * j.params.spp;
    }
    // one capacity serves the whole launch: the largest any job or the
    // launch configuration asks for
    if (launch && launch->active_cap) cap = std::max(cap, launch->active_cap);
    if (cap > vp::kMaxActiveCap) return fail(VP_ERR_INVALID, "active_cap must be at most 65535");
    vp_status st = vp_reset_counters(ctx, ctx->stream);
    if (st != VP_OK) return st;
    double kms = 0.0;
    st = vp_run_jobs(ctx, runs, launch, cap, 0, measure, ctx->stream, &kms, nullptr);
    if (st != VP_OK) return st;
    samples = 0;
    for (uint32_t i = 0; i < njobs; ++i) {
        const size_t bytes = static_cast<size_t>(jobs[i].params.width) * jobs[i].params.height * sizeof(float4);
        VP_HIP_TRY(hipMemcpyAsync(jobs[i].out_rgba, means[i].p, bytes, hipMemcpyDeviceToHost, ctx->stream));
        samples += runs[i].traced;
    }
    st = vp_read_stats(ctx, ctx->stream, stats);
    if (st != VP_OK) return st;
    if (stats) {
        stats->kernel_ms = kms;
        stats->launches = ctx->last_launches;
        stats->paths = samples;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return VP_OK;
}

vp_status vp_transmittance_rays(vp_context* ctx, vp_scene* sc, const float* rays, uint32_t m, float* T) {
    if (!ctx || !sc || !rays || !T) return fail(VP_ERR_INVALID, "null argument");
    if (m == 0) return VP_OK;
    VP_HIP_TRY(hipSetDevice(ctx->device));
    Dev dr, dt;
    VP_HIP_TRY(hipMalloc(&dr.p, sizeof(float) * 6 * m));
    VP_HIP_TRY(hipMalloc(&dt.p, sizeof(float) * m));
    VP_HIP_TRY(hipMemcpyAsync(dr.p, rays, sizeof(float) * 6 * m, hipMemcpyHostToDevice, ctx->stream));
    VP_HIP_TRY(vp_launch_transmittance(sc->view(), static_cast<const float*>(dr.p), m, static_cast<float*>(dt.p),
                                       ctx->stream));
    VP_HIP_TRY(hipMemcpyAsync(T, dt.p, sizeof(float) * m, hipMemcpyDeviceToHost, ctx->stream));
    VP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VP_OK;
}

vp_status vp_collect_overlaps(vp_context* ctx, vp_scene* sc, const float* rays, uint32_t m, uint32_t cap,
                              uint32_t* counts, float* t, uint32_t* idx) {
    if (!ctx || !sc || !rays || !counts || !t || !idx) return fail(VP_ERR_INVALID, "null argument");
    if (m == 0) return VP_OK;
    if (cap == 0) return fail(VP_ERR_INVALID, "cap must be positive");
    VP_HIP_TRY(hipSetDevice(ctx->device));
    Dev dr, dc, dt, di;
    VP_HIP_TRY(hipMalloc(&dr.p, sizeof(float) * 6 * m));
    VP_HIP_TRY(hipMalloc(&dc.p, sizeof(uint32_t) * m));
    VP_HIP_TRY(hipMalloc(&dt.p, sizeof(float) * static_cast<size_t>(m) * cap));
    VP_HIP_TRY(hipMalloc(&di.p, sizeof(uint32_t) * static_cast<size_t>(m) * cap));
    VP_HIP_TRY(hipMemsetAsync(dt.p, 0, sizeof(float) * static_cast<size_t>(m) * cap, ctx->stream));
    VP_HIP_TRY(hipMemsetAsync(di.p, 0, sizeof(uint32_t) * static_cast<size_t>(m) * cap, ctx->stream));
    VP_HIP_TRY(hipMemcpyAsync(dr.p, rays, sizeof(float) * 6 * m, hipMemcpyHostToDevice, ctx->stream));
    VP_HIP_TRY(vp_launch_overlaps(sc->view(), static_cast<const float*>(dr.p), m, cap, static_cast<uint32_t*>(dc.p),
                                  static_cast<float*>(dt.p), static_cast<uint32_t*>(di.p), ctx->stream));
    VP_HIP_TRY(hipMemcpyAsync(counts, dc.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, ctx->stream));
    VP_HIP_TRY(hipMemcpyAsync(t, dt.p, sizeof(float) * static_cast<size_t>(m) * cap, hipMemcpyDeviceToHost, ctx->stream));
    VP_HIP_TRY(hipMemcpyAsync(idx, di.p, sizeof(uint32_t) * static_cast<size_t>(m) * cap, hipMemcpyDeviceToHost,
                              ctx->stream));
    VP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VP_OK;
}

vp_status vp_math_eval(vp_context* ctx, int fn, const float* x, const float* y, uint32_t n, float* out) {
    if (!ctx || !x || !out) return fail(VP_ERR_INVALID, "null argument");
    if (n == 0) return VP_OK;
    VP_HIP_TRY(hipSetDevice(ctx->device));
    Dev dx, dy, dout;
    VP_HIP_TRY(hipMalloc(&dx.p, sizeof(float) * n));
    VP_HIP_TRY(hipMalloc(&dy.p, sizeof(float) * n));
    VP_HIP_TRY(hipMalloc(&dout.p, sizeof(float) * n));
    VP_HIP_TRY(hipMemcpyAsync(dx.p, x, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    if (y) VP_HIP_TRY(hipMemcpyAsync(dy.p, y, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    else VP_HIP_TRY(hipMemsetAsync(dy.p, 0, sizeof(float) * n, ctx->stream));
    VP_HIP_TRY(vp_launch_math(fn, static_cast<const float*>(dx.p), static_cast<const float*>(dy.p), n,
                              static_cast<float*>(dout.p), ctx->stream));
    VP_HIP_TRY(hipMemcpyAsync(out, dout.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    VP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VP_OK;
}

vp_status vp_prim_eval(vp_context* ctx, int fn, const float* record16, const float* args9, uint32_t m, float* out) {
    if (!ctx || !record16 || !args9 || !out) return fail(VP_ERR_INVALID, "null argument");
    if (m == 0) return VP_OK;
    VP_HIP_TRY(hipSetDevice(ctx->device));
    Dev drec, dargs, dout;
    VP_HIP_TRY(hipMalloc(&drec.p, sizeof(vp::PrimRecord)));
    VP_HIP_TRY(hipMalloc(&dargs.p, sizeof(float) * 9 * m));
    VP_HIP_TRY(hipMalloc(&dout.p, sizeof(float) * m));
    VP_HIP_TRY(hipMemcpyAsync(drec.p, record16, sizeof(vp::PrimRecord), hipMemcpyHostToDevice, ctx->stream));
    VP_HIP_TRY(hipMemcpyAsync(dargs.p, args9, sizeof(float) * 9 * m, hipMemcpyHostToDevice, ctx->stream));
    VP_HIP_TRY(vp_launch_prim(fn, static_cast<const vp::PrimRecord*>(drec.p), static_cast<const float*>(dargs.p), m,
                              static_cast<float*>(dout.p), ctx->stream));
    VP_HIP_TRY(hipMemcpyAsync(out, dout.p, sizeof(float) * m, hipMemcpyDeviceToHost, ctx->stream));
    VP_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VP_OK;
}

}  // extern "C"
