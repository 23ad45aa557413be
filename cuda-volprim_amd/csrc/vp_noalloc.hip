// vp_noalloc.hip — a render entry point that allocates nothing.
//
// vp_render_device_noalloc renders one job like vp_render_device, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  Every buffer the render
// needs beyond the caller's scene, environment and output image lives in one
// file-scope device array, `scratch`, of kNoAllocScratchBytes:
//
//   [accumulators]  width * height * 8 u64 pixel sums (<= 512 x 512 pixels)
//   [tables]        one job-table entry, its path prefix, the launch
//                   descriptor, the path counter and the event counters
//   [slab]          the megakernel's per-lane active sets; the launch uses as
//                   many lanes as the slab holds at the request's capacity
//
// Each device that loads the library holds its own copy of the array until
// the process ends.  Host code never passes the array's name to a kernel (in
// host code that name is a host-side stand-in, not a device address): at
// context creation, with the context's device current, hipGetSymbolAddress
// gives the device address, which the context keeps (vp_noalloc_attach).
//
// Calls on one device share its array and leave the last call's contents in
// it; every call clears what must start at zero (accumulators, counters, path
// counter) with kernels, not hipMemsetAsync, so a captured sequence stays
// valid on replay.  Calls that overlap on one device race on the array.
#include "vp_device.h"
#include "vp_internal.h"

namespace vp {
namespace {

constexpr uint32_t kNoAllocMaxPixels = 512u * 512u;
constexpr size_t kNoAllocAccumBytes = size_t(kNoAllocMaxPixels) * kAccumWords * sizeof(unsigned long long);
constexpr size_t kNoAllocTableBytes = 4096;
constexpr size_t kNoAllocSlabBytes = size_t(112) << 20;
constexpr size_t kNoAllocScratchBytes = kNoAllocAccumBytes + kNoAllocTableBytes + kNoAllocSlabBytes;
static_assert(sizeof(JobDesc) + 4 * 8 + kClaimParts * 8 + kNumCounters * 8 <= kNoAllocTableBytes,
              "tables fit their block");

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Offsets into the scratch array.
struct NoAllocLayout {
    unsigned long long* accum;
    JobDesc* job;
    uint64_t* job_begin;     // [1]
    uint64_t* launch_desc;   // [3]: job count, end and start of the path range
    unsigned long long* path_counter;  // [kClaimParts]
    unsigned long long* counters;  // [kNumCounters]
    uint32_t* slab;
};
NoAllocLayout layout(void* base) {
    unsigned char* b = static_cast<unsigned char*>(base);
    NoAllocLayout l;
    l.accum = reinterpret_cast<unsigned long long*>(b);
    unsigned char* t = b + kNoAllocAccumBytes;
    l.job = reinterpret_cast<JobDesc*>(t);
    l.job_begin = reinterpret_cast<uint64_t*>(t + sizeof(JobDesc));
    l.launch_desc = l.job_begin + 1;
    l.path_counter = reinterpret_cast<unsigned long long*>(l.job_begin + 4);
    l.counters = l.path_counter + kClaimParts;
    l.slab = reinterpret_cast<uint32_t*>(b + kNoAllocAccumBytes + kNoAllocTableBytes);
    return l;
}

// Installs the one-entry job table and clears the counters: the entry comes
// in as a kernel argument, so no host staging buffer is needed.
__global__ void k_noalloc_install(JobDesc J, JobDesc* job, uint64_t* job_begin, uint64_t* launch_desc,
                                  unsigned long long* path_counter, unsigned long long* counters) {
    const uint32_t i = threadIdx.x;
    if (i == 0) {
        *job = J;
        job_begin[0] = 0;
        launch_desc[0] = 1;
        launch_desc[1] = J.path_count;
        launch_desc[2] = 0;
    }
    if (i < kClaimParts) path_counter[i] = 0ull;
    if (i < static_cast<uint32_t>(kNumCounters)) counters[i] = 0ull;
}

vp_status fail(vp_status code, const std::string& msg) {
    vp_set_error(msg);
    return code;
}

}  // namespace
}  // namespace vp

// Device address of this device's scratch array, kept by the context.  Called
// from vp_context_create with the context's device current.
vp_status vp_noalloc_attach(vp_context* ctx) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(vp::scratch));
    if (e != hipSuccess || p == nullptr)
        return vp::fail(VP_ERR_HIP, std::string("cannot resolve the no-allocation scratch array: ") +
                                        hipGetErrorString(e));
    ctx->noalloc_scratch = p;
    return VP_OK;
}

extern "C" vp_status vp_render_device_noalloc(vp_context* ctx, vp_scene* sc, vp_env* env, const vp_camera* cam,
                                              const vp_render_params* params, float* out_rgba_dev, void* stream,
                                              vp_render_stats* stats) {
    using namespace vp;
    if (!ctx || !sc || !env || !cam || !params || !out_rgba_dev) return fail(VP_ERR_INVALID, "null argument");
    vp_status st = vp_check_params(params);
    if (st != VP_OK) return st;
    if (params->adaptive_threshold > 0.0f)
        return fail(VP_ERR_INVALID, "vp_render_device_noalloc: adaptive sampling needs per-pixel state beyond the "
                                    "scratch array; use vp_render_device");
    const uint64_t npix = static_cast<uint64_t>(params->width) * params->height;
    if (npix > kNoAllocMaxPixels)
        return fail(VP_ERR_INVALID, "vp_render_device_noalloc: " + std::to_string(params->width) + "x" +
                                        std::to_string(params->height) + " exceeds the scratch limit of " +
                                        std::to_string(kNoAllocMaxPixels) + " pixels (512x512)");
    if (ctx->noalloc_scratch == nullptr) return fail(VP_ERR_HIP, "vp_render_device_noalloc: no scratch address");
    const uint32_t cap = params->active_cap ? params->active_cap : 128u;
    // the megakernel, at the context's tuned (or default) shape, with no
    // more lanes than the slab holds at this capacity
    vp_launch_config c{};
    if (ctx->have_tuned) c = ctx->tuned;
    c.pipeline = VP_PIPELINE_MEGAKERNEL;
    RenderLaunch L = vp_resolve_launch(ctx, &c, cap, 0, params->measure_caps ? 1 : 0);
    const size_t lanes_max = kNoAllocSlabBytes / (sizeof(uint32_t) * kMegaSlabBuffers * static_cast<size_t>(mega_slab_stride(cap)));
    const size_t grid_max = lanes_max / L.block;
    if (grid_max == 0)
        return fail(VP_ERR_INVALID, "vp_render_device_noalloc: active_cap " + std::to_string(cap) +
                                        " leaves no room for one workgroup in the scratch slab");
    if (L.grid > grid_max) L.grid = static_cast<uint32_t>(grid_max);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    VP_HIP_TRY(hipSetDevice(ctx->device));
    const NoAllocLayout lay = layout(ctx->noalloc_scratch);
    JobDesc J = vp_make_job(sc, env, cam, params);
    J.accum = lay.accum;
    J.mean = reinterpret_cast<float4*>(out_rgba_dev);
    J.samples = nullptr;
    J.done = nullptr;
    J.s_begin = 0;
    J.s_count = params->spp;
    J.path_begin = 0;
    J.path_count = static_cast<uint64_t>(J.npad) * J.s_count;
    VP_HIP_TRY(vp_launch_fill_u64(lay.accum, static_cast<uint32_t>(npix * kAccumWords), 0ull, s));
    hipLaunchKernelGGL(k_noalloc_install, dim3(1), dim3(64), 0, s, J, lay.job, lay.job_begin, lay.launch_desc,
                       lay.path_counter, lay.counters);
    VP_HIP_TRY(hipGetLastError());
    VP_HIP_TRY(vp_launch_render(L, lay.job, lay.job_begin, lay.launch_desc, lay.path_counter, lay.slab, lay.counters,
                                s));
    VP_HIP_TRY(vp_launch_resolve_job(lay.job, 0, static_cast<uint32_t>(npix), s));
    if (stats) {
        unsigned long long cnt[kNumCounters];
        VP_HIP_TRY(hipMemcpyAsync(cnt, lay.counters, sizeof(cnt), hipMemcpyDeviceToHost, s));
        VP_HIP_TRY(hipStreamSynchronize(s));
        *stats = vp_render_stats{};
        stats->active_overflow = cnt[kCntActiveOverflow];
        stats->max_hits_per_ray = static_cast<uint32_t>(cnt[kCntMaxHits]);
        stats->max_point_overlap = static_cast<uint32_t>(cnt[kCntMaxOverlap]);
        stats->launches = 1;
        stats->paths = npix * params->spp;
    }
    return VP_OK;
}
