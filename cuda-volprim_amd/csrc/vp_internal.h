// vp_internal.h — device-side data layout and host-side object definitions
// shared by the build kernels (vp_bvh.hip), the render kernels
// (vp_render.hip) and the C ABI (vp_api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/vp_api.h"
#include "vp_core.h"
#include "vp_walk.h"

namespace vp {

// ---------------------------------------------------------------------------
// BVH: 2N-1 nodes in depth-first preorder, 32 bytes each (two float4):
//   a = (min.x, min.y, min.z, skip)   skip = index of the next node when the
//                                     subtree is skipped (2N-1 ends the walk)
//   b = (max.x, max.y, max.z, prim)   prim = stored primitive index of a leaf,
//                                     0xffffffff for an inner node
// A hit inner node continues at i + 1 (its left child).  Leaves appear in
// stored-primitive order, so a walk meets primitives in ascending index.
// ---------------------------------------------------------------------------
constexpr uint32_t kInnerNode = 0xffffffffu;

// The walk tree the render kernels traverse is defined in vp_walk.h; its leaf
// boxes are the rope tree's, so a walk of either meets the same leaves.

// Conservative leaf box of record r (input scale s, centre c): the 3-sigma
// ellipsoid with forward rotation R and scale s has half extent
// 3 sqrt(sum_j (R_ij s_j)^2) along world axis i; the box is widened by a
// relative 2^-10 and an absolute 2^-20 (|c| + h) so every point the FP32
// entry/containment tests accept lies inside it.  Evaluated in double.
VP_HD void conservative_box(const PrimRecord& r, const float s[3], float lo[3], float hi[3]) {
    // forward rotation = conjugate of the stored one
    const double qw = r.qw, qx = -(double)r.qx, qy = -(double)r.qy, qz = -(double)r.qz;
    const double R[3][3] = {
        {1.0 - 2.0 * (qy * qy + qz * qz), 2.0 * (qx * qy - qw * qz), 2.0 * (qx * qz + qw * qy)},
        {2.0 * (qx * qy + qw * qz), 1.0 - 2.0 * (qx * qx + qz * qz), 2.0 * (qy * qz - qw * qx)},
        {2.0 * (qx * qz - qw * qy), 2.0 * (qy * qz + qw * qx), 1.0 - 2.0 * (qx * qx + qy * qy)},
    };
    const double sv[3] = {s[0], s[1], s[2]};
    const double cv[3] = {r.cx, r.cy, r.cz};
    for (int a = 0; a < 3; ++a) {
        double h2 = 0.0;
        for (int j = 0; j < 3; ++j) h2 += (R[a][j] * sv[j]) * (R[a][j] * sv[j]);
        double h = 3.0 * sqrt(h2);
        h = h * (1.0 + 0x1p-10) + 0x1p-20 * (fabs(cv[a]) + h);
        lo[a] = (float)(cv[a] - h);
        hi[a] = (float)(cv[a] + h);
    }
}

struct SceneView {
    const PrimRecord* prims;
    const float4* nodes;
    uint32_t n_prims;
    uint32_t n_nodes;
    const float4* w4;    // walk-tree records (kW4RecordF4 float4 each, or 16-bit records)
    uint32_t w4_levels;  // its level count L
    uint32_t w4_half;    // 1: w4 holds 16-bit records (kW4hRecordB bytes each)
};

struct EnvDev {
    const float* tex;
    const AliasEntry* alias;
    const float* joint;
    uint32_t w, h;
    float total;
    uint32_t has_table;
};

// One job slice of the persistent renderer: samples [s_begin, s_begin+s_count)
// of one image.  Path q of the slice (0 <= q < path_count = npad s_count) is
// sample s_begin + q % s_count of the pixel at tile position q / s_count
// (8x8 tiles, in bands of tile rows, path_begin).
// A finished path adds its radiance to its pixel's accumulator (kAccumWords
// words per pixel, vp_core.h); the mean is resolved from it after the last
// slice.  Entries are aligned to the 128-byte cache line so a render launch
// that is reading table entries never shares a line with entries appended
// after it started (the server appends while earlier launches run).
struct alignas(128) JobDesc {
    SceneView scene;
    EnvDev env;
    CameraPod cam;
    PhaseParams hg;
    uint32_t width, height, spp, seed;
    uint32_t max_depth, rr_depth;
    float rr_max_survival, firefly_clamp, filter_stddev, shadow_rr;
    uint32_t use_ris, ris_candidates;
    uint32_t tiles_x, npad;
    uint32_t s_begin, s_count;
    uint64_t path_begin, path_count;
    unsigned long long* accum;  // width * height * kAccumWords radiance sums of the job
    float4* mean;               // width * height mean, written by the resolve kernel
    float4* samples;            // optional (null = none): per-sample radiance of this slice, sample-major,
                                // s_count * W * H (sample s_begin + k of pixel i at k * W * H + i)
    unsigned long long* done;   // optional (null = none): finished paths of this entry (server)
    // adaptive sampling (null adapt_state = off): the slices are the batches
    // of the reference's launch loop; a pixel whose state has bit 31 set has
    // converged and traces no more paths.  The resolve then folds the slice's
    // samples (samples, required) into the pixel's Welford mean (in mean) and
    // M2, and sets bit 31 for the next slice once the criterion holds; the low
    // 31 bits count the pixel's samples.
    float adapt_threshold;
    uint32_t adapt_min;
    float4* adapt_m2;
    uint32_t* adapt_state;
};
static_assert(sizeof(JobDesc) % 128 == 0, "job-table entries fill whole cache lines");

// Per-slot slab of the wavefront pipeline: two active-set buffers and one
// buffer of rebuild candidates, each of the active-set capacity.
constexpr uint32_t kSlabBuffers = 3;
// Per-lane slab of the megakernel, buffers of the active-set capacity:
//   0, 1  active sets (ping-pong)
//   2     rebuild candidates (stored primitive index)
//   3, 4  their entry and exit distances along the ray (float bits)
//   5     exit distance of each member of the current active set along the
//         current ray
// The argmin computes the values of 3, 4 and 5 anyway, so the rebuild only
// compares them with the scatter distance.
constexpr uint32_t kMegaSlabBuffers = 6;
// Words between the megakernel's slab buffers: the capacity rounded up to a
// multiple of 4, so every buffer starts 16-byte aligned and can be read four
// entries at a time.
VP_HD uint32_t mega_slab_stride(uint32_t cap) { return (cap + 3u) & ~3u; }
// Path counters of one megakernel launch: its path range is split into this
// many parts, one per XCD of the MI355X, and a wavefront claims paths from its
// own XCD's part first (so each XCD's L2 serves a narrower band of the
// image), then from the others once its part is used up.
constexpr uint32_t kClaimParts = 8;
// Largest active-set capacity: the walks count set sizes in 16 bits.
constexpr uint32_t kMaxActiveCap = 65535;

// Path pool of the wavefront pipeline (structure of arrays over P slots).
struct WfPool {
    float4* ob;    // origin.xyz, bounce
    float4* dj;    // direction.xyz, job
    float4* bt;    // throughput.xyz, scatter distance
    float4* lr;    // radiance.xyz, active-set size
    uint4* rng;    // PCG32 state, increment
    uint4* misc;   // key (2 words), pixel index, flags (set buffer, state, pending terms)
    float4* pw;    // [3][P] pending term factor.xyz (the factor its transmittance multiplies)
    float4* po;    // [3][P] shadow-ray origin.xyz, set buffer | count << 1
    float4* pd;    // [3][P] shadow-ray direction.xyz, transmittance
    uint32_t* live;      // [P] live-slot queue
    uint32_t* shadow;    // [3P] shadow-ray queue (slot << 2 | term)
    uint32_t* free_q;    // [P] free-slot queue
    unsigned int* qc;    // [2][8] per parity: live, shadow, free, claimed, extend and shadow fetch heads
    unsigned long long* claim_base;  // [2] first path number claimed per parity
    unsigned long long* finished;    // paths retired (or padding) so far
    unsigned long long* job_done;    // per job-table entry: paths retired so far
    uint32_t* slab;      // [P][3][cap]: each slot's two active sets and its rebuild candidates
    uint32_t* ncand;     // [P] rebuild candidates recorded by the last extend walk
    uint32_t P, cap;
};
struct WfLaunch {
    uint32_t grid_regen, grid_extend, grid_scatter, grid_shadow;
    uint32_t fetch_min;  // idle lanes a wavefront waits for before taking new walks
};

// Counter slots of the render kernel.
enum CounterSlot : int {
    kCntActiveOverflow = 0,
    kCntMaxHits = 1,
    kCntMaxOverlap = 2,
    kCntPaths = 3,
    kCntTrace = 4,         // collect traversals (one per bounce)
    kCntScatter = 5,
    kCntShadowRays = 6,
    kCntArgHit = 7,        // entered primitives evaluated by the argmin
    kCntArgActive = 8,     // origin-set primitives evaluated by the argmin
    kCntArgAccept = 9,     // candidates that beat the running minimum
    kCntRbActive = 10,     // rebuild: origin-set members tested
    kCntRbHit = 11,        // rebuild: point-query candidates tested as entries
    kCntAlb = 12,          // albedo terms
    kCntShActive = 13,     // shadow: origin-set chords integrated
    kCntShHit = 14,        // shadow: entered chords integrated
    kCntLeafTests = 15,    // entry tests at BVH leaves (all traversals)
    kCntNodeVisits = 16,   // BVH nodes loaded (all traversals)
    kCntContain = 17,      // containment tests (point queries)
    kCntVisitsShadow = 18, // BVH nodes loaded by shadow-ray walks (part of kCntNodeVisits)
    kCntVisitsPoint = 19,  // BVH nodes loaded by point queries (part of kCntNodeVisits)
    kNumCounters = 20,
};

}  // namespace vp

// ---------------------------------------------------------------------------
// Host objects behind the opaque C handles.
// ---------------------------------------------------------------------------
struct vp_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    // persistent render resources, grown on demand
    uint32_t* slab = nullptr;
    size_t slab_words = 0;
    unsigned long long* counters = nullptr;  // vp::kNumCounters
    unsigned long long* path_counter = nullptr;
    vp::JobDesc* jobs_dev = nullptr;
    uint32_t jobs_cap = 0;
    uint64_t* job_begin_dev = nullptr;
    unsigned long long* job_done_dev = nullptr;  // per job-table entry (wavefront pipeline)
    uint64_t* launch_desc = nullptr;  // [0] job slices, [1] end of the launch's path range, [2] its start
    // pinned staging of the job table, the slice begins and the launch descriptor
    vp::JobDesc* jobs_pinned = nullptr;
    uint64_t* begin_pinned = nullptr;
    uint64_t* desc_pinned = nullptr;
    uint32_t pinned_cap = 0;
    // captured step graph and the shape it was captured for
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_key[6] = {0, 0, 0, 0, 0, 0};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // wavefront pipeline: path pool, its allocation, the captured
    // two-iteration graph and the pinned completion counters
    vp::WfPool wf{};
    void* wf_mem = nullptr;
    hipGraphExec_t wf_graph = nullptr;
    uint64_t wf_graph_key[6] = {0, 0, 0, 0, 0, 0};
    unsigned long long* wf_done_pinned = nullptr;  // [2]
    hipEvent_t wf_ev[2] = {nullptr, nullptr};
    std::mutex mu;
    uint32_t last_launches = 0;
    vp_launch_config tuned{};
    bool have_tuned = false;
    // device address of this device's no-allocation scratch array
    // (vp_noalloc.hip), resolved once at context creation
    void* noalloc_scratch = nullptr;
};

struct vp_scene {
    vp_context* ctx = nullptr;
    uint32_t n = 0;
    uint32_t depth = 0;
    vp::PrimRecord* prims = nullptr;  // stored (Morton) order
    float4* nodes = nullptr;           // 2 * (2n - 1)
    uint32_t* order = nullptr;         // order[k] = input index
    uint32_t* morton = nullptr;        // codes in input order
    float4* w4 = nullptr;              // walk tree, w4_records(w4_levels) records
    uint32_t w4_levels = 0;
    uint32_t w4_half = 0;              // 1: 16-bit records
    double w4_growth = 0.0;            // summed leaf-box surface of the 16-bit records / 32-bit - 1
    vp::SceneView view() const {
        return vp::SceneView{prims, nodes, n, n ? 2 * n - 1 : 0, w4, w4_levels, w4_half};
    }
};

struct vp_env {
    vp_context* ctx = nullptr;
    uint32_t w = 0, h = 0;
    float total = 0.0f;
    float* tex = nullptr;
    float* joint = nullptr;
    vp::AliasEntry* alias = nullptr;
    std::vector<float> joint_host;
    std::vector<vp::AliasEntry> alias_host;
    vp::EnvDev view() const {
        return vp::EnvDev{tex, alias, joint, w, h, total, (alias != nullptr && total > 0.0f) ? 1u : 0u};
    }
};

// error plumbing (vp_api.cpp)
void vp_set_error(const std::string& msg);
#define VP_HIP_TRY(expr)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            vp_set_error(std::string(#expr) + ": " + hipGetErrorString(e_));               \
            return VP_ERR_HIP;                                                             \
        }                                                                                  \
    } while (0)

// launchers (vp_bvh.hip)
vp_status vp_build_scene_device(vp_context* ctx, const vp_primitive* host_prims, uint32_t n, vp_scene* scene);
// host tree replacing the device LBVH's nodes, records and order (vp_bvh_sah.cpp):
// binned SAH, or (median) balanced median splits on the walk tree's record
// boundaries
vp_status vp_build_scene_sah(vp_context* ctx, const vp_primitive* host_prims, uint32_t n, vp_scene* scene,
                             bool median);
// walk tree over the scene's stored order, from its rope tree's leaf boxes (vp_bvh.hip)
// walk_format: 0 (16-bit records when the leaf boxes grow little),
// VP_SCENE_WALK_F32 or VP_SCENE_WALK_F16
vp_status vp_build_walk_tree(vp_context* ctx, vp_scene* scene, uint32_t walk_format);

// launchers (vp_render.hip)
struct RenderLaunch {
    uint32_t block;
    uint32_t grid;
    uint32_t active_cap;
    uint32_t regen_min;
    int count_events;
    int measure;
    uint32_t pipeline;  // VP_PIPELINE_MEGAKERNEL or VP_PIPELINE_WAVEFRONT
    uint32_t pool;      // wavefront path slots
    vp::WfLaunch wf;
};
hipError_t vp_launch_step(vp_context* ctx, const RenderLaunch& L, uint32_t max_jobs, uint32_t max_npix,
                          hipStream_t s);
// One render launch over an explicit job table, descriptor, path counter and
// slab (the server runs up to two at once, each with its own slab).
hipError_t vp_launch_render(const RenderLaunch& L, const vp::JobDesc* jobs, const uint64_t* job_begin,
                            const uint64_t* launch_desc, unsigned long long* counter, uint32_t* slab,
                            unsigned long long* counters, hipStream_t s);

// execution of job lists (vp_exec.cpp)
struct JobRun {
    vp::JobDesc desc;
    float4* mean = nullptr;         // device W*H mean (output)
    float4* samples_all = nullptr;  // optional caller buffer spp*W*H of per-sample radiance
    uint32_t issued = 0;            // samples launched so far
    uint32_t slice_spp = 0;         // samples per launch (0 = all at once)
    float4* batch = nullptr;        // adaptive sampling: per-sample radiance of one batch
    uint64_t traced = 0;            // samples traced (set when the run completes)
};
// samples per batch of adaptive sampling (BATCH_SIZE, renderer.cpp:32)
constexpr uint32_t kAdaptiveBatch = 16;
vp_status vp_ensure_slab(vp_context* ctx, size_t words);
vp_status vp_ensure_jobs(vp_context* ctx, uint32_t n);
void vp_drop_graph(vp_context* ctx);
RenderLaunch vp_resolve_launch(vp_context* ctx, const vp_launch_config* cfg, uint32_t active_cap, int count,
                               int measure);
// One step over the first nslices entries of the pinned job table (capacity max_jobs).
vp_status vp_run_step(vp_context* ctx, const RenderLaunch& L, uint32_t nslices, uint64_t total_paths,
                      uint32_t max_jobs, uint32_t max_npix, bool use_graph, hipStream_t s);
vp_status vp_run_jobs(vp_context* ctx, std::vector<JobRun>& runs, const vp_launch_config* cfg, uint32_t active_cap,
                      int count, int measure, hipStream_t s, double* kernel_ms, uint64_t* paths);
vp_status vp_read_stats(vp_context* ctx, hipStream_t s, vp_render_stats* stats);
vp_status vp_reset_counters(vp_context* ctx, hipStream_t s);
vp::JobDesc vp_make_job(const vp_scene* sc, const vp_env* env, const vp_camera* cam, const vp_render_params* p);
vp_status vp_check_params(const vp_render_params* p);
// Resolves ctx->noalloc_scratch on ctx's device (which must be current).
vp_status vp_noalloc_attach(vp_context* ctx);
hipError_t vp_launch_fill_u64(unsigned long long* p, uint32_t n, unsigned long long v, hipStream_t s);
hipError_t vp_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
hipError_t vp_launch_normalize(float4* out, const float4* in, size_t n, float scale, hipStream_t s);
hipError_t vp_launch_transmittance(vp::SceneView sc, const float* rays, uint32_t m, float* T, hipStream_t s);
hipError_t vp_launch_overlaps(vp::SceneView sc, const float* rays, uint32_t m, uint32_t cap, uint32_t* counts,
                              float* t, uint32_t* idx, hipStream_t s);
hipError_t vp_launch_math(int fn, const float* x, const float* y, uint32_t n, float* out, hipStream_t s);
hipError_t vp_launch_prim(int fn, const vp::PrimRecord* rec, const float* args, uint32_t m, float* out,
                          hipStream_t s);
int vp_render_occupancy(uint32_t block, uint32_t active_cap, int count_events, int* blocks_per_cu);
// wavefront pipeline launchers (vp_wavefront.hip)
hipError_t vp_wf_launch_init(const vp::WfPool& w, hipStream_t s);
hipError_t vp_wf_launch_iterations(vp_context* ctx, const vp::WfPool& w, const vp::WfLaunch& L, hipStream_t s);
int vp_wf_occupancy(int which, int* blocks_per_cu);
// Means of every entry of ctx's job table from their accumulators.
hipError_t vp_launch_resolve(vp_context* ctx, uint32_t max_jobs, uint32_t max_npix, hipStream_t s);
// Means of job-table entry j of the table at jobs.
hipError_t vp_launch_resolve_job(const vp::JobDesc* jobs, uint32_t j, uint32_t npix, hipStream_t s);
void vp_free_wf(vp_context* ctx);
vp_status vp_ensure_wf(vp_context* ctx, uint32_t P, uint32_t cap);
vp_status vp_wf_prepare(vp_context* ctx, const RenderLaunch& L, hipStream_t s);
