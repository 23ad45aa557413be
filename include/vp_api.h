/* vp_api.h -- C ABI of the MI355X Gaussian-mixture volumetric path tracer.
 *
 * Plain C types only (pointers, sizes, PODs).  Every call returns a
 * vp_status; vp_last_error() gives the message of the last failure on the
 * calling thread.  Calls on one context are externally synchronised (one
 * host thread drives a context, as the reference's Renderer does); the
 * server (vp_server_*) owns its own thread.
 *
 * Reference interfaces each entry point replaces (kfernandez31/cuda-volprim):
 *   vp_scene_create      Renderer::initPrimsAndGAS: Morton sort + GAS/IAS build
 *                        (src/thesis/host/app/renderer.cpp:85-173)
 *   vp_env_create        EnvironmentMap ctor: texture + importance CDF
 *                        (include/thesis/host/params/environment_map.h:45-151)
 *   vp_camera_*          host::params::Camera::createPerspective/createOrthographic
 *                        (include/thesis/host/params/camera.h:40-150)
 *   vp_render            Renderer::render batch loop + optixLaunch(W x H)
 *                        (src/thesis/host/app/renderer.cpp:341-439)
 *   vp_normalize_accumulator
 *                        launch_normalize_accumulator_kernel
 *                        (device/kernels/normalize_accumulator.h:11-13)
 *   vp_render_jobs       (no counterpart: the batched persistent form)
 *   vp_server_*          (no counterpart: the serving form)
 *   vp_transmittance_rays, vp_collect_overlaps, vp_math_eval, vp_prim_eval
 *                        test entry points exposing compute_transmittance_to_env
 *                        (device/core/sampling.cuh:548-583), collect_hits
 *                        (sampling.cuh:287-305) and the primitive math
 *                        (include/thesis/device/params/primitive.h:127-318)
 */
#ifndef VP_API_H
#define VP_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    VP_OK = 0,
    VP_ERR_INVALID = 1,   /* bad argument (null pointer, size, range) */
    VP_ERR_HIP = 2,       /* HIP runtime failure */
    VP_ERR_OOM = 3,       /* device allocation failed or admission refused */
    VP_ERR_NOT_READY = 4, /* server request still running */
    VP_ERR_STATE = 5      /* object in the wrong state for the call */
} vp_status;

typedef struct vp_context vp_context;
typedef struct vp_scene vp_scene;
typedef struct vp_env vp_env;
typedef struct vp_server vp_server;

/* One Gaussian: centre, FORWARD rotation quaternion (w, x, y, z; normalised by
 * the library), per-axis scale (linear, > 0), albedo, total mass M
 * (PLY sigma_t_0 x sigma multiplier, src/thesis/host/utils/io/ply.cpp:107-116). */
typedef struct {
    float center[3];
    float quat_wxyz[4];
    float scale[3];
    float albedo[3];
    float mass;
} vp_primitive;

/* Camera as the device consumes it (include/thesis/device/params/camera.h:17-26). */
typedef struct {
    float eye[3];
    float pixel00_rel[3];
    float pixel_du[3];
    float pixel_dv[3];
    float view_dir[3];
    float orthographic; /* 0 = perspective, 1 = orthographic */
} vp_camera;

/* Render settings: the test_runner flag surface (test/test_runner.cpp:101-129). */
typedef struct {
    uint32_t width, height;
    uint32_t spp;             /* --spp */
    uint32_t seed;            /* --seed */
    uint32_t max_depth;       /* --max-depth (default 128) */
    uint32_t rr_depth;        /* --rr-depth (default 12) */
    float rr_max_survival;    /* --rr-max-survival (default 0.99) */
    float firefly_clamp;      /* --firefly-clamp (0 = off) */
    float filter_stddev;      /* --filter-width (0 = box) */
    float phase_g;            /* --phase-g (default 0.85) */
    uint32_t use_ris;         /* --ris */
    uint32_t ris_candidates;  /* --ris-candidates (default 6) */
    uint32_t measure_caps;    /* --measure-caps */
    uint32_t active_cap;      /* origin/scatter active-set capacity (0 = 128) */
    float shadow_rr;          /* --shadow-rr: NEE shadow-ray roulette threshold eps (0 = off, the reference's
                                 estimator); a shadow ray whose term has fallen below eps survives with
                                 probability 1/4 and counts 4x (unbiased; no counterpart in the reference) */
    float adaptive_threshold; /* adaptive sampling (device/entry/raygen.cuh:37-69,214-233, compile-time off
                                 in the reference: ENABLE_ADAPTIVE_SAMPLING, constants.cuh:196-216); 0 = off.
                                 On, the samples run in batches of 16 (renderer.cpp:32; a launch's batch_spp
                                 overrides it) and a pixel stops once it has adaptive_min_samples or more
                                 and max over channels of sqrt(M2/(n-1)) / max(mean, 1e-6) < threshold; the
                                 image is then the per-pixel Welford mean.  Megakernel renders only; not
                                 with per-sample output, the server or vp_render_device_noalloc. */
    uint32_t adaptive_min_samples; /* ADAPTIVE_MIN_SAMPLES (default 32) */
} vp_render_params;

typedef struct {
    uint64_t active_overflow;   /* active-set inserts dropped at capacity */
    uint64_t hit_overflow;      /* always 0: entry hits are streamed, never dropped */
    uint32_t max_hits_per_ray;  /* --measure-caps: max entries on one traced ray */
    uint32_t max_point_overlap; /* --measure-caps: max primitives containing a vertex */
    double kernel_ms;           /* GPU time of the render-kernel launches (HIP events) */
    uint32_t launches;          /* render-kernel launches the call made */
    double wall_ms;             /* host wall time of the call */
    uint64_t paths;             /* camera paths traced */
} vp_render_stats;

/* Launch configuration chosen by the autotuner or forced by the caller. */
typedef struct {
    uint32_t block;        /* threads per workgroup (64..1024, multiple of 64) */
    uint32_t blocks_per_cu;/* persistent workgroups per CU (0 = occupancy max) */
    uint32_t active_cap;   /* active-set capacity (entries per path) */
    uint32_t batch_spp;    /* samples per pixel per launch (0 = all) */
    uint32_t regen_min;    /* idle lanes a wavefront waits for before claiming new paths (0 = 1) */
    uint32_t pipeline;     /* VP_PIPELINE_MEGAKERNEL (0) or VP_PIPELINE_WAVEFRONT (1) */
    uint32_t pool_paths;   /* wavefront: path slots in flight (0 = 1 Mi) */
} vp_launch_config;
#define VP_PIPELINE_MEGAKERNEL 0u
#define VP_PIPELINE_WAVEFRONT 1u

/* ---- context ---------------------------------------------------------- */
vp_status vp_context_create(int device, vp_context** out);
vp_status vp_context_destroy(vp_context* ctx);
const char* vp_last_error(void);
vp_status vp_device_info(vp_context* ctx, char* name, size_t name_len, uint32_t* num_cus, double* clock_mhz);

/* ---- scene / environment / camera -------------------------------------- */
vp_status vp_scene_create(vp_context* ctx, const vp_primitive* prims, uint32_t n, vp_scene** out);
/* vp_scene_create with build flags (no counterpart: the reference's OptiX
 * builder has one mode).  VP_SCENE_BVH_SAH replaces the GPU LBVH's tree with
 * a binned surface-area-heuristic tree built on the host, in the same
 * layout; renders of the two differ only through the stored primitive order.
 * VP_SCENE_BVH_MEDIAN instead builds, on the host, a balanced tree of median
 * splits that fall on the 4-wide walk tree's record boundaries (each split
 * on the axis whose halves have the smaller summed box surface), so each
 * walk-tree record's children are spatial quarters of its primitives.  The
 * two flags exclude each other.
 * The 4-wide walk tree the render kernels traverse is stored in 64-byte
 * records of binary16 bounds, rounded outward, when that grows the leaf
 * boxes' summed surface by at most 1/16 (scenes whose primitives are small
 * against their distance from the origin keep 128-byte records of binary32
 * bounds).  VP_SCENE_WALK_F32 and VP_SCENE_WALK_F16 force one form; they
 * exclude each other.  The form decides only which primitives a walk tests,
 * never a result: renders are bit-identical either way. */
#define VP_SCENE_BVH_SAH 1u
#define VP_SCENE_BVH_MEDIAN 2u
#define VP_SCENE_WALK_F32 4u
#define VP_SCENE_WALK_F16 8u
vp_status vp_scene_create_ex(vp_context* ctx, const vp_primitive* prims, uint32_t n, uint32_t flags,
                             vp_scene** out);
vp_status vp_scene_destroy(vp_scene* scene);
/* Order the scene stores primitives in (order[k] = input index of the k-th),
 * their 30-bit Morton codes (input order) and the 64-byte device records
 * (16 floats each, stored order).  Any pointer may be null. */
vp_status vp_scene_export(vp_scene* scene, uint32_t* order, uint32_t* morton, float* records16, uint32_t* depth);
/* Walk-tree form of a scene (no counterpart): half_records = 1 for the
 * binary16 records; box_growth = the summed leaf-box surface of the binary16
 * records over that of the binary32 ones, minus 1 (0 when the scene was
 * forced to VP_SCENE_WALK_F32).  Either pointer may be null. */
vp_status vp_scene_walk_info(vp_scene* scene, uint32_t* half_records, double* box_growth);
vp_status vp_env_create(vp_context* ctx, const float* rgba, uint32_t width, uint32_t height, vp_env** out);
vp_status vp_env_destroy(vp_env* env);
/* joint weights (W*H), alias thresholds (W*H), alias indices (W*H), total. */
vp_status vp_env_export(vp_env* env, float* joint, float* prob, uint32_t* alias, float* total);
vp_status vp_camera_perspective(uint32_t width, uint32_t height, const float eye[3], const float target[3],
                                const float up[3], float vfov_deg, vp_camera* out);
vp_status vp_camera_orthographic(uint32_t width, uint32_t height, const float eye[3], const float target[3],
                                 const float up[3], float ortho_height, vp_camera* out);
void vp_default_params(vp_render_params* p);

/* ---- plain renderer ---------------------------------------------------- */
/* Render to a host RGBA float image (width*height*4). */
vp_status vp_render(vp_context* ctx, vp_scene* scene, vp_env* env, const vp_camera* cam,
                    const vp_render_params* params, float* out_rgba, vp_render_stats* stats);
/* Same, output stays on the device: out_rgba_dev (W*H*4 floats), and, if not
 * null, out_samples_dev (spp*W*H*4 floats, sample-major) receives every
 * sample's radiance (the megakernel writes it; a wavefront launch
 * configuration falls back to the megakernel).  stream is a hipStream_t
 * (null = the context stream).  A pixel's mean is the exact 32.32 fixed-point
 * sum of its samples over their count (the reference keeps a running Welford
 * mean, raygen.cuh:214-238; the sum makes the result independent of the order
 * in which paths finish). */
vp_status vp_render_device(vp_context* ctx, vp_scene* scene, vp_env* env, const vp_camera* cam,
                           const vp_render_params* params, const vp_launch_config* launch,
                           float* out_rgba_dev, float* out_samples_dev, void* stream, vp_render_stats* stats);
/* vp_render_device without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * buffer beyond the caller's scene, environment and output image lives in a
 * file-scope device array of 128 MiB + 4 KiB that each device holds from the
 * library's first use until the process ends: the pixel accumulators (at most
 * 512x512 pixels; larger images are rejected with VP_ERR_INVALID), a one-entry
 * job table with its counters, and the megakernel's active-set slab, which
 * bounds the lanes of the launch (fewer lanes at larger active_cap; a
 * capacity that leaves no room for one workgroup is rejected).  The image is
 * bit-identical to vp_render_device's.  Calls on one device share the array
 * and race if they overlap; each call clears what must start at zero with
 * kernels (no hipMemsetAsync).  With stats null nothing synchronises the
 * stream; with stats it reads the counters back and synchronises. */
vp_status vp_render_device_noalloc(vp_context* ctx, vp_scene* scene, vp_env* env, const vp_camera* cam,
                                   const vp_render_params* params, float* out_rgba_dev, void* stream,
                                   vp_render_stats* stats);
/* out[i] = mean[i] * scale over n RGBA pixels, device pointers. */
vp_status vp_normalize_accumulator(float* out_rgba_dev, const float* mean_rgba_dev, size_t n, float scale,
                                   void* stream);

/* ---- batched persistent form ------------------------------------------- */
typedef struct {
    vp_scene* scene;
    vp_env* env;
    vp_camera camera;
    vp_render_params params;
    float* out_rgba; /* host, width*height*4 */
} vp_job;
vp_status vp_render_jobs(vp_context* ctx, vp_job* jobs, uint32_t njobs, const vp_launch_config* launch,
                         vp_render_stats* stats);
/* Search launch configurations on a job mix, cache and return the best. */
vp_status vp_autotune(vp_context* ctx, vp_job* jobs, uint32_t njobs, vp_launch_config* best);

/* ---- serving form ------------------------------------------------------ */
/* Continuous batching on the persistent megakernel: a server thread appends
 * each admitted request to its job table while earlier render launches are
 * still running and starts a launch over the longer table on a second
 * stream; launches share one path counter, so the new one takes the new
 * request's paths while the older one drains and the GPU stays full.
 * Finished paths add into the request's per-pixel accumulator (64 bytes per
 * pixel, whatever the spp); a request is resolved and copied out as soon as
 * its last path finishes.  Scenes and environments stay resident in their
 * handles.  The server owns its job table, path counter and slabs, so other
 * render calls on the same context stay possible (they share the GPU). */
typedef struct {
    uint64_t mem_budget_bytes; /* admission limit: accumulator bytes of the requests in flight (0 = 8 GiB) */
    uint32_t max_inflight;     /* requests on the GPU at once (0 = 64) */
    uint32_t active_cap;       /* largest active-set capacity a request may ask for (0 = the context's
                                  autotuned capacity, else 128); sizes the server's slabs */
    uint32_t reserve_cus;      /* 1: keep one CU of each XCD out of the render launches, so progress
                                  polls and result copies (blit kernels) run while a launch is
                                  resident and a request is returned as soon as its own paths finish;
                                  0 (default): every CU renders, and those copies wait for CUs that
                                  a launch frees (saturated throughput 11 % higher on the
                                  request-stream workload) */
} vp_server_config;
typedef struct {
    uint64_t completed;
    uint64_t samples;
    double p50_ms, p99_ms, mean_ms; /* submit-to-result latency */
    double busy_ms;                 /* time with requests in flight */
    uint64_t launches;              /* render launches started */
} vp_server_stats;
/* The launch shape is the context's autotuned megakernel one (else the
 * defaults); the larger of cfg->active_cap and the tuned capacity is the
 * largest a request may ask for.
 * A request above 1024x1024 pixels or above the memory budget is rejected
 * at submit. */
vp_status vp_server_create(vp_context* ctx, const vp_server_config* cfg, vp_server** out);
/* Queue a request; *ticket identifies it.  The output buffer must stay valid
 * until vp_server_poll reports completion. */
vp_status vp_server_submit(vp_server* srv, const vp_job* job, uint64_t* ticket);
/* VP_OK when done (latency_ms filled), VP_ERR_NOT_READY otherwise. */
vp_status vp_server_poll(vp_server* srv, uint64_t ticket, double* latency_ms);
vp_status vp_server_drain(vp_server* srv);
vp_status vp_server_stats_get(vp_server* srv, vp_server_stats* out);
vp_status vp_server_destroy(vp_server* srv);

/* ---- test entry points -------------------------------------------------- */
/* rays: m x (ox, oy, oz, dx, dy, dz), host.  T[m]: exp(-tau) with the origin
 * set = primitives whose 3-sigma bound contains the origin. */
vp_status vp_transmittance_rays(vp_context* ctx, vp_scene* scene, const float* rays, uint32_t m, float* T);
/* Entry hits of each ray sorted by (t, stored index): counts[m], t[m*cap],
 * idx[m*cap] (stored indices). */
vp_status vp_collect_overlaps(vp_context* ctx, vp_scene* scene, const float* rays, uint32_t m, uint32_t cap,
                              uint32_t* counts, float* t, uint32_t* idx);
/* Render once with the event-counting kernel variant and return the event
 * counters (n_events <= 18, order of the kernel's counter slots: active
 * overflow, max hits, max overlap, paths, traces, scatters, shadow rays,
 * argmin hits, argmin origin-set terms, argmin accepts, rebuild origin-set
 * tests, rebuild entry tests, albedo terms, shadow origin-set chords, shadow
 * entered chords, leaf entry tests, node visits, containment tests). */
vp_status vp_render_count_events(vp_context* ctx, vp_scene* scene, vp_env* env, const vp_camera* cam,
                                 const vp_render_params* params, uint64_t* events, uint32_t n_events);
/* Evaluate math function fn (0 exp 1 log 2 sin 3 cos 4 atan2 5 acos 6 erf
 * 7 erfc 8 erfinv 9 erf_diff 10 sqrt 11 div 12 chord_poly) on the device. */
vp_status vp_math_eval(vp_context* ctx, int fn, const float* x, const float* y, uint32_t n, float* out);
/* Evaluate primitive function fn (0 inv_cdf 1 inv_cdf_span 2 optical_depth
 * 3 exit_inside 4 entry_t 5 exit_entry 6 point_inside 7 pdf 8 chord_tau (the
 * optical depth of the whole entered chord) 9 entered_flight (free flight of
 * depth tau into the entered primitive; 1e20 when it passes) 10 inside_tau
 * (optical depth from an origin inside the primitive to its exit)
 * 11 origin_flight (free flight of depth tau from such an origin)) for one
 * record over m argument rows (o3 d3 t0 t1 tau). */
vp_status vp_prim_eval(vp_context* ctx, int fn, const float* record16, const float* args9, uint32_t m,
                       float* out);

#ifdef __cplusplus
}
#endif

#endif /* VP_API_H */
