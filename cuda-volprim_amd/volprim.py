"""ctypes binding of libvolprim.so (C ABI in include/vp_api.h).

The library is built in-tree by `make -C cuda-volprim_amd` (or
__graft_entry__.build()).  Loading fails loudly when it is missing: there is
no CPU fallback for the render path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VP_LIBRARY selects another build of the same library (used for sweeps).
LIB_PATH = os.environ.get("VP_LIBRARY", os.path.join(HERE, "libvolprim.so"))

VP_OK = 0
_STATUS = {0: "VP_OK", 1: "VP_ERR_INVALID", 2: "VP_ERR_HIP", 3: "VP_ERR_OOM", 4: "VP_ERR_NOT_READY",
           5: "VP_ERR_STATE"}
MATH_FNS = {"exp": 0, "log": 1, "sin": 2, "cos": 3, "atan2": 4, "acos": 5, "erf": 6, "erfc": 7,
            "erfinv": 8, "erf_diff": 9, "sqrt": 10, "div": 11, "chord_poly": 12}
PRIM_FNS = {"inv_cdf": 0, "inv_cdf_span": 1, "optical_depth": 2, "exit_inside": 3, "entry_t": 4,
            "exit_entry": 5, "point_inside": 6, "pdf": 7, "chord_tau": 8, "entered_flight": 9,
            "inside_tau": 10, "origin_flight": 11}


class VpError(RuntimeError):
    pass


class Primitive(ctypes.Structure):
    _fields_ = [("center", ctypes.c_float * 3), ("quat_wxyz", ctypes.c_float * 4),
                ("scale", ctypes.c_float * 3), ("albedo", ctypes.c_float * 3), ("mass", ctypes.c_float)]


class Camera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("pixel00_rel", ctypes.c_float * 3),
                ("pixel_du", ctypes.c_float * 3), ("pixel_dv", ctypes.c_float * 3),
                ("view_dir", ctypes.c_float * 3), ("orthographic", ctypes.c_float)]

    def pod(self):
        """16 floats: eye, p00, du, dv, view_dir, orthographic (the oracle's layout)."""
        return np.array(list(self.eye) + list(self.pixel00_rel) + list(self.pixel_du) +
                        list(self.pixel_dv) + list(self.view_dir) + [self.orthographic], np.float32)


class RenderParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("max_depth", ctypes.c_uint32), ("rr_depth", ctypes.c_uint32),
                ("rr_max_survival", ctypes.c_float), ("firefly_clamp", ctypes.c_float),
                ("filter_stddev", ctypes.c_float), ("phase_g", ctypes.c_float),
                ("use_ris", ctypes.c_uint32), ("ris_candidates", ctypes.c_uint32),
                ("measure_caps", ctypes.c_uint32), ("active_cap", ctypes.c_uint32), ("shadow_rr", ctypes.c_float),
                ("adaptive_threshold", ctypes.c_float), ("adaptive_min_samples", ctypes.c_uint32)]


class RenderStats(ctypes.Structure):
    _fields_ = [("active_overflow", ctypes.c_uint64), ("hit_overflow", ctypes.c_uint64),
                ("max_hits_per_ray", ctypes.c_uint32), ("max_point_overlap", ctypes.c_uint32),
                ("kernel_ms", ctypes.c_double), ("launches", ctypes.c_uint32), ("wall_ms", ctypes.c_double),
                ("paths", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LaunchConfig(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("active_cap", ctypes.c_uint32), ("batch_spp", ctypes.c_uint32), ("regen_min", ctypes.c_uint32),
                ("pipeline", ctypes.c_uint32), ("pool_paths", ctypes.c_uint32)]


PIPELINE_MEGAKERNEL = 0
PIPELINE_WAVEFRONT = 1


class Job(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("env", ctypes.c_void_p), ("camera", Camera),
                ("params", RenderParams), ("out_rgba", ctypes.c_void_p)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("mem_budget_bytes", ctypes.c_uint64), ("max_inflight", ctypes.c_uint32),
                ("active_cap", ctypes.c_uint32), ("reserve_cus", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("samples", ctypes.c_uint64), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double), ("busy_ms", ctypes.c_double),
                ("launches", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every symbol include/vp_api.h declares (checked by the CPU test suite).
EXPORTED = [
    "vp_context_create", "vp_context_destroy", "vp_last_error", "vp_device_info", "vp_scene_create",
    "vp_scene_destroy", "vp_scene_export", "vp_env_create", "vp_env_destroy", "vp_env_export",
    "vp_camera_perspective", "vp_camera_orthographic", "vp_default_params", "vp_render",
    "vp_render_device", "vp_normalize_accumulator", "vp_render_jobs", "vp_autotune", "vp_server_create",
    "vp_server_submit", "vp_server_poll", "vp_server_drain", "vp_server_stats_get", "vp_server_destroy",
    "vp_transmittance_rays", "vp_collect_overlaps", "vp_render_count_events", "vp_math_eval", "vp_prim_eval",
    "vp_scene_create_ex", "vp_render_device_noalloc", "vp_scene_walk_info",
]
SCENE_BVH_SAH = 1
SCENE_BVH_MEDIAN = 2
SCENE_WALK_F32 = 4
SCENE_WALK_F16 = 8
EVENT_NAMES = ["active_overflow", "max_hits", "max_overlap", "paths", "traces", "scatters", "shadow_rays",
               "arg_hit", "arg_active", "arg_accept", "rb_active", "rb_hit", "alb", "sh_active", "sh_hit",
               "leaf_tests", "node_visits", "contain", "visits_shadow", "visits_point"]

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_LIB = None


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise VpError(f"{LIB_PATH} is not built: run `make -C {HERE}` (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    lib.vp_last_error.restype = ctypes.c_char_p
    lib.vp_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    lib.vp_context_destroy.argtypes = [_vp]
    lib.vp_device_info.argtypes = [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32),
                                   ctypes.POINTER(ctypes.c_double)]
    lib.vp_scene_create.argtypes = [_vp, ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(_vp)]
    lib.vp_scene_create_ex.argtypes = [_vp, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_vp)]
    lib.vp_scene_destroy.argtypes = [_vp]
    lib.vp_scene_export.argtypes = [_vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_uint32)]
    lib.vp_scene_walk_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)]
    lib.vp_env_create.argtypes = [_vp, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_vp)]
    lib.vp_env_destroy.argtypes = [_vp]
    lib.vp_env_export.argtypes = [_vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_float)]
    lib.vp_camera_perspective.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p, _f32p, ctypes.c_float,
                                          ctypes.POINTER(Camera)]
    lib.vp_camera_orthographic.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p, _f32p, ctypes.c_float,
                                           ctypes.POINTER(Camera)]
    lib.vp_default_params.argtypes = [ctypes.POINTER(RenderParams)]
    lib.vp_default_params.restype = None
    lib.vp_render.argtypes = [_vp, _vp, _vp, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams), _f32p,
                              ctypes.POINTER(RenderStats)]
    lib.vp_render_device.argtypes = [_vp, _vp, _vp, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams),
                                     ctypes.POINTER(LaunchConfig), _vp, _vp, _vp, ctypes.POINTER(RenderStats)]
    lib.vp_normalize_accumulator.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_float, _vp]
    lib.vp_render_device_noalloc.argtypes = [_vp, _vp, _vp, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams),
                                             _vp, _vp, ctypes.POINTER(RenderStats)]
    lib.vp_render_jobs.argtypes = [_vp, ctypes.POINTER(Job), ctypes.c_uint32, ctypes.POINTER(LaunchConfig),
                                   ctypes.POINTER(RenderStats)]
    lib.vp_autotune.argtypes = [_vp, ctypes.POINTER(Job), ctypes.c_uint32, ctypes.POINTER(LaunchConfig)]
    lib.vp_server_create.argtypes = [_vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
    lib.vp_server_submit.argtypes = [_vp, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_uint64)]
    lib.vp_server_poll.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    lib.vp_server_drain.argtypes = [_vp]
    lib.vp_server_stats_get.argtypes = [_vp, ctypes.POINTER(ServerStats)]
    lib.vp_server_destroy.argtypes = [_vp]
    lib.vp_transmittance_rays.argtypes = [_vp, _vp, _f32p, ctypes.c_uint32, _f32p]
    lib.vp_collect_overlaps.argtypes = [_vp, _vp, _f32p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _f32p, _u32p]
    lib.vp_render_count_events.argtypes = [_vp, _vp, _vp, ctypes.POINTER(Camera), ctypes.POINTER(RenderParams),
                                           np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS"), ctypes.c_uint32]
    lib.vp_math_eval.argtypes = [_vp, ctypes.c_int, _f32p, _f32p, ctypes.c_uint32, _f32p]
    lib.vp_prim_eval.argtypes = [_vp, ctypes.c_int, _f32p, _f32p, ctypes.c_uint32, _f32p]
    _LIB = lib
    return lib


def _check(rc, what):
    if rc != VP_OK:
        msg = load().vp_last_error().decode(errors="replace")
        raise VpError(f"{what} failed with {_STATUS.get(rc, rc)}: {msg}")


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def primitives_from_rows(rows14):
    """rows: N x 14 = centre(3) quat wxyz(4) scale(3) albedo(3) mass -> ctypes array."""
    rows14 = f32(rows14).reshape(-1, 14)
    arr = (Primitive * rows14.shape[0])()
    ctypes.memmove(arr, rows14.ctypes.data, rows14.nbytes)
    return arr


def default_params(**kw):
    p = RenderParams()
    load().vp_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def camera_perspective(w, h, eye, target, up, vfov):
    c = Camera()
    _check(load().vp_camera_perspective(w, h, f32(eye), f32(target), f32(up), vfov, ctypes.byref(c)),
           "vp_camera_perspective")
    return c


def camera_orthographic(w, h, eye, target, up, ortho_h):
    c = Camera()
    _check(load().vp_camera_orthographic(w, h, f32(eye), f32(target), f32(up), ortho_h, ctypes.byref(c)),
           "vp_camera_orthographic")
    return c


class Scene:
    def __init__(self, ctx, rows14, bvh="lbvh", walk="auto"):
        self.ctx = ctx
        self.rows = f32(rows14).reshape(-1, 14)
        self.n = self.rows.shape[0]
        prims = primitives_from_rows(self.rows)
        h = _vp()
        flags = {"lbvh": 0, "sah": SCENE_BVH_SAH, "median": SCENE_BVH_MEDIAN}[bvh]
        flags |= {"auto": 0, "f32": SCENE_WALK_F32, "f16": SCENE_WALK_F16}[walk]
        _check(load().vp_scene_create_ex(ctx.handle, ctypes.cast(prims, _vp), self.n, flags, ctypes.byref(h)),
               "vp_scene_create_ex")
        self.handle = h

    def export(self):
        order = np.empty(self.n, np.uint32)
        morton = np.empty(self.n, np.uint32)
        rec = np.empty((self.n, 16), np.float32)
        depth = ctypes.c_uint32(0)
        _check(load().vp_scene_export(self.handle, order.ctypes.data, morton.ctypes.data, rec.ctypes.data,
                                      ctypes.byref(depth)), "vp_scene_export")
        return order, morton, rec, depth.value

    def walk_info(self):
        """(half_records, box_growth): whether the walk tree holds binary16
        records, and the growth of the summed leaf-box surface they bring."""
        half = ctypes.c_uint32(0)
        growth = ctypes.c_double(0.0)
        _check(load().vp_scene_walk_info(self.handle, ctypes.byref(half), ctypes.byref(growth)),
               "vp_scene_walk_info")
        return bool(half.value), growth.value

    def close(self):
        if self.handle:
            load().vp_scene_destroy(self.handle)
            self.handle = None


class Env:
    def __init__(self, ctx, rgba):
        rgba = f32(rgba)
        if rgba.ndim != 3 or rgba.shape[2] != 4:
            raise ValueError("env must be H x W x 4")
        self.h, self.w = rgba.shape[0], rgba.shape[1]
        self.rgba = rgba
        h = _vp()
        _check(load().vp_env_create(ctx.handle, rgba.reshape(-1), self.w, self.h, ctypes.byref(h)), "vp_env_create")
        self.handle = h

    def export(self):
        n = self.w * self.h
        joint = np.empty(n, np.float32)
        prob = np.empty(n, np.float32)
        alias = np.empty(n, np.uint32)
        total = ctypes.c_float(0)
        _check(load().vp_env_export(self.handle, joint.ctypes.data, prob.ctypes.data, alias.ctypes.data,
                                    ctypes.byref(total)), "vp_env_export")
        return joint, prob, alias, total.value

    def close(self):
        if self.handle:
            load().vp_env_destroy(self.handle)
            self.handle = None


class Context:
    def __init__(self, device=0):
        h = _vp()
        _check(load().vp_context_create(device, ctypes.byref(h)), "vp_context_create")
        self.handle = h

    def close(self):
        if self.handle:
            load().vp_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_info(self):
        buf = ctypes.create_string_buffer(256)
        cus = ctypes.c_uint32(0)
        mhz = ctypes.c_double(0)
        _check(load().vp_device_info(self.handle, buf, 256, ctypes.byref(cus), ctypes.byref(mhz)), "vp_device_info")
        return buf.value.decode(), cus.value, mhz.value

    def scene(self, rows14, bvh="lbvh", walk="auto"):
        """bvh: "lbvh" (GPU build, default), "sah" (host SAH tree, same layout) or
        "median" (host tree of median splits on the walk tree's record boundaries).
        walk: "auto" (binary16 walk-tree records when the leaf boxes grow by at
        most 1/16), "f16" or "f32"; renders are bit-identical for all three."""
        return Scene(self, rows14, bvh, walk)

    def env(self, rgba):
        return Env(self, rgba)

    def render(self, scene, env, cam, params):
        out = np.empty((params.height, params.width, 4), np.float32)
        st = RenderStats()
        _check(load().vp_render(self.handle, scene.handle, env.handle, ctypes.byref(cam), ctypes.byref(params),
                                out.reshape(-1), ctypes.byref(st)), "vp_render")
        return out, st.as_dict()

    def render_device(self, scene, env, cam, params, out_ptr, samples_ptr=None, stream=None, launch=None):
        st = RenderStats()
        _check(load().vp_render_device(self.handle, scene.handle, env.handle, ctypes.byref(cam), ctypes.byref(params),
                                       ctypes.byref(launch) if launch is not None else None, out_ptr, samples_ptr,
                                       stream, ctypes.byref(st)), "vp_render_device")
        return st.as_dict()

    def render_device_noalloc(self, scene, env, cam, params, out_ptr, stream=None, stats=True):
        """vp_render_device_noalloc: no allocation, scratch in a static device array."""
        st = RenderStats()
        _check(load().vp_render_device_noalloc(self.handle, scene.handle, env.handle, ctypes.byref(cam),
                                               ctypes.byref(params), out_ptr, stream,
                                               ctypes.byref(st) if stats else None), "vp_render_device_noalloc")
        return st.as_dict() if stats else None

    def render_jobs(self, jobs, launch=None):
        """jobs: list of (scene, env, cam, params).  Returns (images, stats)."""
        arr = (Job * len(jobs))()
        outs = []
        for i, (sc, env, cam, p) in enumerate(jobs):
            o = np.empty((p.height, p.width, 4), np.float32)
            outs.append(o)
            arr[i].scene = sc.handle
            arr[i].env = env.handle
            arr[i].camera = cam
            arr[i].params = p
            arr[i].out_rgba = o.ctypes.data
        st = RenderStats()
        _check(load().vp_render_jobs(self.handle, arr, len(jobs), ctypes.byref(launch) if launch else None,
                                     ctypes.byref(st)), "vp_render_jobs")
        return outs, st.as_dict()

    def normalize_accumulator(self, out_ptr, mean_ptr, n, scale, stream=None):
        """out[i] = mean[i] * scale over n RGBA pixels (device pointers)."""
        _check(load().vp_normalize_accumulator(out_ptr, mean_ptr, n, scale, stream), "vp_normalize_accumulator")

    def count_events(self, scene, env, cam, params):
        ev = np.zeros(len(EVENT_NAMES), np.uint64)
        _check(load().vp_render_count_events(self.handle, scene.handle, env.handle, ctypes.byref(cam),
                                             ctypes.byref(params), ev, len(EVENT_NAMES)), "vp_render_count_events")
        return dict(zip(EVENT_NAMES, [int(v) for v in ev]))

    def autotune(self, jobs):
        """jobs: list of (scene, env, cam, params).  Returns the chosen LaunchConfig."""
        arr = (Job * len(jobs))()
        for i, (sc, env, cam, p) in enumerate(jobs):
            arr[i].scene, arr[i].env, arr[i].camera, arr[i].params = sc.handle, env.handle, cam, p
        best = LaunchConfig()
        _check(load().vp_autotune(self.handle, arr, len(jobs), ctypes.byref(best)), "vp_autotune")
        return best

    def server(self, mem_budget_bytes=8 << 30, max_inflight=64, active_cap=0, reserve_cus=False):
        return Server(self, mem_budget_bytes, max_inflight, active_cap, reserve_cus)

    def transmittance_rays(self, scene, rays6):
        rays6 = f32(rays6).reshape(-1, 6)
        T = np.empty(rays6.shape[0], np.float32)
        _check(load().vp_transmittance_rays(self.handle, scene.handle, rays6.reshape(-1), rays6.shape[0], T),
               "vp_transmittance_rays")
        return T

    def collect_overlaps(self, scene, rays6, cap):
        rays6 = f32(rays6).reshape(-1, 6)
        m = rays6.shape[0]
        counts = np.empty(m, np.uint32)
        t = np.empty((m, cap), np.float32)
        idx = np.empty((m, cap), np.uint32)
        _check(load().vp_collect_overlaps(self.handle, scene.handle, rays6.reshape(-1), m, cap, counts,
                                          t.reshape(-1), idx.reshape(-1)), "vp_collect_overlaps")
        return counts, t, idx

    def math_eval(self, name, x, y=None):
        x = f32(x)
        y = f32(np.zeros_like(x) if y is None else y)
        out = np.empty_like(x)
        _check(load().vp_math_eval(self.handle, MATH_FNS[name], x.reshape(-1), y.reshape(-1), x.size,
                                   out.reshape(-1)), "vp_math_eval")
        return out

    def prim_eval(self, name, rec16, args9):
        args9 = f32(args9).reshape(-1, 9)
        out = np.empty(args9.shape[0], np.float32)
        _check(load().vp_prim_eval(self.handle, PRIM_FNS[name], f32(rec16).reshape(16), args9.reshape(-1),
                                   args9.shape[0], out), "vp_prim_eval")
        return out


class Server:
    """Continuous-batching request loop (vp_server_*).  submit() returns a
    ticket and the host image the result lands in; poll() reports completion."""

    def __init__(self, ctx, mem_budget_bytes, max_inflight, active_cap=0, reserve_cus=False):
        self.ctx = ctx
        cfg = ServerConfig(mem_budget_bytes, max_inflight, active_cap, 1 if reserve_cus else 0)
        h = _vp()
        _check(load().vp_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)), "vp_server_create")
        self.handle = h
        self._keep = {}

    def submit(self, scene, env, cam, params):
        out = np.empty((params.height, params.width, 4), np.float32)
        job = Job()
        job.scene, job.env, job.camera, job.params, job.out_rgba = scene.handle, env.handle, cam, params, \
            out.ctypes.data
        t = ctypes.c_uint64(0)
        _check(load().vp_server_submit(self.handle, ctypes.byref(job), ctypes.byref(t)), "vp_server_submit")
        self._keep[t.value] = out
        return t.value, out

    def poll(self, ticket):
        """None while running, else the latency in ms."""
        lat = ctypes.c_double(0)
        rc = load().vp_server_poll(self.handle, ticket, ctypes.byref(lat))
        if rc == 4:
            return None
        _check(rc, "vp_server_poll")
        self._keep.pop(ticket, None)
        return lat.value

    def drain(self):
        _check(load().vp_server_drain(self.handle), "vp_server_drain")

    def stats(self):
        st = ServerStats()
        _check(load().vp_server_stats_get(self.handle, ctypes.byref(st)), "vp_server_stats_get")
        return st.as_dict()

    def close(self):
        if self.handle:
            load().vp_server_destroy(self.handle)
            self.handle = None
