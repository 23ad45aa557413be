"""ctypes binding of the CPU oracle (oracle/libvpo.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvpo.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")

MATH_FNS = {"exp": 0, "log": 1, "sin": 2, "cos": 3, "atan2": 4, "acos": 5, "erf": 6,
            "erfc": 7, "erfinv": 8, "erf_diff": 9, "sqrt": 10, "div": 11, "chord_poly": 12}
PRIM_FNS = {"inv_cdf": 0, "inv_cdf_span": 1, "optical_depth": 2, "exit_inside": 3,
            "entry_t": 4, "exit_entry": 5, "point_inside": 6, "pdf": 7, "chord_tau": 8, "entered_flight": 9,
            "inside_tau": 10, "origin_flight": 11}
COUNTER_NAMES = ["n_sh_active", "n_sh_hit", "n_arg_active", "n_arg_hit", "n_arg_accept",
                 "n_rb_active", "n_rb_hit", "n_alb", "n_scatter", "n_sample", "n_shadow_rays",
                 "n_trace", "active_overflow", "max_hits", "max_overlap"]


class Params(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("spp", ctypes.c_uint32), ("seed", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("rr_depth", ctypes.c_uint32),
                ("rr_max_survival", ctypes.c_float), ("firefly_clamp", ctypes.c_float),
                ("filter_stddev", ctypes.c_float), ("phase_g", ctypes.c_float),
                ("use_ris", ctypes.c_uint32), ("ris_candidates", ctypes.c_uint32),
                ("active_cap", ctypes.c_uint32), ("measure", ctypes.c_uint32),
                ("x0", ctypes.c_uint32), ("y0", ctypes.c_uint32),
                ("x1", ctypes.c_uint32), ("y1", ctypes.c_uint32),
                ("sample_begin", ctypes.c_uint32), ("sample_end", ctypes.c_uint32),
                ("stride", ctypes.c_uint32), ("use_bvh", ctypes.c_uint32), ("shadow_rr", ctypes.c_float),
                ("adaptive_threshold", ctypes.c_float), ("adaptive_min", ctypes.c_uint32),
                ("adaptive_batch", ctypes.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.vpo_make_prims.argtypes = [_f32p, ctypes.c_long, _f32p]
    lib.vpo_prim_eval.argtypes = [ctypes.c_int, _f32p, _f32p, ctypes.c_long, _f32p]
    lib.vpo_math_eval.argtypes = [ctypes.c_int, _f32p, _f32p, ctypes.c_long, _f32p]
    lib.vpo_expand_bits.argtypes = [ctypes.c_uint32]
    lib.vpo_expand_bits.restype = ctypes.c_uint32
    lib.vpo_morton_codes.argtypes = [_f32p, ctypes.c_long, _u32p]
    lib.vpo_sort_order.argtypes = [_u32p, ctypes.c_long, _u32p]
    lib.vpo_pcg_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_long, _u32p]
    lib.vpo_pcg_raw.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_long, _u32p]
    lib.vpo_hash64.argtypes = [ctypes.c_uint64]
    lib.vpo_hash64.restype = ctypes.c_uint64
    lib.vpo_transmittance_rays.argtypes = [_f32p, ctypes.c_long, _f32p, ctypes.c_long, _f32p]
    lib.vpo_overlaps.argtypes = [_f32p, ctypes.c_long, _f32p, ctypes.c_long, ctypes.c_long,
                                 _u32p, _f32p, _u32p]
    lib.vpo_overlaps.restype = ctypes.c_long
    lib.vpo_env_tables.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _f32p, _f32p, _u32p,
                                   ctypes.POINTER(ctypes.c_float)]
    lib.vpo_env_ref_cdf.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f32p,
                                    ctypes.POINTER(ctypes.c_float)]
    lib.vpo_camera_perspective.argtypes = [ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f32p,
                                           ctypes.c_float, _f32p]
    lib.vpo_camera_orthographic.argtypes = [ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f32p,
                                            ctypes.c_float, _f32p]
    lib.vpo_render.argtypes = [_f32p, ctypes.c_long, _f32p, _f32p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                               _f32p, ctypes.POINTER(Params), ctypes.c_void_p, ctypes.c_void_p,
                               _u64p, ctypes.c_int, ctypes.c_void_p]
    lib.vpo_render.restype = ctypes.c_int
    lib.vpo_shadow_rr_T.argtypes = [ctypes.c_float, ctypes.c_float, _u64p, ctypes.c_long, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.c_double, _f32p, _u32p]
    lib.vpo_shadow_rr_T.restype = ctypes.c_uint64
    lib.vpo_phase_eval.argtypes = [ctypes.c_float, _f32p, _f32p, ctypes.c_long, _f32p]
    lib.vpo_phase_sample.argtypes = [ctypes.c_float, _f32p, ctypes.c_uint64, ctypes.c_long, _f32p, _f32p]
    lib.vpo_env_sample.argtypes = [_f32p, _f32p, _f32p, _u32p, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_uint64, ctypes.c_long, _f32p, _f32p]
    lib.vpo_env_pdf.argtypes = [_f32p, _f32p, _f32p, _u32p, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                _f32p, ctypes.c_long, _f32p]
    lib.vpo_env_lookup.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, _f32p, ctypes.c_long, _f32p]
    lib.vpo_camera_ray.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                   _f32p, _f32p]
    lib.vpo_prim_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    lib.vpo_first_scatter.argtypes = [_f32p, ctypes.c_long, _f32p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_long,
                                      _f32p]
    lib.vpo_prim_uniform.restype = ctypes.c_float
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def math_eval(name, x, y=None):
    x = f32(x)
    y = f32(np.zeros_like(x) if y is None else y)
    out = np.empty_like(x)
    lib().vpo_math_eval(MATH_FNS[name], x, y, x.size, out)
    return out


def make_prims(rows14):
    rows14 = f32(rows14).reshape(-1, 14)
    out = np.empty((rows14.shape[0], 16), np.float32)
    lib().vpo_make_prims(rows14, rows14.shape[0], out)
    return out


def prim_eval(name, prim16, args9):
    args9 = f32(args9).reshape(-1, 9)
    out = np.empty(args9.shape[0], np.float32)
    lib().vpo_prim_eval(PRIM_FNS[name], f32(prim16).reshape(16), args9, args9.shape[0], out)
    return out


def morton_codes(centers):
    centers = f32(centers).reshape(-1, 3)
    codes = np.empty(centers.shape[0], np.uint32)
    lib().vpo_morton_codes(centers, centers.shape[0], codes)
    return codes


def sort_order(codes):
    codes = np.ascontiguousarray(codes, np.uint32)
    order = np.empty(codes.size, np.uint32)
    lib().vpo_sort_order(codes, codes.size, order)
    return order


def scene_order(rows14):
    """Permutation the renderer applies: stable sort of Morton codes."""
    rows14 = f32(rows14).reshape(-1, 14)
    return sort_order(morton_codes(rows14[:, 0:3]))


def pcg_stream(seed, seq, n):
    out = np.empty(n, np.uint32)
    lib().vpo_pcg_stream(seed, seq, n, out)
    return out


def pcg_raw(state, inc, n):
    out = np.empty(n, np.uint32)
    lib().vpo_pcg_raw(state, inc, n, out)
    return out


def first_scatter(prims16, ray6, keys):
    """Bounce-0 argmin free-flight distance of independent path keys along one ray (inf = escape)."""
    prims16 = f32(prims16).reshape(-1, 16)
    keys = np.ascontiguousarray(keys, np.uint64)
    out = np.empty(keys.size, np.float32)
    lib().vpo_first_scatter(prims16.reshape(-1), prims16.shape[0], f32(ray6).reshape(6),
                            keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), keys.size, out)
    return out


def transmittance_rays(prims16, rays6):
    prims16 = f32(prims16).reshape(-1, 16)
    rays6 = f32(rays6).reshape(-1, 6)
    T = np.empty(rays6.shape[0], np.float32)
    lib().vpo_transmittance_rays(prims16, prims16.shape[0], rays6, rays6.shape[0], T)
    return T


def overlaps(prims16, rays6, cap):
    prims16 = f32(prims16).reshape(-1, 16)
    rays6 = f32(rays6).reshape(-1, 6)
    m = rays6.shape[0]
    counts = np.empty(m, np.uint32)
    t = np.zeros((m, cap), np.float32)
    idx = np.zeros((m, cap), np.uint32)
    lib().vpo_overlaps(prims16, prims16.shape[0], rays6, m, cap, counts, t, idx)
    return counts, t, idx


def env_tables(rgba):
    rgba = f32(rgba)
    h, w = rgba.shape[0], rgba.shape[1]
    joint = np.empty(w * h, np.float32)
    prob = np.empty(w * h, np.float32)
    alias = np.empty(w * h, np.uint32)
    total = ctypes.c_float(0.0)
    lib().vpo_env_tables(rgba.reshape(-1), w, h, joint, prob, alias, ctypes.byref(total))
    return joint, prob, alias, total.value


def env_ref_cdf(rgba):
    rgba = f32(rgba)
    h, w = rgba.shape[0], rgba.shape[1]
    marg = np.empty(h, np.float32)
    cond = np.empty(w * h, np.float32)
    joint = np.empty(w * h, np.float32)
    total = ctypes.c_float(0.0)
    lib().vpo_env_ref_cdf(rgba.reshape(-1), w, h, marg, cond, joint, ctypes.byref(total))
    return marg, cond.reshape(h, w), joint.reshape(h, w), total.value


def camera_perspective(w, h, eye, target, up, vfov):
    pod = np.empty(16, np.float32)
    lib().vpo_camera_perspective(w, h, f32(eye), f32(target), f32(up), vfov, pod)
    return pod


def camera_orthographic(w, h, eye, target, up, ortho_h):
    pod = np.empty(16, np.float32)
    lib().vpo_camera_orthographic(w, h, f32(eye), f32(target), f32(up), ortho_h, pod)
    return pod


def render(prims16, env_rgba, cam_pod, width, height, spp, seed=42, max_depth=128, rr_depth=12,
           rr_max_survival=0.99, firefly_clamp=0.0, filter_stddev=0.0, phase_g=0.85,
           use_ris=False, ris_candidates=6, active_cap=128, measure=False, crop=None,
           samples=None, keep_samples=False, threads=None, env_tables_=None, stride=1, bvh=False, shadow_rr=0.0,
           adaptive_threshold=0.0, adaptive_min_samples=32, adaptive_batch=16):
    """CPU restatement of the estimator.  Returns (mean[h,w,4], samples or None,
    counters dict).  crop = (x0, y0, x1, y1); samples = (begin, end).
    bvh=True answers the ray and point queries through a CPU BVH instead of
    loops over every primitive (identical images, faster at large N).
    adaptive_threshold > 0: adaptive sampling in batches of adaptive_batch;
    the mean is the Welford mean and the counters dict gains
    "samples_per_pixel" (uint32 [h, w])."""
    prims16 = f32(prims16).reshape(-1, 16)
    env_rgba = f32(env_rgba)
    eh, ew = env_rgba.shape[0], env_rgba.shape[1]
    if env_tables_ is None:
        env_tables_ = env_tables(env_rgba)
    joint, prob, alias, total = env_tables_
    x0, y0, x1, y1 = crop if crop is not None else (0, 0, width, height)
    s0, s1 = samples if samples is not None else (0, spp)
    p = Params(width, height, spp, seed, max_depth, rr_depth, rr_max_survival, firefly_clamp,
               filter_stddev, phase_g, int(use_ris), ris_candidates, active_cap, int(measure),
               x0, y0, x1, y1, s0, s1, stride, int(bvh), float(shadow_rr), float(adaptive_threshold),
               int(adaptive_min_samples), int(adaptive_batch))
    mean = np.zeros((y1 - y0, x1 - x0, 4), np.float32)
    smp = np.zeros((s1 - s0, y1 - y0, x1 - x0, 4), np.float32) if keep_samples else None
    counters = np.zeros(len(COUNTER_NAMES), np.uint64)
    counts = np.zeros((y1 - y0, x1 - x0), np.uint32) if adaptive_threshold > 0 else None
    threads = threads or os.cpu_count() or 1
    rc = lib().vpo_render(prims16, prims16.shape[0], env_rgba.reshape(-1), joint,
                          prob.ctypes.data, alias.ctypes.data, total, ew, eh, f32(cam_pod),
                          ctypes.byref(p), mean.ctypes.data,
                          smp.ctypes.data if smp is not None else None, counters, threads,
                          counts.ctypes.data if counts is not None else None)
    if rc != 0:
        raise ValueError("vpo_render rejected its arguments")
    k = dict(zip(COUNTER_NAMES, counters.tolist()))
    if counts is not None:
        k["samples_per_pixel"] = counts
    return mean, smp, k


def shadow_rr_T(w, eps, keys, tau, bounce=0, ray=0):
    """The shadow-ray roulette's estimates of T = e^-tau for one ray of
    weight w per key, walked in full: (T[m], first failing level[m], tau_c)."""
    keys = np.ascontiguousarray(keys, np.uint64)
    T = np.empty(keys.size, np.float32)
    fail = np.empty(keys.size, np.uint32)
    thr = lib().vpo_shadow_rr_T(w, eps, keys, keys.size, bounce, ray, tau, T, fail)
    return T, fail, thr / 2.0 ** 40


def phase_eval(g, wi, wo):
    wi, wo = f32(wi).reshape(-1, 3), f32(wo).reshape(-1, 3)
    out = np.empty(wi.shape[0], np.float32)
    lib().vpo_phase_eval(g, wi, wo, wi.shape[0], out)
    return out


def phase_sample(g, wi, seed=1):
    wi = f32(wi).reshape(-1, 3)
    wo = np.empty_like(wi)
    pdf = np.empty(wi.shape[0], np.float32)
    lib().vpo_phase_sample(g, wi, seed, wi.shape[0], wo, pdf)
    return wo, pdf


def env_sample(rgba, tables, n, seed=1):
    rgba = f32(rgba)
    joint, prob, alias, total = tables
    wo = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    lib().vpo_env_sample(rgba.reshape(-1), joint, prob, alias, total, rgba.shape[1], rgba.shape[0], seed, n, wo, pdf)
    return wo, pdf


def env_pdf(rgba, tables, wo):
    rgba = f32(rgba)
    joint, prob, alias, total = tables
    wo = f32(wo).reshape(-1, 3)
    pdf = np.empty(wo.shape[0], np.float32)
    lib().vpo_env_pdf(rgba.reshape(-1), joint, prob, alias, total, rgba.shape[1], rgba.shape[0], wo, wo.shape[0], pdf)
    return pdf


def env_lookup(rgba, dirs):
    rgba = f32(rgba)
    dirs = f32(dirs).reshape(-1, 3)
    out = np.empty_like(dirs)
    lib().vpo_env_lookup(rgba.reshape(-1), rgba.shape[1], rgba.shape[0], dirs, dirs.shape[0], out)
    return out


def camera_ray(pod, px, py, jx, jy):
    o = np.empty(3, np.float32)
    d = np.empty(3, np.float32)
    lib().vpo_camera_ray(f32(pod), px, py, jx, jy, o, d)
    return o, d
