"""Benchmark: Msamples/s of the Gaussian-mixture volumetric path tracer on one
MI355X, on the scattering workload (4096 anisotropic Gaussians, 512x512,
64 spp, max depth 64, Russian roulette from depth 12, albedo 0.9, isotropic
phase, 2048x1024 procedural sky environment with a 0.5 degree sun).

A step is one full render of that image (512*512*64 = 16.8 M camera paths)
with the scene, BVH and environment already resident in HBM.  Prints one
JSON line with the throughput, the roofline of the render kernel (FP32 VALU:
algorithmic FLOPs from the event-counting kernel variant weighted as
documented in DESIGN.md, over the kernel time measured with HIP events on
the library stream) and the CPU baseline (the oracle's C restatement of the
estimator on a bounded pixel sample, all host cores).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "cuda-volprim_amd")

# FP32 VALU peak of MI355X (published vector FP32, dual-issue/packed) and the
# rate without packed FP32 (which this build disables).
PEAK_FP32_TFLOPS = 157.3
PEAK_FP32_NONPACKED_TFLOPS = 78.6

# FLOP weight of each counted event: FMA = 2, add/mul = 1, each transcendental
# or rcp/sqrt = 1, erf = 16, erfinv = 20 (derivations in DESIGN.md).
FLOP_WEIGHTS = {
    "node_visits": 12, "leaf_tests": 105, "contain": 41, "arg_active": 230, "arg_hit": 141,
    "arg_accept": 122, "rb_active": 92, "rb_hit": 122, "sh_active": 238, "sh_hit": 268, "alb": 52,
    "scatters": 220, "paths": 35,
}


def load_pkg():
    if "volprim_amd" in sys.modules:
        return sys.modules["volprim_amd"]
    spec = importlib.util.spec_from_file_location("volprim_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["volprim_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(scenes):
    rows = scenes.c1_scattering(n=4096, seed=1, sigma_multiplier=10.0, albedo=0.9)
    env = scenes.sky_env(2048, 1024)
    cam_args = scenes.default_camera_args(512, 512, dist=3.5, fov=40.0)
    params = dict(width=512, height=512, spp=64, seed=0, max_depth=64, rr_depth=12, phase_g=0.0)
    return rows, env, cam_args, params


def cpu_baseline(rows, env, cam_pod, params, threads, budget_s=20.0):
    """The oracle (brute force over all primitives, pthreads over rows) on a
    strided pixel sample of the same image; Msamples/s is spp-invariant."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import vpo
    order = vpo.scene_order(rows)
    rec = vpo.make_prims(rows)[order]
    tabs = vpo.env_tables(env)
    stride, spp_used = 16, 1
    t0 = time.perf_counter()
    vpo.render(rec, env, cam_pod, params["width"], params["height"], params["spp"], seed=params["seed"],
               max_depth=params["max_depth"], rr_depth=params["rr_depth"], phase_g=params["phase_g"],
               stride=stride, samples=(0, spp_used), threads=threads, env_tables_=tabs)
    probe = time.perf_counter() - t0
    npx = ((params["width"] + stride - 1) // stride) * ((params["height"] + stride - 1) // stride)
    # grow the sample (more samples per pixel) to about budget_s of work
    spp_used = int(max(1, min(params["spp"], budget_s / max(probe, 1e-3))))
    t0 = time.perf_counter()
    vpo.render(rec, env, cam_pod, params["width"], params["height"], params["spp"], seed=params["seed"],
               max_depth=params["max_depth"], rr_depth=params["rr_depth"], phase_g=params["phase_g"],
               stride=stride, samples=(0, spp_used), threads=threads, env_tables_=tabs)
    dt = time.perf_counter() - t0
    n = npx * spp_used
    return {"value": n / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{npx} pixels (every {stride}th row and column of 512x512) x {spp_used} spp = {n} paths "
                      f"in {dt:.1f} s; brute force over all primitives"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup (0 = library default)")
    ap.add_argument("--bpc", type=int, default=0, help="persistent workgroups per CU (0 = occupancy limit)")
    ap.add_argument("--regen", type=int, default=0, help="idle lanes per wavefront before claiming new paths")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design is in DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    import torch
    pkg = load_pkg()
    from volprim_amd import scenes, volprim

    rows, env, cam_args, pp = workload(scenes)
    ctx = volprim.Context(0)
    dev_name, n_cu, clk = ctx.device_info()
    scene = ctx.scene(rows)
    envh = ctx.env(env)
    cam = volprim.camera_perspective(**cam_args)
    params = volprim.default_params(**pp)
    out = torch.empty((pp["height"], pp["width"], 4), dtype=torch.float32, device="cuda")
    launch = None
    if args.block or args.bpc or args.regen:
        launch = volprim.LaunchConfig(args.block, args.bpc, 0, 0, args.regen)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        ctx.render_device(scene, envh, cam, params, out.data_ptr(), launch=launch)
    torch.cuda.synchronize()
    kernel_ms, launches = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.render_device(scene, envh, cam, params, out.data_ptr(), launch=launch)
        kernel_ms += st["kernel_ms"]
        launches += st["launches"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    samples = pp["width"] * pp["height"] * pp["spp"]
    ms_per_step = dt / args.steps * 1e3
    value = samples * args.steps / dt / 1e6
    img = out.cpu().numpy()
    assert np.isfinite(img).all(), "non-finite pixels"

    roofline = None
    events = None
    if not args.no_count:
        events = ctx.count_events(scene, envh, cam, params)
        flops = sum(FLOP_WEIGHTS[k] * events[k] for k in FLOP_WEIGHTS)
        avg_launch_s = kernel_ms / max(launches, 1) / 1e3
        flops_per_launch = flops * args.steps / max(launches, 1)
        achieved = flops_per_launch / avg_launch_s / 1e12
        roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                    "peak_nonpacked": PEAK_FP32_NONPACKED_TFLOPS, "kernel": "k_render",
                    "avg_launch_ms": round(avg_launch_s * 1e3, 3), "flops_per_launch": flops_per_launch}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(rows, env, cam.pod(), pp, threads=min(os.cpu_count() or 1, 16))

    line = {
        "metric": "Msamples/s at <=1% RMSE vs reference image; p50/p99 request latency",
        "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "scattering: 4096 anisotropic Gaussians (seed 1), sigma multiplier 10, albedo 0.9, "
                               "isotropic phase, 2048x1024 sky + sun environment, 512x512, 64 spp, max depth 64, "
                               "RR from depth 12, MIS NEE", "width": 512, "height": 512, "spp": 64,
                   "primitives": 4096},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device": dev_name,
        "kernel_ms_per_step": round(kernel_ms / args.steps, 3),
        "events_per_step": events,
        "image_mean": float(img[..., 0:3].mean()),
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
