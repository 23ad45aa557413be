"""Benchmark of the Gaussian-mixture volumetric path tracer on one MI355X.

Default: the served request stream.  Render requests drawn from the batched
job mix (resolution uniform over {64,...,1024}^2, spp round(exp U[ln 4,
ln 256]) divided by --spp-div, scene uniform over the 4096-Gaussian
scattering scene, the 10^6-Gaussian cloud and the 2.5*10^5-Gaussian torus,
camera on a sphere of radius 3.5) go to the server all at once; one step is
one request (--steps requests are timed, --warmup smaller ones go first).
value = the served aggregate Msamples/s from the first submission to the
last result, with scenes, BVHs and environment tables resident.  The same
requests then arrive as a Poisson stream at --load times that capacity for
the p50/p99 latencies.  Requests use the unbiased shadow-ray
roulette at eps = 1e-3 (--shadow-rr; 0 is the reference's estimator, which
is also served once for comparison).  Beside it: plain one-launch-per-request
rendering and one batched launch of the same requests (all three images are
checked bit-identical), p50/p99 latency, the roofline of the batched
k_render launch (FP32 VALU: algorithmic FLOPs from the event-counting kernel
variant with per-event weights derived from the reference's formulas, over the launch's kernel time
from HIP events on the library stream), a per-scene image-error estimate and
the CPU baseline (the oracle's C restatement of the estimator with its CPU
BVH, on a bounded pixel sample, all host cores).

--single {scattering,cloud,torus}: one scene rendered whole per step
(512x512x64 for scattering, 1024x1024x4 for cloud and torus), for profiling.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--spp-div D] [--no-cpu]
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
# or rcp/sqrt = 1, erf = 16, erfinv = 20 (from the reference's formulas).  A node
# visit loads one 4-wide walk-tree record and slab-tests its 4 child boxes
# (6 FMA each).
FLOP_WEIGHTS = {
    "node_visits": 48, "leaf_tests": 105, "contain": 41, "arg_active": 230, "arg_hit": 141,
    "arg_accept": 0, "rb_active": 1, "rb_hit": 122, "sh_active": 238, "sh_hit": 268, "alb": 52,
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


# FLOP weights that differ from the per-event derivations of the reference
# formulas, and why (the counted events are this build's own work):
WEIGHTS_NOTE = ("arg_accept 0 (not 122) and rb_active 1 (not 92): the exit distance a candidate or an origin-set "
                "member is compared with is computed once during the argmin walk and charged there, so the accept "
                "and the rebuild's membership test are a compare each")


def workload(scenes, name):
    """--single workloads: rows, camera arguments, render parameters, description."""
    if name == "scattering":
        rows = scenes.c1_scattering(n=4096, seed=1, sigma_multiplier=10.0, albedo=0.9)
        return (rows, scenes.default_camera_args(512, 512, dist=3.5, fov=40.0),
                dict(width=512, height=512, spp=64, seed=0, max_depth=64, rr_depth=12, phase_g=0.0, active_cap=128),
                "scattering: 4096 anisotropic Gaussians (seed 1), sigma multiplier 10, albedo 0.9, isotropic "
                "phase, 2048x1024 sky + sun environment, 512x512, 64 spp, max depth 64, RR from depth 12")
    if name == "cloud":
        rows = scenes.c2_cloud(n=1_000_000, seed=2, sigma_multiplier=7.5, albedo=0.99)
        return (rows, scenes.default_camera_args(1024, 1024, dist=3.5, fov=40.0),
                dict(width=1024, height=1024, spp=4, seed=0, max_depth=64, rr_depth=12, phase_g=0.85, active_cap=384),
                "cloud: 10^6 Gaussians (fBm noise, seed 2), sigma multiplier 7.5, albedo 0.99, HG g 0.85, "
                "2048x1024 sky + sun environment, 1024x1024, 4 spp, max depth 64, RR from depth 12")
    rows = scenes.torus_shell(n=250_000, seed=3, sigma_multiplier=10.0, albedo=0.95)
    return (rows, scenes.default_camera_args(1024, 1024, dist=3.5, fov=40.0),
            dict(width=1024, height=1024, spp=4, seed=0, max_depth=64, rr_depth=12, phase_g=0.0, active_cap=384),
            "torus: 2.5*10^5 Gaussians on a torus shell (seed 3), sigma multiplier 10, albedo 0.95, isotropic "
            "phase, 2048x1024 sky + sun environment, 1024x1024, 4 spp, max depth 64, RR from depth 12")


def cpu_baseline(rows, env, cam_pod, params, threads, budget_s=20.0, repeats=5):
    """The oracle (with its CPU BVH for ray and point queries, pthreads over
    rows) on a strided pixel sample of the same image; Msamples/s is
    spp-invariant.  One warm-up render sizes the sample, then `repeats` timed
    renders (median and spread reported)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import vpo
    order = vpo.scene_order(rows)
    rec = vpo.make_prims(rows)[order]
    tabs = vpo.env_tables(env)
    stride = 16

    def run(spp_used):
        t0 = time.perf_counter()
        vpo.render(rec, env, cam_pod, params["width"], params["height"], params["spp"], seed=params["seed"],
                   max_depth=params["max_depth"], rr_depth=params["rr_depth"], phase_g=params["phase_g"],
                   use_ris=bool(params.get("use_ris", 0)), stride=stride, samples=(0, spp_used), threads=threads,
                   env_tables_=tabs, bvh=True, active_cap=params.get("active_cap", 128),
                   shadow_rr=params.get("shadow_rr", 0.0))
        return time.perf_counter() - t0

    probe = run(1)
    npx = ((params["width"] + stride - 1) // stride) * ((params["height"] + stride - 1) // stride)
    # samples per pixel so that each timed repeat is about budget_s / repeats
    spp_used = int(max(1, min(params["spp"], budget_s / repeats / max(probe, 1e-3))))
    times = [run(spp_used) for _ in range(repeats)]
    n = npx * spp_used
    rates = sorted(n / t / 1e6 for t in times)
    return {"value": float(np.median(rates)), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "spread": [rates[0], rates[-1]],
            "sample": f"{npx} pixels (every {stride}th row and column of {params['width']}x{params['height']}) x "
                      f"{spp_used} spp = {n} paths per "
                      f"repeat, 1 warm-up + {repeats} timed repeats (median); CPU BVH (binary, Morton-ordered) queries"}


def measure_traffic(args):
    """HBM bytes of one k_render launch from two rocprofv3 --pmc runs of this
    script (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by
    what the same counter reads for an elementwise kernel that reads and
    writes 256 MiB.  Returns (dict, note); dict is None when a run fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    child = [sys.executable, os.path.abspath(__file__), "--single", args.single, "--steps", "1", "--warmup", "0",
             "--no-cpu", "--no-count", "--no-quality", "--no-traffic", "--copy-calibration", "--bvh", args.bvh]
    if args.block or args.bpc or args.regen:
        child += ["--block", str(args.block), "--bpc", str(args.bpc), "--regen", str(args.regen)]
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="vp_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
        try:
            subprocess.run([prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--"] + child,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=180)
        except subprocess.TimeoutExpired:
            return None, f"rocprofv3 --pmc {counter} timed out"
        path = os.path.join(d, "run_counter_collection.csv")
        if not os.path.exists(path):
            return None, f"rocprofv3 --pmc {counter} wrote no counters"
        per = {}
        names = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            per[r["Dispatch_Id"]] = per.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
            names[r["Dispatch_Id"]] = r["Kernel_Name"]
        shutil.rmtree(d, ignore_errors=True)
        render = [v for k, v in per.items() if "k_render" in names[k]]
        calib = [v for k, v in per.items() if "elementwise" in names[k]]
        if not render or not calib:
            return None, f"{counter}: kernels not found in the counter dump"
        scale = (256 * 1024) / calib[-1]
        out[counter] = (render[-1], calib[-1], scale)
    f, w = out["FETCH_SIZE"], out["WRITE_SIZE"]
    total = (f[0] * f[2] + w[0] * w[2]) * 1024.0
    return {"bytes": total, "fetch_kib": f[0], "write_kib": w[0], "fetch_scale": f[2], "write_scale": w[2]}, None


MIX_ACTIVE_CAP = {"scattering": 128, "cloud": 384, "torus": 384}
MIX_SCENES = {"scattering": "4096 anisotropic Gaussians (seed 1, sigma multiplier 10, albedo 0.9, isotropic phase)",
              "cloud": "10^6 Gaussians from 5-octave fBm noise in a 2x1x1 box (seed 2, sigma multiplier 7.5, "
                       "albedo 0.99, HG g 0.85)",
              "torus": "2.5*10^5 Gaussians on a torus shell R 1 r 0.3 (seed 3, sigma multiplier 10, albedo 0.95)"}


def mix_scenes(ctx, scenes, bvh, names, walk="auto"):
    gens = {"scattering": lambda: scenes.c1_scattering(n=4096, seed=1, sigma_multiplier=10.0, albedo=0.9),
            "cloud": lambda: scenes.c2_cloud(n=1_000_000, seed=2, sigma_multiplier=7.5, albedo=0.99),
            "torus": lambda: scenes.torus_shell(n=250_000, seed=3, sigma_multiplier=10.0, albedo=0.95)}
    rows = {k: gens[k]() for k in names}
    return {k: ctx.scene(r, bvh=bvh, walk=walk) for k, r in rows.items()}, rows


def mix_requests(volprim, jobs, handles, envh, shadow_rr=0.0):
    reqs = []
    for j in jobs:
        p = volprim.default_params(width=j["width"], height=j["height"], spp=j["spp"], seed=j["seed"], max_depth=64,
                                   rr_depth=12, phase_g=j["phase_g"], active_cap=MIX_ACTIVE_CAP[j["scene"]],
                                   shadow_rr=shadow_rr)
        reqs.append((handles[j["scene"]], envh, volprim.camera_perspective(**j["cam"]), p))
    return reqs


def serve(ctx, reqs, arrivals, mem_gib=16, max_inflight=64):
    """Submit reqs at the given arrival times (seconds from the start) to a
    vp_server; returns (seconds from the first submission to the last
    result, per-request latencies in ms, server stats, images)."""
    import threading
    srv = ctx.server(mem_budget_bytes=mem_gib << 30, max_inflight=max_inflight,
                     active_cap=max(r[3].active_cap for r in reqs))
    tickets = [None] * len(reqs)
    t0 = time.perf_counter()

    def feeder():
        for i, r in enumerate(reqs):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(*r)

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    srv.drain()
    dt = time.perf_counter() - t0
    lat = [srv.poll(t[0]) for t in tickets]
    st = srv.stats()
    srv.close()
    return dt, lat, st, [t[1] for t in tickets]


def rel_rmse_est(ctx, torch, req):
    """Relative RMSE of one request's image from the per-pixel sample variance
    of the same render (an unbiased estimate of its MSE against the converged
    image), and the spp that would reach 1 %."""
    sc, envh, cam, p = req
    out = torch.empty((p.height, p.width, 4), dtype=torch.float32, device="cuda")
    smp = torch.empty((p.spp, p.height, p.width, 4), dtype=torch.float32, device="cuda")
    ctx.render_device(sc, envh, cam, p, out.data_ptr(), samples_ptr=smp.data_ptr())
    torch.cuda.synchronize()
    rgb = smp[..., :3].double()
    var_mean = rgb.var(dim=0, unbiased=True) / p.spp if p.spp > 1 else torch.zeros_like(rgb[0])
    rel = float(var_mean.mean().sqrt() / out[..., :3].double().mean())
    return {"res": p.width, "spp": p.spp, "rel_rmse_est": round(rel, 5),
            "spp_for_1pct_est": int(np.ceil(p.spp * (rel / 0.01) ** 2))}


def run_stream(args):
    """The default measurement: the served Poisson request stream of the mixed
    workload, with plain one-launch-per-request rendering and the batched
    one-launch form of the same requests beside it.  All three return
    bit-identical images (checked), so they are compared at matched RMSE."""
    import torch
    from volprim_amd import scenes, volprim

    t_setup = time.perf_counter()
    ctx = volprim.Context(0)
    dev_name = ctx.device_info()[0]
    K, W = args.steps, args.warmup
    jobs = scenes.mix_jobs(n_jobs=K, seed=args.mix_seed, res_div=args.res_div, spp_div=args.spp_div)
    # warm-up requests: the same distribution at 1/8 resolution and 1 spp
    # (first launches load code objects and grow the allocator pools)
    wjobs = scenes.mix_jobs(n_jobs=W, seed=args.mix_seed + 100, res_div=8, spp_div=256) if W else []
    names = sorted({j["scene"] for j in jobs + wjobs})
    handles, rows = mix_scenes(ctx, scenes, args.bvh, names, args.walk)
    env_rgba = scenes.sky_env(2048, 1024)
    envh = ctx.env(env_rgba)
    reqs = mix_requests(volprim, jobs, handles, envh, args.shadow_rr)
    wreqs = mix_requests(volprim, wjobs, handles, envh, args.shadow_rr)
    samples = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    print(f"[stream] {K} requests ({samples} samples) over {names}, scenes resident after "
          f"{time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    # warm-up: the server path and the plain path once each
    if wreqs:
        serve(ctx, wreqs, np.zeros(len(wreqs)))
        for r in wreqs:
            ctx.render(*r)
    torch.cuda.synchronize()

    # plain: one vp_render per request, in arrival order (the reference's
    # one-launch-per-request renderer, scenes kept resident)
    t0 = time.perf_counter()
    plain_imgs, plain_ms = [], []
    for r in reqs:
        s = time.perf_counter()
        img, st = ctx.render(*r)
        if st["active_overflow"]:
            raise RuntimeError(f"active-set overflow in the plain run: {st}")
        plain_imgs.append(img)
        plain_ms.append((time.perf_counter() - s) * 1e3)
    t_plain = time.perf_counter() - t0
    print(f"[stream] plain {t_plain:.2f} s", file=sys.stderr, flush=True)

    # batched: every request in one call, launch shape from the autotuner
    t0 = time.perf_counter()
    tuned = ctx.autotune(reqs)
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch_imgs, bst = ctx.render_jobs(reqs, launch=tuned)
    t_batch = time.perf_counter() - t0
    print(f"[stream] batched {t_batch:.2f} s (autotune {t_tune:.1f} s, "
          f"{'wavefront' if tuned.pipeline else 'megakernel'})", file=sys.stderr, flush=True)

    # served, saturated: every request submitted at once, so the server's
    # continuous batching, not the arrival process, sets the pace (the value)
    t_served, _, sst, served_imgs = serve(ctx, reqs, np.zeros(K), mem_gib=args.mem_gib)
    print(f"[stream] served (saturated) {t_served:.2f} s, {sst['launches']} launches", file=sys.stderr, flush=True)
    same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) and
               np.array_equal(a.view(np.uint32), c.view(np.uint32))
               for a, b, c in zip(plain_imgs, batch_imgs, served_imgs))

    # served, Poisson arrivals at the stated offered load of that capacity:
    # the request latencies (p50/p99) under load
    lam = args.load * K / t_served
    arrivals = np.cumsum(np.random.default_rng(args.mix_seed + 1).exponential(1.0 / lam, size=K))
    arrivals -= arrivals[0]
    t_poisson, lat, pst, poisson_imgs = serve(ctx, reqs, arrivals, mem_gib=args.mem_gib)
    print(f"[stream] served (Poisson, load {args.load}) {t_poisson:.2f} s", file=sys.stderr, flush=True)
    same = same and all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                        for a, b in zip(plain_imgs, poisson_imgs))

    # the saturated served stream with the reference's estimator (no
    # shadow-ray roulette), for comparison on the same requests
    ref_served = None
    if args.shadow_rr > 0.0 and not args.no_reference:
        rreqs = mix_requests(volprim, jobs, handles, envh, 0.0)
        t_ref, _, rst, _ = serve(ctx, rreqs, np.zeros(K), mem_gib=args.mem_gib)
        ref_served = {"seconds": round(t_ref, 3), "msamples_s": round(samples / t_ref / 1e6, 4),
                      "launches": rst["launches"], "note": "saturated, shadow_rr 0"}
        print(f"[stream] served (saturated), reference estimator {t_ref:.2f} s", file=sys.stderr, flush=True)

    # roofline of the dominant kernel: k_render over the batched table.
    # Algorithmic FLOPs from the event-counting build run on the same jobs
    # (counts are per path, so they do not depend on what shares a launch),
    # over the launch's kernel time from HIP events on the library stream.
    # Each request is counted on its first ceil(spp / count_div) samples of
    # every pixel and scaled by the spp ratio (samples are independent and
    # identically distributed per pixel).
    roofline, events = None, None
    if not args.no_count:
        events = {k: 0.0 for k in volprim.EVENT_NAMES}
        t0 = time.perf_counter()
        for i, (sc, ev_h, cam, p) in enumerate(reqs):
            spp_c = max(1, -(-p.spp // args.count_div))
            pc = type(p).from_buffer_copy(p)
            pc.spp = spp_c
            ev = ctx.count_events(sc, ev_h, cam, pc)
            for k in events:
                events[k] += ev[k] * (p.spp / spp_c)
            print(f"[stream] event counts: request {i + 1}/{K} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
                  flush=True)
        events = {k: int(round(v)) for k, v in events.items()}
        flops = sum(FLOP_WEIGHTS[k] * events[k] for k in FLOP_WEIGHTS)
        kms = bst["kernel_ms"] / max(bst["launches"], 1)
        achieved = flops / max(bst["launches"], 1) / (kms / 1e3) / 1e12
        roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                    "traffic_note": "HBM bytes of this launch are not measured in the default run; "
                                    "profiles/ holds rocprofv3 FETCH_SIZE/WRITE_SIZE runs per workload",
                    "peak_nonpacked": PEAK_FP32_NONPACKED_TFLOPS, "kernel": "k_render (batched launch of the stream)",
                    "avg_launch_ms": round(kms, 3), "flops_per_launch": flops / max(bst["launches"], 1),
                    "weights_note": WEIGHTS_NOTE,
                    "counts_note": f"event counts of every request on its first ceil(spp/{args.count_div}) samples "
                                   f"per pixel, scaled by the spp ratio"}

    quality = None
    if not args.no_quality:
        quality = {}
        for name in names:
            i = next((k for k, j in enumerate(jobs) if j["scene"] == name), None)
            if i is not None:
                quality[name] = rel_rmse_est(ctx, torch, reqs[i])
                if args.shadow_rr > 0.0:
                    sc, ev_h, cam, p = reqs[i]
                    p0 = type(p).from_buffer_copy(p)
                    p0.shadow_rr = 0.0
                    quality[name]["rel_rmse_est_reference_estimator"] = rel_rmse_est(ctx, torch,
                                                                                      (sc, ev_h, cam, p0))[
                        "rel_rmse_est"]
        quality["method"] = ("per-pixel sample variance / spp of the first request of each scene (an unbiased "
                             "estimate of its MSE against the converged image), with the stream's estimator and "
                             "with the reference's (no shadow-ray roulette)")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_mix(rows, env_rgba, handles, jobs, reqs, threads=min(os.cpu_count() or 1, 16))

    ms = lambda v: round(float(v), 2)
    pct = lambda v, q: ms(np.percentile(v, q))
    value = samples / t_served / 1e6
    line = {
        "metric": "Msamples/s at ≤1% RMSE vs reference image; p50/p99 request latency",
        "value": round(value, 4), "unit": "Msamples/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(t_served / K * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"served Poisson stream of {K} render requests (one step = one request) drawn from "
                               f"the batched job mix (seed {args.mix_seed}): resolution U{{64,128,256,512,1024}}^2"
                               + (f" / {args.res_div}" if args.res_div > 1 else "") +
                               f", spp round(exp U[ln 4, ln 256]) / {args.spp_div} (at least 1), scene uniform over "
                               f"scattering 4096 / cloud 10^6 / torus 2.5*10^5, camera on a sphere of radius 3.5, "
                               f"FOV 40, max depth 64, RR from 12, MIS NEE"
                               + (f" with shadow-ray roulette eps {args.shadow_rr}" if args.shadow_rr > 0 else "")
                               + ", 2048x1024 sky + 0.5 deg sun; all requests submitted at once (latency from a "
                                 f"Poisson stream at {args.load} x that capacity)",
                   "requests": K, "samples": samples, "spp_div": args.spp_div, "res_div": args.res_div,
                   "lambda_req_s_latency_run": round(lam, 4), "scenes": {k: MIX_SCENES[k] for k in names}},
        "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "mean": ms(np.mean(lat)),
                       "note": f"served Poisson stream at load {args.load} x the saturated served capacity "
                               f"(lambda {lam:.4f} requests/s), submit to result"},
        "speedup_vs_plain": round(t_plain / t_served, 3),
        "plain": {"seconds": round(t_plain, 3), "msamples_s": round(samples / t_plain / 1e6, 4),
                  "p50_ms": pct(plain_ms, 50), "p99_ms": pct(plain_ms, 99),
                  "note": "one vp_render per request in order; latency = the request's own render time"},
        "batched": {"seconds": round(t_batch, 3), "msamples_s": round(samples / t_batch / 1e6, 4),
                    "speedup_vs_plain": round(t_plain / t_batch, 3), "autotune_s": round(t_tune, 2),
                    "launch": {"pipeline": "wavefront" if tuned.pipeline else "megakernel", "block": tuned.block,
                               "blocks_per_cu": tuned.blocks_per_cu, "active_cap": tuned.active_cap,
                               "regen_min": tuned.regen_min, "pool_paths": tuned.pool_paths},
                    "kernel_ms": round(bst["kernel_ms"], 2)},
        "served": {"seconds": round(t_served, 3), "msamples_s": round(value, 4),
                   "launches": sst["launches"], "busy_ms": round(sst["busy_ms"], 1),
                   "note": "every request submitted at once; value"},
        "served_poisson": {"seconds": round(t_poisson, 3), "msamples_s": round(samples / t_poisson / 1e6, 4),
                           "lambda_req_s": round(lam, 4), "load": args.load, "launches": pst["launches"],
                           "note": "arrival-bound: the last request arrives near K / lambda"},
        "served_reference_estimator": ref_served,
        "images_identical": bool(same),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "quality": quality,
        "device": dev_name,
        "events": events,
        "mrays_per_s": (round((events["traces"] + events["shadow_rays"]) / t_served / 1e6, 2) if events else None),
    }
    print(json.dumps(line), flush=True)
    for h in handles.values():
        h.close()
    envh.close()
    ctx.close()
    if not same:
        sys.exit(1)


def cpu_baseline_mix(rows, env_rgba, handles, jobs, reqs, threads, budget_s=20.0, repeats=3):
    """The oracle (with its CPU BVH) on a bounded sample of the stream: for
    each scene, the first request of that scene on every 16th row and
    column, at a sample count sized to the budget.  Msamples/s is the
    sample-weighted rate over the three scenes' samples in the stream."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import vpo
    tabs = vpo.env_tables(env_rgba)
    stride = 16
    per_scene, notes = {}, []
    for name in sorted(handles):
        i = next((k for k, j in enumerate(jobs) if j["scene"] == name), None)
        if i is None:
            continue
        rec = handles[name].export()[2]
        _, _, cam, p = reqs[i]

        def run(spp_used):
            t0 = time.perf_counter()
            vpo.render(rec, env_rgba, cam.pod(), p.width, p.height, p.spp, seed=p.seed, max_depth=p.max_depth,
                       rr_depth=p.rr_depth, phase_g=p.phase_g, active_cap=p.active_cap, stride=stride,
                       shadow_rr=p.shadow_rr,
                       samples=(0, spp_used), threads=threads, env_tables_=tabs, bvh=True)
            return time.perf_counter() - t0

        npx = ((p.width + stride - 1) // stride) * ((p.height + stride - 1) // stride)
        probe = run(1)
        spp_used = int(max(1, min(p.spp, budget_s / len(handles) / repeats / max(probe, 1e-3))))
        times = sorted(run(spp_used) for _ in range(repeats))
        per_scene[name] = npx * spp_used / times[len(times) // 2] / 1e6
        notes.append(f"{name}: {npx} px x {spp_used} spp")
    # time per sample of the stream's scene mix
    w = {k: sum(j["width"] * j["height"] * j["spp"] for j in jobs if j["scene"] == k) for k in per_scene}
    tot = sum(w.values())
    rate = tot / sum(w[k] / per_scene[k] for k in per_scene) if tot else 0.0
    return {"value": round(rate, 5), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "per_scene": {k: round(v, 5) for k, v in per_scene.items()},
            "sample": "the oracle (CPU BVH) on every 16th row and column of the first request of each scene ("
                      + "; ".join(notes) + f"), 1 warm-up + {repeats} timed repeats (median); the stream's rate "
                      "weights each scene by its share of the stream's samples"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="requests timed (--single: renders timed)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--count-div", type=int, default=8,
                    help="stream: event counts on ceil(spp / this) samples per pixel, scaled up")
    ap.add_argument("--no-quality", action="store_true", help="skip the image-error estimate")
    ap.add_argument("--spp-div", type=int, default=8, help="stream: every request's spp is divided by this")
    ap.add_argument("--res-div", type=int, default=1, help="stream: every request's resolution is divided by this")
    ap.add_argument("--load", type=float, default=0.7,
                    help="stream: Poisson offered load as a multiple of the saturated served capacity")
    ap.add_argument("--mix-seed", type=int, default=4)
    ap.add_argument("--mem-gib", type=int, default=16, help="stream: server admission budget")
    ap.add_argument("--single", choices=["scattering", "cloud", "torus"], default=None,
                    help="render one scene whole per step instead of the request stream")
    ap.add_argument("--traffic", action="store_true", help="--single: two rocprofv3 --pmc runs for HBM traffic")
    ap.add_argument("--no-traffic", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--rmse-ref-spp", type=int, default=0, help="--single: also measure relative RMSE against a "
                                                                   "render with this many spp (untimed)")
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup (0 = library default)")
    ap.add_argument("--bpc", type=int, default=0, help="persistent workgroups per CU (0 = occupancy limit)")
    ap.add_argument("--regen", type=int, default=0, help="idle lanes per wavefront before claiming new paths")
    ap.add_argument("--pipeline", choices=["mega", "wave"], default="mega",
                    help="persistent megakernel or wavefront-queue pipeline")
    ap.add_argument("--pool", type=int, default=0, help="wavefront path slots (0 = library default)")
    ap.add_argument("--bvh", choices=["lbvh", "sah", "median"], default="median",
                    help="leaf order of the walk tree: host median splits on the walk tree's record boundaries "
                         "(default), host SAH tree, or the GPU LBVH's Morton order")
    ap.add_argument("--walk", choices=["auto", "f16", "f32"], default="auto",
                    help="walk-tree records: binary16 when the leaf boxes grow little (auto), or forced")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB elementwise copy kernel first (HBM counter calibration)")
    ap.add_argument("--ris", action="store_true",
                    help="NEE by resampled importance sampling (the reference's --ris) instead of MIS")
    ap.add_argument("--shadow-rr", type=float, default=1e-3,
                    help="shadow-ray roulette threshold eps (0 = off, the reference's estimator)")
    ap.add_argument("--no-reference", action="store_true",
                    help="stream: skip the served run with the reference's estimator")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build renders on one GPU",
              file=sys.stderr)
        sys.exit(2)

    import torch
    load_pkg()
    if args.single is None:
        run_stream(args)
        return
    run_single(args, torch)


def run_single(args, torch):
    from volprim_amd import scenes, volprim

    rows, cam_args, pp, desc = workload(scenes, args.single)
    env = scenes.sky_env(2048, 1024)
    if args.ris:
        pp = dict(pp, use_ris=1)
    if args.shadow_rr:
        pp = dict(pp, shadow_rr=args.shadow_rr)
    ctx = volprim.Context(0)
    dev_name, n_cu, clk = ctx.device_info()
    scene = ctx.scene(rows, bvh=args.bvh, walk=args.walk)
    walk_half, walk_growth = scene.walk_info()
    envh = ctx.env(env)
    cam = volprim.camera_perspective(**cam_args)
    params = volprim.default_params(**pp)
    out = torch.empty((pp["height"], pp["width"], 4), dtype=torch.float32, device="cuda")
    launch = None
    if args.block or args.bpc or args.regen or args.pipeline != "mega" or args.pool:
        launch = volprim.LaunchConfig(args.block, args.bpc, 0, 0, args.regen,
                                      volprim.PIPELINE_WAVEFRONT if args.pipeline == "wave" else 0, args.pool)
    torch.cuda.synchronize()

    if args.copy_calibration:
        # a plain elementwise kernel of known traffic (reads and writes
        # 256 MiB each) for scaling the HBM counters of a --pmc run
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b
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
        traffic, traffic_note = None, "not measured (pass --traffic)"
        if args.traffic and not args.no_traffic and args.pipeline == "mega":
            tr, traffic_note = measure_traffic(args)
            if tr is not None:
                traffic = int(tr["bytes"])
                traffic_note = (f"rocprofv3 --pmc: FETCH_SIZE {tr['fetch_kib']:.0f} KiB x {tr['fetch_scale']:.3f}, "
                                f"WRITE_SIZE {tr['write_kib']:.0f} KiB x {tr['write_scale']:.3f} (scales from a "
                                f"256 MiB read + 256 MiB write elementwise kernel)")
        elif args.pipeline != "mega":
            traffic_note = "measured for the megakernel only"
        roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic, "traffic_note": traffic_note,
                    "peak_nonpacked": PEAK_FP32_NONPACKED_TFLOPS,
                    "kernel": "k_render" if args.pipeline == "mega" else "k_wf_* (one render step)",
                    "avg_launch_ms": round(avg_launch_s * 1e3, 3), "flops_per_launch": flops_per_launch,
                    "weights_note": WEIGHTS_NOTE}

    # image error: relative RMSE of the timed image from its per-pixel sample
    # variance (an unbiased estimate of the MSE against the converged image),
    # and optionally measured against a high-spp render of the same view
    quality = None
    if not args.no_quality:
        smp = torch.empty((pp["spp"], pp["height"], pp["width"], 4), dtype=torch.float32, device="cuda")
        ctx.render_device(scene, envh, cam, params, out.data_ptr(), samples_ptr=smp.data_ptr(), launch=launch)
        torch.cuda.synchronize()
        rgb = smp[..., :3].double()
        var_mean = rgb.var(dim=0, unbiased=True) / pp["spp"]
        mean_img = out[..., :3].double()
        rel_est = float(var_mean.mean().sqrt() / mean_img.mean())
        quality = {"rel_rmse_est": round(rel_est, 5), "spp": pp["spp"],
                   "spp_for_1pct_est": int(np.ceil(pp["spp"] * (rel_est / 0.01) ** 2)),
                   "method": "per-pixel sample variance / spp"}
        del smp, rgb, var_mean
        if args.rmse_ref_spp:
            ref = torch.empty_like(out)
            pr = volprim.default_params(**dict(pp, spp=args.rmse_ref_spp, seed=pp["seed"] + 1))
            t_ref = time.perf_counter()
            ctx.render_device(scene, envh, cam, pr, ref.data_ptr(), launch=launch)
            torch.cuda.synchronize()
            t_ref = time.perf_counter() - t_ref
            R = ref[..., :3].double()
            mse = float(((mean_img - R) ** 2).mean())
            quality.update({"rel_rmse_vs_ref": round(float(np.sqrt(mse)) / float(R.mean()), 5),
                            "ref_spp": args.rmse_ref_spp, "ref_seconds": round(t_ref, 2),
                            "ref_note": "reference rendered with seed+1 (independent samples); its own noise "
                                        "adds spp/ref_spp of the variance"})
        del mean_img

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(rows, env, cam.pod(), pp, threads=min(os.cpu_count() or 1, 16))

    line = {
        "metric": f"Msamples/s of one {args.single} render (profiling workload, not the request-stream metric)",
        "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc + ", " + ("RIS NEE (6 candidates)" if args.ris else "MIS NEE")
                               + (f", shadow-ray roulette eps {args.shadow_rr}" if args.shadow_rr else ""),
                   "width": pp["width"], "height": pp["height"], "spp": pp["spp"], "primitives": int(rows.shape[0])},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device": dev_name,
        "kernel_ms_per_step": round(kernel_ms / args.steps, 3),
        "pipeline": args.pipeline,
        "bvh": args.bvh,
        "walk_records": {"form": "binary16" if walk_half else "binary32", "leaf_box_surface_growth":
                         round(walk_growth, 5)},
        "events_per_step": events,
        # rays traced per second: camera and continuation rays (one argmin
        # walk each) plus shadow rays, from the counted events of one render
        "mrays_per_s": (round((events["traces"] + events["shadow_rays"]) * args.steps / dt / 1e6, 1)
                        if events else None),
        "quality": quality,
        "image_mean": float(img[..., 0:3].mean()),
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
