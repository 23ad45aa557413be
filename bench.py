"""Benchmark: Msamples/s of the Gaussian-mixture volumetric path tracer on one
MI355X, on the scattering workload (4096 anisotropic Gaussians, 512x512,
64 spp, max depth 64, Russian roulette from depth 12, albedo 0.9, isotropic
phase, 2048x1024 procedural sky environment with a 0.5 degree sun).

A step is one full render of that image (512*512*64 = 16.8 M camera paths)
with the scene, BVH (the median-split tree unless --bvh says otherwise) and environment already
resident in HBM.  Prints one
JSON line with the throughput, the roofline of the render kernel (FP32 VALU:
algorithmic FLOPs from the event-counting kernel variant weighted as
documented in DESIGN.md, over the kernel time measured with HIP events on
the library stream) and the CPU baseline (the oracle's C restatement of the
estimator with its CPU BVH, on a bounded pixel sample, all host cores).

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
# or rcp/sqrt = 1, erf = 16, erfinv = 20 (derivations in DESIGN.md).  A node
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


def workload(scenes):
    rows = scenes.c1_scattering(n=4096, seed=1, sigma_multiplier=10.0, albedo=0.9)
    env = scenes.sky_env(2048, 1024)
    cam_args = scenes.default_camera_args(512, 512, dist=3.5, fov=40.0)
    params = dict(width=512, height=512, spp=64, seed=0, max_depth=64, rr_depth=12, phase_g=0.0)
    return rows, env, cam_args, params


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
                   env_tables_=tabs, bvh=True)
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
            "sample": f"{npx} pixels (every {stride}th row and column of 512x512) x {spp_used} spp = {n} paths per "
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
    child = [sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "0", "--no-cpu", "--no-count",
             "--no-quality", "--no-traffic", "--copy-calibration", "--bvh", args.bvh]
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


def run_mix(args):
    """The job mix three ways on resident scenes: plain (one vp_render per
    request, in order), batched (every job in one vp_render_jobs call with
    the autotuned launch), served (a Poisson request stream into vp_server).
    All three must return bit-identical images."""
    import threading
    from volprim_amd import scenes, volprim

    t0 = time.perf_counter()
    gens = {"scattering": lambda: scenes.c1_scattering(n=4096, seed=1, sigma_multiplier=10.0, albedo=0.9),
            "cloud": lambda: scenes.c2_cloud(n=1_000_000, seed=2, sigma_multiplier=7.5, albedo=0.99),
            "torus": lambda: scenes.torus_shell(n=250_000, seed=3, sigma_multiplier=10.0, albedo=0.95)}
    jobs = scenes.mix_jobs(n_jobs=args.mix_jobs, seed=4, res_div=args.mix_res_div, spp_div=args.mix_spp_div)
    ctx = volprim.Context(0)
    dev_name = ctx.device_info()[0]
    handles = {k: ctx.scene(g(), bvh=args.bvh) for k, g in gens.items() if any(j["scene"] == k for j in jobs)}
    envh = ctx.env(scenes.sky_env(2048, 1024))
    print(f"[mix] {len(jobs)} jobs, scenes and environment resident after {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    reqs = []
    for j in jobs:
        p = volprim.default_params(width=j["width"], height=j["height"], spp=j["spp"], seed=j["seed"], max_depth=64,
                                   rr_depth=12, phase_g=j["phase_g"], active_cap=MIX_ACTIVE_CAP[j["scene"]])
        reqs.append((handles[j["scene"]], envh, volprim.camera_perspective(**j["cam"]), p))
    samples = sum(j["width"] * j["height"] * j["spp"] for j in jobs)

    # plain: one request at a time, one render call each
    t0 = time.perf_counter()
    plain_imgs, plain_lat = [], []
    for r in reqs:
        s = time.perf_counter()
        img, st = ctx.render(*r)
        if st["active_overflow"] or st["hit_overflow"]:
            raise RuntimeError(f"active-set overflow in the plain run: {st}")
        plain_imgs.append(img)
        plain_lat.append((time.perf_counter() - s) * 1e3)
    t_plain = time.perf_counter() - t0
    print(f"[mix] plain {t_plain:.2f} s", file=sys.stderr, flush=True)

    # batched: all jobs in one call, launch shape from the autotuner
    t0 = time.perf_counter()
    tuned = ctx.autotune(reqs)
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch_imgs, bst = ctx.render_jobs(reqs, launch=tuned)
    t_batch = time.perf_counter() - t0
    print(f"[mix] batched {t_batch:.2f} s (autotune {t_tune:.1f} s)", file=sys.stderr, flush=True)

    # served: (1) every request submitted at once (capacity), (2) a Poisson
    # stream at half the batched capacity (latency under load)
    def serve(arrivals):
        srv = ctx.server(mem_budget_bytes=args.mix_mem_gib << 30, max_inflight=64)
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

    t_sat, lat_sat, st_sat, served_imgs = serve(np.zeros(len(reqs)))
    print(f"[mix] served, all at once {t_sat:.2f} s", file=sys.stderr, flush=True)
    rate = args.mix_load * len(reqs) / t_batch
    arr = np.cumsum(np.random.default_rng(5).exponential(1.0 / rate, size=len(reqs)))
    t_served, lat, sst, served_imgs2 = serve(arr)
    print(f"[mix] served, Poisson {t_served:.2f} s", file=sys.stderr, flush=True)

    same = all(np.array_equal(a, b) and np.array_equal(a, c) and np.array_equal(a, d)
               for a, b, c, d in zip(plain_imgs, batch_imgs, served_imgs, served_imgs2))
    ms = lambda v: round(float(v), 2)
    line = {
        "metric": "Msamples/s on the mixed request stream (plain vs batched vs served); p50/p99 request latency",
        "value": round(samples / t_sat / 1e6, 4), "unit": "Msamples/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic", "device": dev_name,
        "config": {"workload": f"{len(jobs)} jobs (seed 4): resolution U{{64..1024}}^2 / {args.mix_res_div}, spp "
                               f"round(exp U[ln4, ln256]) / {args.mix_spp_div}, scenes U{{scattering 4096, cloud 1M, "
                               f"torus 250k}}, max depth 64, RR 12, sky environment 2048x1024",
                   "samples": samples, "jobs": len(jobs)},
        "plain": {"seconds": round(t_plain, 3), "msamples_s": round(samples / t_plain / 1e6, 4),
                  "p50_ms": ms(np.percentile(plain_lat, 50)), "p99_ms": ms(np.percentile(plain_lat, 99))},
        "batched": {"seconds": round(t_batch, 3), "msamples_s": round(samples / t_batch / 1e6, 4),
                    "launch": {"pipeline": "wavefront" if tuned.pipeline else "megakernel", "block": tuned.block,
                               "blocks_per_cu": tuned.blocks_per_cu, "active_cap": tuned.active_cap,
                               "regen_min": tuned.regen_min, "pool_paths": tuned.pool_paths},
                    "autotune_s": round(t_tune, 2), "launches": bst["launches"],
                    "speedup_vs_plain": round(t_plain / t_batch, 3)},
        "served_saturated": {"seconds": round(t_sat, 3), "msamples_s": round(samples / t_sat / 1e6, 4),
                             "p50_ms": ms(np.percentile(lat_sat, 50)), "p99_ms": ms(np.percentile(lat_sat, 99)),
                             "speedup_vs_plain": round(t_plain / t_sat, 3), "server_stats": st_sat},
        "served_poisson": {"seconds": round(t_served, 3), "msamples_s": round(samples / t_served / 1e6, 4),
                           "lambda_req_s": round(rate, 3), "offered_load_vs_batched": args.mix_load,
                           "requests": len(reqs), "p50_ms": ms(np.percentile(lat, 50)),
                           "p99_ms": ms(np.percentile(lat, 99)), "server_stats": sst},
        "images_identical": bool(same),
    }
    print(json.dumps(line), flush=True)
    for h in handles.values():
        h.close()
    envh.close()
    ctx.close()
    if not same:
        sys.exit(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--no-quality", action="store_true", help="skip the image-error estimate")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 --pmc runs for HBM traffic")
    ap.add_argument("--rmse-ref-spp", type=int, default=0, help="also measure relative RMSE against a render "
                                                                   "with this many spp (untimed)")
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup (0 = library default)")
    ap.add_argument("--bpc", type=int, default=0, help="persistent workgroups per CU (0 = occupancy limit)")
    ap.add_argument("--regen", type=int, default=0, help="idle lanes per wavefront before claiming new paths")
    ap.add_argument("--pipeline", choices=["mega", "wave"], default="mega",
                    help="persistent megakernel or wavefront-queue pipeline")
    ap.add_argument("--pool", type=int, default=0, help="wavefront path slots (0 = library default)")
    ap.add_argument("--bvh", choices=["lbvh", "sah", "median"], default="median",
                    help="leaf order of the walk tree: host median splits on the walk tree's record boundaries "
                         "(default), host SAH tree, or the GPU LBVH's Morton order")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB elementwise copy kernel first (HBM counter calibration)")
    ap.add_argument("--ris", action="store_true",
                    help="NEE by resampled importance sampling (the reference's --ris) instead of MIS")
    ap.add_argument("--mix", action="store_true", help="run the job mix (plain vs batched vs served) instead")
    ap.add_argument("--mix-jobs", type=int, default=64)
    ap.add_argument("--mix-res-div", type=int, default=4)
    ap.add_argument("--mix-spp-div", type=int, default=4)
    ap.add_argument("--mix-load", type=float, default=0.5,
                    help="Poisson offered load as a multiple of the batched capacity")
    ap.add_argument("--mix-mem-gib", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design is in DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    import torch
    pkg = load_pkg()
    if args.mix:
        run_mix(args)
        return
    from volprim_amd import scenes, volprim

    rows, env, cam_args, pp = workload(scenes)
    if args.ris:
        pp = dict(pp, use_ris=1)
    ctx = volprim.Context(0)
    dev_name, n_cu, clk = ctx.device_info()
    scene = ctx.scene(rows, bvh=args.bvh)
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
        traffic, traffic_note = None, "not measured (--no-traffic)"
        if not args.no_traffic and args.pipeline == "mega":
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
                    "avg_launch_ms": round(avg_launch_s * 1e3, 3), "flops_per_launch": flops_per_launch}

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
        "metric": "Msamples/s at <=1% RMSE vs reference image; p50/p99 request latency",
        "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "scattering: 4096 anisotropic Gaussians (seed 1), sigma multiplier 10, albedo 0.9, "
                               "isotropic phase, 2048x1024 sky + sun environment, 512x512, 64 spp, max depth 64, "
                               "RR from depth 12, " + ("RIS NEE (6 candidates)" if args.ris else "MIS NEE"),
                   "width": 512, "height": 512, "spp": 64,
                   "primitives": 4096},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device": dev_name,
        "kernel_ms_per_step": round(kernel_ms / args.steps, 3),
        "pipeline": args.pipeline,
        "bvh": args.bvh,
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
