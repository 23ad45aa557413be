"""GPU parity at the benchmark configurations' sizes, and the checks that pin
the measurement: the single-Gaussian workload at its full 128x128x16, a
camera inside the medium (non-empty bounce-0 origin sets), the 10^6-Gaussian
cloud and the 2.5*10^5-Gaussian torus at 1024x1024, the 64-job batched mix
over the three resident scenes, a served Poisson stream of 256 requests, the
furnace at 1e-3 for MIS and RIS, the --measure-caps maxima, the
event counters behind the roofline numerator and vp_normalize_accumulator.

Images are compared bit for bit: megakernel against wavefront pipeline,
batched and served against solo renders, and strided pixels against the CPU
oracle on the same seeds.  Each case prints its name and sizes before it
touches the GPU.
"""
import os
import threading
import time

import numpy as np
import pytest

import vpo
from volprim_amd import scenes, volprim

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)
# active-set capacities of the dense scenes: the reference's rule
# (ceil(1.125 v) rounded up to 16) on their measured maximum overlaps
# (cloud 319, torus 267), rounded up to 384
DENSE_CAP = 384


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitwise(gpu, ref, what):
    g, r = np.asarray(gpu, np.float32), np.asarray(ref, np.float32)
    assert g.shape == r.shape, what
    diff = bits(g) != bits(r)
    if diff.any():
        i = np.flatnonzero(diff.reshape(-1))[:5]
        raise AssertionError(f"{what}: {int(diff.sum())} of {g.size} differ; first {i}: gpu={g.reshape(-1)[i]} "
                             f"ref={r.reshape(-1)[i]}")


def oracle_render(rec, env, cam, p, tables=None, **kw):
    ref, _, k = vpo.render(rec, env, cam.pod(), p.width, p.height, p.spp, seed=p.seed, max_depth=p.max_depth,
                           rr_depth=p.rr_depth, rr_max_survival=p.rr_max_survival, firefly_clamp=p.firefly_clamp,
                           filter_stddev=p.filter_stddev, phase_g=p.phase_g, use_ris=bool(p.use_ris),
                           ris_candidates=p.ris_candidates, active_cap=p.active_cap, threads=THREADS, shadow_rr=p.shadow_rr,
                           env_tables_=tables, **kw)
    return ref, k


# ---------------------------------------------------------------------------
# single Gaussian at full size; camera inside the medium
# ---------------------------------------------------------------------------
def test_c0_full_size_bitexact(ctx):
    # one isotropic Gaussian (s = 0.25, central optical depth 2), absorption
    # only, white environment, 128x128 at 16 spp, eye at distance 3.5
    rows = scenes.c0_single_gaussian()
    env = scenes.constant_env()
    cam = volprim.camera_perspective(**scenes.default_camera_args(128, 128))
    p = volprim.default_params(width=128, height=128, spp=16, seed=42)
    print("[case] C0 single Gaussian 128x128x16, megakernel + wavefront vs oracle", flush=True)
    sc = ctx.scene(rows)
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    img, st = ctx.render(sc, ev, cam, p)
    wf = volprim.LaunchConfig(256, 0, 0, 0, 0, volprim.PIPELINE_WAVEFRONT, 1 << 16)
    img_wf = ctx.render_jobs([(sc, ev, cam, p)], launch=wf)[0][0]
    sc.close()
    ev.close()
    ref, _ = oracle_render(rec, env, cam, p)
    assert_bitwise(img, ref, "C0 megakernel vs oracle")
    assert_bitwise(img_wf, ref, "C0 wavefront vs oracle")
    # the central pixel sees exp(-2) of the white field (up to the 16-sample
    # spread of the jittered rays over the pixel)
    c = img[63:65, 63:65, 0].mean()
    assert abs(c - np.exp(-2.0)) < 0.01, c


INSIDE_CASES = {
    # eye at the origin inside a cluster of large Gaussians, MIS and RIS
    "mis": dict(kw={}),
    "ris": dict(kw=dict(use_ris=1, ris_candidates=4)),
}


@pytest.mark.parametrize("case", sorted(INSIDE_CASES))
def test_camera_inside_medium_bitexact(ctx, case):
    rows = scenes.random_scene(60, 31, extent=0.6, smin=0.15, smax=0.45, albedo=0.85)
    rec_in = vpo.make_prims(rows)
    eye = np.zeros(3, np.float32)
    # the eye lies inside several primitives' 3-sigma bounds: bounce 0 starts
    # with a non-empty origin set (raygen.cuh:148-160, sampling.cuh:322-341)
    probe = np.array([[0, 0, 0, 0, 0, 1, 0, 0, 0]], np.float32)
    inside = sum(int(vpo.prim_eval("point_inside", r, probe)[0] > 0) for r in rec_in)
    assert inside >= 2, inside
    cam = volprim.camera_perspective(24, 20, eye, np.array([0.3, 0.1, 1.0], np.float32),
                                     np.array([0, 1, 0], np.float32), 70.0)
    env = scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    p = volprim.default_params(width=24, height=20, spp=3, seed=7, phase_g=0.4, **INSIDE_CASES[case]["kw"])
    print(f"[case] camera inside the medium ({case}) 24x20x3, origin set {inside}", flush=True)
    sc = ctx.scene(rows)
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    img, st = ctx.render(sc, ev, cam, p)
    wf = volprim.LaunchConfig(256, 0, 0, 0, 1, volprim.PIPELINE_WAVEFRONT, 4096)
    img_wf = ctx.render_jobs([(sc, ev, cam, p)], launch=wf)[0][0]
    sc.close()
    ev.close()
    ref, k = oracle_render(rec, env, cam, p)
    assert_bitwise(img, ref, f"inside ({case}) megakernel vs oracle")
    assert_bitwise(img_wf, ref, f"inside ({case}) wavefront vs oracle")
    assert st["active_overflow"] == k["active_overflow"]
    # the origin-set terms were exercised at bounce 0 on every path
    assert k["n_arg_active"] >= p.width * p.height * p.spp


# ---------------------------------------------------------------------------
# dense scenes at 1024x1024
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def sky(ctx):
    rgba = scenes.sky_env(2048, 1024)
    h = ctx.env(rgba)
    yield rgba, h, vpo.env_tables(rgba)
    h.close()


@pytest.mark.parametrize("name", ["cloud", "torus"])
def test_dense_scene_1024_bitexact(ctx, sky, name):
    # cloud: 10^6 Gaussians from fBm noise, albedo 0.99, HG g = 0.85;
    # torus: 2.5*10^5 Gaussians on a torus shell, albedo 0.95, isotropic.
    # 1024x1024 at 2 spp, depth 64, RR from 12, median-split tree.  The
    # megakernel and the wavefront pipeline agree on every pixel; every 64th
    # row and column matches the oracle (with its CPU BVH); no active-set
    # overflow at the capacity the measured overlaps give.
    rows = scenes.c2_cloud() if name == "cloud" else scenes.torus_shell()
    g = 0.85 if name == "cloud" else 0.0
    env_rgba, envh, tables = sky
    cam = volprim.camera_perspective(**scenes.default_camera_args(1024, 1024, view="diag"))
    p = volprim.default_params(width=1024, height=1024, spp=2, seed=3, max_depth=64, rr_depth=12, phase_g=g,
                               active_cap=DENSE_CAP)
    print(f"[case] {name} {rows.shape[0]} Gaussians 1024x1024x2 megakernel + wavefront", flush=True)
    sc = ctx.scene(rows, bvh="median")
    _, _, rec, _ = sc.export()
    img, st = ctx.render(sc, envh, cam, p)
    wf = volprim.LaunchConfig(256, 0, DENSE_CAP, 0, 0, volprim.PIPELINE_WAVEFRONT, 1 << 20)
    imgs_wf, st_wf = ctx.render_jobs([(sc, envh, cam, p)], launch=wf)
    sc.close()
    assert st["active_overflow"] == 0 and st_wf["active_overflow"] == 0
    assert_bitwise(imgs_wf[0], img, f"{name}: wavefront vs megakernel, full image")
    s = 64
    print(f"[case] {name}: oracle on every {s}th row and column", flush=True)
    ref, k = oracle_render(rec, env_rgba, cam, p, tables=tables, stride=s, bvh=True)
    assert_bitwise(img[::s, ::s], ref[::s, ::s], f"{name} vs oracle (strided pixels)")
    assert k["active_overflow"] == 0
    assert np.isfinite(img).all() and img[..., :3].mean() > 0.0


# ---------------------------------------------------------------------------
# batched mix and served stream over the three resident scenes
# ---------------------------------------------------------------------------
MIX_CAP = {"scattering": 128, "cloud": DENSE_CAP, "torus": DENSE_CAP}


@pytest.fixture(scope="module")
def resident(ctx, sky):
    gens = {"scattering": lambda: scenes.c1_scattering(n=4096, seed=1, sigma_multiplier=10.0, albedo=0.9),
            "cloud": scenes.c2_cloud, "torus": scenes.torus_shell}
    hs, recs = {}, {}
    for k, gen in gens.items():
        print(f"[setup] resident scene {k}", flush=True)
        hs[k] = ctx.scene(gen(), bvh="median")
        recs[k] = hs[k].export()[2]
    yield hs, recs
    for h in hs.values():
        h.close()


def _requests(jobs, hs, envh):
    out = []
    for j in jobs:
        p = volprim.default_params(width=j["width"], height=j["height"], spp=j["spp"], seed=j["seed"], max_depth=64,
                                   rr_depth=12, phase_g=j["phase_g"], active_cap=MIX_CAP[j["scene"]])
        out.append((hs[j["scene"]], envh, volprim.camera_perspective(**j["cam"]), p))
    return out


def test_c3_batched_mix_bitexact(ctx, sky, resident):
    # the 64-job mix (resolution uniform over {64..1024}^2 / 4, spp
    # round(exp U[ln 4, ln 256]) / 16, scenes uniform over the three): one
    # batched call equals 64 solo renders bit for bit; one job of each scene
    # matches the oracle on every 8th row and column
    env_rgba, envh, tables = sky
    hs, recs = resident
    jobs = scenes.mix_jobs(n_jobs=64, seed=4, res_div=4, spp_div=16)
    reqs = _requests(jobs, hs, envh)
    print(f"[case] C3 batched mix: {len(jobs)} jobs, "
          f"{sum(j['width'] * j['height'] * j['spp'] for j in jobs)} samples", flush=True)
    batched, st = ctx.render_jobs(reqs)
    assert st["active_overflow"] == 0
    print("[case] C3: solo renders", flush=True)
    for i, r in enumerate(reqs):
        assert_bitwise(batched[i], ctx.render(*r)[0], f"C3 job {i} ({jobs[i]['scene']}) batched vs solo")
    for name in ("scattering", "cloud", "torus"):
        i = next(k for k, j in enumerate(jobs) if j["scene"] == name)
        _, _, cam, p = reqs[i]
        print(f"[case] C3 job {i} ({name} {p.width}x{p.height}x{p.spp}) vs oracle, every 8th row/column", flush=True)
        ref, _ = oracle_render(recs[name], env_rgba, cam, p, tables=tables, stride=8, bvh=True)
        assert_bitwise(batched[i][::8, ::8], ref[::8, ::8], f"C3 job {i} vs oracle")


def test_c4_served_poisson_stream_bitexact(ctx, sky, resident):
    # 256 requests with the mix's size and scene distribution (resolution / 8,
    # spp / 16), arriving as a Poisson stream faster than the server finishes
    # them, so requests join while others run; every 4th result equals its
    # solo render bit for bit.  Completion is polled against a deadline, so a
    # stalled server fails the test instead of blocking it.
    env_rgba, envh, tables = sky
    hs, recs = resident
    jobs = scenes.mix_jobs(n_jobs=256, seed=5, res_div=8, spp_div=16)
    reqs = _requests(jobs, hs, envh)
    arrivals = np.cumsum(np.random.default_rng(5).exponential(1.0 / 400.0, size=len(reqs)))
    print(f"[case] C4 served stream: {len(reqs)} requests, lambda 400/s", flush=True)
    srv = ctx.server(mem_budget_bytes=4 << 30, max_inflight=64, active_cap=DENSE_CAP)
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
    lats = [None] * len(reqs)
    deadline = t0 + 150.0
    next_report = t0 + 10.0
    while any(l is None for l in lats):
        for i, (t, _) in enumerate(tickets):
            if lats[i] is None:
                lats[i] = srv.poll(t)
        done = sum(l is not None for l in lats)
        now = time.perf_counter()
        if now >= next_report:
            print(f"[case] C4: {done}/{len(reqs)} done after {now - t0:.1f} s", flush=True)
            next_report = now + 10.0
        assert now < deadline, f"served stream stalled: {done}/{len(reqs)} requests done after {now - t0:.0f} s"
        time.sleep(0.05)
    srv.drain()
    st = srv.stats()
    srv.close()
    print(f"[case] C4: served in {time.perf_counter() - t0:.2f} s, {st['launches']} launches, "
          f"p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms; solo renders of every 4th request", flush=True)
    assert st["completed"] == len(reqs)
    assert st["launches"] >= 2, "no request joined while another was running"
    for i in range(0, len(reqs), 4):
        assert_bitwise(tickets[i][1], ctx.render(*reqs[i])[0], f"C4 request {i} ({jobs[i]['scene']}) served vs solo")
        if i % 64 == 0:
            print(f"[case] C4: solo {i + 1}/{len(reqs)} ({time.perf_counter() - t0:.1f} s)", flush=True)


# ---------------------------------------------------------------------------
# furnace, measuring run, event counters, normalize
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ris", [0, 1])
def test_furnace_1e3(ctx, ris):
    # albedo-1 medium in a constant white environment reproduces the field:
    # |mean - 1| < 1e-3 (experiments/mitsuba-reference/furnace_check.py:36-40),
    # with enough samples (256x256x256, per-sample spread about 0.7) that
    # 5 standard errors stay below 1e-3
    rows = scenes.random_scene(64, 11, albedo=1.0)
    rows[:, 10:13] = 1.0
    sc = ctx.scene(rows)
    ev = ctx.env(scenes.constant_env())
    cam = volprim.camera_perspective(**scenes.default_camera_args(256, 256))
    p = volprim.default_params(width=256, height=256, spp=256, max_depth=512, rr_depth=512, phase_g=0.3,
                               use_ris=ris)
    print(f"[case] furnace 256x256x256 {'RIS' if ris else 'MIS'}", flush=True)
    img, _ = ctx.render(sc, ev, cam, p)
    sc.close()
    ev.close()
    px = img[..., 0:3].mean(axis=2).astype(np.float64)
    m = px.mean()
    sem = px.std() / np.sqrt(px.size)
    print(f"furnace mean {m:.6f} sem {sem:.2e}", flush=True)
    assert 5 * sem < 1e-3, sem
    assert abs(m - 1.0) < 1e-3, (m, sem)


def test_measure_caps_match_oracle(ctx):
    # --measure-caps (device/core/sampling.cuh:338-340,521-530;
    # src/thesis/host/app/renderer.cpp:405-418): the largest entry count of
    # one traced ray and the largest number of primitives containing one
    # vertex equal the oracle's, and the image is unchanged
    rows = scenes.random_scene(120, 13, smax=0.4, albedo=0.9)
    env = scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    cam = volprim.camera_perspective(**scenes.default_camera_args(24, 24))
    p = volprim.default_params(width=24, height=24, spp=3, seed=5, phase_g=0.2, measure_caps=1)
    print("[case] measure caps 24x24x3, 120 Gaussians", flush=True)
    sc = ctx.scene(rows)
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    img, st = ctx.render(sc, ev, cam, p)
    p0 = volprim.default_params(width=24, height=24, spp=3, seed=5, phase_g=0.2)
    img0, _ = ctx.render(sc, ev, cam, p0)
    sc.close()
    ev.close()
    ref, k = oracle_render(rec, env, cam, p, measure=True)
    assert st["max_hits_per_ray"] == k["max_hits"] > 0
    assert st["max_point_overlap"] == k["max_overlap"] > 0
    assert_bitwise(img, ref, "measuring render vs oracle")
    assert_bitwise(img0, img, "measuring render vs plain render")


# Counters that are properties of the estimator (not of how a walk visits
# the tree): the GPU's event-counting kernel and the brute-force oracle must
# agree exactly.  Walk-dependent counts (entered primitives the argmin
# evaluates before pruning, rebuild candidates, node visits, leaf tests) are
# the counting kernel's own work and have no brute-force counterpart.
ESTIMATOR_COUNTERS = {"paths": None, "traces": "n_trace", "scatters": "n_scatter", "shadow_rays": "n_shadow_rays",
                      "arg_active": "n_arg_active", "rb_active": "n_rb_active", "alb": "n_alb",
                      "sh_active": "n_sh_active", "sh_hit": "n_sh_hit"}


@pytest.mark.parametrize("seed,ris", [(14, 0), (15, 1)])
def test_event_counters_match_oracle(ctx, seed, ris):
    rows = scenes.random_scene(90, seed, smax=0.35, albedo=0.9)
    env = scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    cam = volprim.camera_perspective(**scenes.default_camera_args(20, 16))
    p = volprim.default_params(width=20, height=16, spp=4, seed=seed, phase_g=0.3, use_ris=ris)
    print(f"[case] event counters seed={seed} ris={ris}", flush=True)
    sc = ctx.scene(rows)
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    events = ctx.count_events(sc, ev, cam, p)
    sc.close()
    ev.close()
    _, k = oracle_render(rec, env, cam, p)
    for ours, theirs in ESTIMATOR_COUNTERS.items():
        want = p.width * p.height * p.spp if theirs is None else k[theirs]
        assert events[ours] == want, (ours, events[ours], want)


def test_normalize_accumulator(ctx):
    # vp_normalize_accumulator (device/kernels/normalize_accumulator.cu:16-35):
    # out = mean * scale per RGBA component, one rounding each
    import torch
    rng = np.random.default_rng(3)
    n = 100_003
    mean = rng.uniform(-3, 3, (n, 4)).astype(np.float32)
    scale = np.float32(0.37)
    print(f"[case] normalize n={n}", flush=True)
    d_in = torch.from_numpy(mean).cuda()
    d_out = torch.empty_like(d_in)
    ctx.normalize_accumulator(d_out.data_ptr(), d_in.data_ptr(), n, float(scale))
    torch.cuda.synchronize()
    assert_bitwise(d_out.cpu().numpy(), mean * scale, "normalize")
