"""Batched and served forms: every job's image is bit-identical to rendering
it alone, whatever it shares a launch with and whatever the launch shape."""
import time

import numpy as np
import pytest

from volprim_amd import scenes, volprim

pytestmark = pytest.mark.gpu


def _mix(ctx):
    envs = [ctx.env(scenes.constant_env()), ctx.env(scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0))]
    scs = [ctx.scene(scenes.random_scene(40, 21)), ctx.scene(scenes.random_scene(90, 22, albedo=0.95)),
           ctx.scene(scenes.c0_single_gaussian(albedo=0.9))]
    jobs = []
    for i, (w, h, spp) in enumerate([(16, 16, 3), (24, 8, 5), (9, 13, 2), (32, 32, 1), (20, 20, 4), (8, 8, 7)]):
        cam = volprim.camera_perspective(**scenes.default_camera_args(w, h, view=["negz", "posx", "diag"][i % 3]))
        p = volprim.default_params(width=w, height=h, spp=spp, seed=100 + i, phase_g=[0.0, 0.85, -0.3][i % 3],
                                   use_ris=int(i == 4), max_depth=32)
        jobs.append((scs[i % 3], envs[i % 2], cam, p))
    return jobs, scs, envs


def test_batched_jobs_equal_individual_renders(ctx):
    print("[case] batched mix of 6 jobs", flush=True)
    jobs, scs, envs = _mix(ctx)
    alone = [ctx.render(*j)[0] for j in jobs]
    batched, st = ctx.render_jobs(jobs)
    for a, b in zip(alone, batched):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # a different launch shape and slicing does not change any image
    batched2, _ = ctx.render_jobs(jobs, launch=volprim.LaunchConfig(64, 2, 0, 2, 8))
    for a, b in zip(alone, batched2):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # the wavefront pipeline over the same job table
    wf = volprim.LaunchConfig(256, 0, 0, 0, 1, volprim.PIPELINE_WAVEFRONT, 4096)
    batched3, _ = ctx.render_jobs(jobs, launch=wf)
    for a, b in zip(alone, batched3):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for o in scs + envs:
        o.close()


def test_batched_jobs_mixing_scenes_within_wavefronts(ctx):
    # 30 tiny jobs (48 paths each) over three scenes of different sizes, so
    # most wavefronts run paths of two scenes at once and share leaf work
    # between them
    print("[case] 30 tiny jobs over 3 scenes, mixed within wavefronts", flush=True)
    env = ctx.env(scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0))
    scs = [ctx.scene(scenes.random_scene(60, 41, albedo=0.9)), ctx.scene(scenes.random_scene(7, 42, smax=0.5)),
           ctx.scene(scenes.c0_single_gaussian(albedo=0.9))]
    jobs = []
    for i in range(30):
        cam = volprim.camera_perspective(**scenes.default_camera_args(4, 4, view=["negz", "posx", "diag"][i % 3]))
        p = volprim.default_params(width=4, height=4, spp=3, seed=200 + i, phase_g=0.3 * (i % 2),
                                   use_ris=int(i % 5 == 0), max_depth=24)
        jobs.append((scs[i % 3], env, cam, p))
    alone = [ctx.render(*j)[0] for j in jobs]
    wf = volprim.LaunchConfig(256, 0, 0, 0, 1, volprim.PIPELINE_WAVEFRONT, 4096)
    for launch in (None, wf):
        batched, _ = ctx.render_jobs(jobs, launch=launch)
        for k, (a, b) in enumerate(zip(alone, batched)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"job {k} ({'wavefront' if launch else 'megakernel'})"
    for o in scs + [env]:
        o.close()


@pytest.mark.parametrize("reserve_cus", [False, True])
def test_server_result_same_alone_and_in_mix(ctx, reserve_cus):
    print(f"[case] server alone vs mix, reserve_cus={reserve_cus}", flush=True)
    jobs, scs, envs = _mix(ctx)
    srv = ctx.server(max_inflight=4, mem_budget_bytes=1 << 20, reserve_cus=reserve_cus)
    t0, out0 = srv.submit(*jobs[0])
    srv.drain()
    lat0 = srv.poll(t0)
    assert lat0 is not None and lat0 > 0
    tickets = [srv.submit(*j) for j in jobs[1:3]] + [srv.submit(*jobs[0])] + [srv.submit(*j) for j in jobs[3:]]
    deadline = time.time() + 60
    pending = {t: o for t, o in tickets}
    lats = {}
    while pending and time.time() < deadline:
        for t in list(pending):
            lat = srv.poll(t)
            if lat is not None:
                lats[t] = lat
                pending.pop(t)
        time.sleep(0.001)
    assert not pending, "requests did not finish"
    mixed0 = tickets[2][1]
    assert np.array_equal(out0.view(np.uint32), mixed0.view(np.uint32))
    for (t, o), j in zip(tickets[:2] + tickets[3:], jobs[1:3] + jobs[3:]):
        ref = ctx.render(*j)[0]
        assert np.array_equal(o.view(np.uint32), ref.view(np.uint32))
    st = srv.stats()
    assert st["completed"] == len(jobs) + 1 and st["p99_ms"] >= st["p50_ms"] > 0 and st["launches"] >= 2
    srv.close()
    for o in scs + envs:
        o.close()


def test_autotune_picks_a_config_and_keeps_results(ctx):
    print("[case] autotune", flush=True)
    jobs, scs, envs = _mix(ctx)
    before = ctx.render(*jobs[1])[0]
    best = ctx.autotune(jobs)
    assert best.block in (64, 128, 256) and best.active_cap >= 16 and best.regen_min >= 1
    assert best.pipeline in (volprim.PIPELINE_MEGAKERNEL, volprim.PIPELINE_WAVEFRONT)
    after = ctx.render(*jobs[1])[0]
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
    for o in scs + envs:
        o.close()
