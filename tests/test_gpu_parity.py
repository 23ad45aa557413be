"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Integer and index work (Morton codes, order, overlap lists) and, thanks to
the deterministic FP32 math on both sides, the floating-point results too are
compared bit for bit.  Each case prints its name and sizes before touching
the GPU.
"""
import numpy as np
import pytest

import vpo
from volprim_amd import scenes, volprim

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitwise(gpu, ref, what):
    g, r = np.asarray(gpu, np.float32), np.asarray(ref, np.float32)
    assert g.shape == r.shape, what
    both_nan = np.isnan(g) & np.isnan(r)
    diff = (bits(g) != bits(r)) & ~both_nan
    if diff.any():
        i = np.flatnonzero(diff.reshape(-1))[:5]
        raise AssertionError(f"{what}: {int(diff.sum())} of {g.size} differ; first {i}: gpu={g.reshape(-1)[i]} "
                             f"ref={r.reshape(-1)[i]}")


MATH_CASES = {
    "exp": lambda rng, n: (rng.uniform(-110, 90, n), None),
    "log": lambda rng, n: (np.exp(rng.uniform(-103, 88, n)), None),
    "sin": lambda rng, n: (rng.uniform(-30, 30, n), None),
    "cos": lambda rng, n: (rng.uniform(-30, 30, n), None),
    "atan2": lambda rng, n: (rng.normal(size=n), rng.normal(size=n)),
    "acos": lambda rng, n: (rng.uniform(-1, 1, n), None),
    "erf": lambda rng, n: (rng.uniform(-6, 6, n), None),
    "erfc": lambda rng, n: (rng.uniform(-4, 11, n), None),
    "erfinv": lambda rng, n: (rng.uniform(-1, 1, n), None),
    "erf_diff": lambda rng, n: tuple(np.sort(rng.uniform(-6, 6, (2, n)), axis=0)),
    "sqrt": lambda rng, n: (np.exp(rng.uniform(-80, 80, n)), None),
    "div": lambda rng, n: (rng.normal(size=n), rng.normal(size=n)),
    "chord_poly": lambda rng, n: (rng.uniform(0, 9, n), None),
}
SPECIALS = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 1e-30, -1e-30, 1e-40, 88.7, -87.5, -103.9, 3.92, 10.1, 10.2,
                     0.9999999, -0.9999999, np.inf, -np.inf, np.nan], np.float32)


@pytest.mark.parametrize("name", sorted(MATH_CASES))
def test_math_bitexact(ctx, name):
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    x, y = MATH_CASES[name](rng, 200_000)
    x = np.concatenate([np.float32(x), SPECIALS]).astype(np.float32)
    y = np.concatenate([np.float32(y), np.flip(SPECIALS)]).astype(np.float32) if y is not None else None
    print(f"[case] math {name} n={x.size}", flush=True)
    assert_bitwise(ctx.math_eval(name, x, y), vpo.math_eval(name, x, y), f"math {name}")


def _random_rays(rng, m, spread=2.0):
    o = rng.uniform(-spread, spread, (m, 3))
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_prim_functions_bitexact(ctx, seed):
    rng = np.random.default_rng(seed)
    rows = scenes.random_scene(1, seed + 10)
    rec = vpo.make_prims(rows)[0]
    m = 50_000
    rays = _random_rays(rng, m, 1.0)
    t0 = rng.uniform(0, 2, m).astype(np.float32)
    t1 = (t0 + rng.uniform(0, 2, m)).astype(np.float32)
    tau = rng.exponential(1.0, m).astype(np.float32)
    args = np.concatenate([rays, t0[:, None], t1[:, None], tau[:, None]], 1).astype(np.float32)
    for fn in volprim.PRIM_FNS:
        print(f"[case] prim {fn} seed={seed} m={m}", flush=True)
        assert_bitwise(ctx.prim_eval(fn, rec, args), vpo.prim_eval(fn, rec, args), f"prim {fn}")


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (7, 2), (300, 3), (4096, 4)])
def test_scene_build_matches_oracle(ctx, n, seed):
    rows = scenes.random_scene(n, seed)
    if n == 4096:
        rows = scenes.c1_scattering()
    print(f"[case] scene build n={n}", flush=True)
    sc = ctx.scene(rows)
    order, morton, rec, depth = sc.export()
    sc.close()
    codes = vpo.morton_codes(rows[:, 0:3])
    assert np.array_equal(morton, codes), "Morton codes differ from the CPU restatement"
    assert np.array_equal(order, vpo.sort_order(codes)), "stable sort order differs"
    assert_bitwise(rec, vpo.make_prims(rows)[order], "primitive records")
    assert depth < 64


def test_scene_build_duplicate_codes(ctx):
    # collinear stack of identical Gaussians (SG_CLUSTER_MODE=stress): all Morton codes tie
    rows = np.repeat(scenes.c0_single_gaussian(albedo=0.5), 33, axis=0)
    print("[case] scene build duplicates n=33", flush=True)
    sc = ctx.scene(rows)
    order, morton, rec, depth = sc.export()
    sc.close()
    assert np.array_equal(order, np.arange(33, dtype=np.uint32))


@pytest.mark.parametrize("n,seed", [(1, 0), (60, 1), (500, 2)])
def test_overlap_lists_bitexact(ctx, n, seed):
    rng = np.random.default_rng(100 + seed)
    rows = scenes.random_scene(n, seed)
    sc = ctx.scene(rows)
    order, _, rec, _ = sc.export()
    rays = _random_rays(rng, 4000, 1.5)
    cap = 512
    print(f"[case] overlaps n={n} rays={rays.shape[0]}", flush=True)
    cg, tg, ig = ctx.collect_overlaps(sc, rays, cap)
    sc.close()
    co, to, io = vpo.overlaps(rec, rays, cap)
    assert np.array_equal(cg, co), "entry counts differ"
    assert cg.max() < cap
    for r in range(rays.shape[0]):
        k = int(cg[r])
        assert np.array_equal(ig[r, :k], io[r, :k]), f"ray {r}: indices differ"
        assert_bitwise(tg[r, :k], to[r, :k], f"ray {r} entry t")


@pytest.mark.parametrize("n,seed", [(1, 0), (60, 1), (500, 2)])
def test_transmittance_bitexact(ctx, n, seed):
    rng = np.random.default_rng(200 + seed)
    rows = scenes.random_scene(n, seed)
    sc = ctx.scene(rows)
    _, _, rec, _ = sc.export()
    rays = _random_rays(rng, 4000, 1.5)
    print(f"[case] transmittance n={n} rays={rays.shape[0]}", flush=True)
    Tg = ctx.transmittance_rays(sc, rays)
    sc.close()
    assert_bitwise(Tg, vpo.transmittance_rays(rec, rays), "transmittance")


def _render_both(ctx, rows, env, cam, **kw):
    sc = ctx.scene(rows)
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    p = volprim.default_params(**kw)
    img, st = ctx.render(sc, ev, cam, p)
    sc.close()
    ev.close()
    ref, _, k = vpo.render(rec, env, cam.pod(), p.width, p.height, p.spp, seed=p.seed, max_depth=p.max_depth,
                           rr_depth=p.rr_depth, rr_max_survival=p.rr_max_survival, firefly_clamp=p.firefly_clamp,
                           filter_stddev=p.filter_stddev, phase_g=p.phase_g, use_ris=bool(p.use_ris),
                           ris_candidates=p.ris_candidates, active_cap=p.active_cap, shadow_rr=p.shadow_rr)
    return img, ref, st, k


RENDER_CASES = {
    "c0_absorb": dict(rows=lambda: scenes.c0_single_gaussian(), env="const", w=32, h=32, spp=4, kw={}),
    "c0_scatter": dict(rows=lambda: scenes.c0_single_gaussian(albedo=0.9), env="const", w=24, h=24, spp=4, kw={}),
    "mis_sky": dict(rows=lambda: scenes.random_scene(48, 5), env="sky", w=24, h=20, spp=3, kw={}),
    "ris_sky": dict(rows=lambda: scenes.random_scene(48, 6), env="sky", w=20, h=20, spp=3,
                    kw=dict(use_ris=1, ris_candidates=4)),
    "iso_filter_clamp": dict(rows=lambda: scenes.random_scene(40, 7), env="sky", w=20, h=16, spp=3,
                             kw=dict(phase_g=0.0, filter_stddev=0.7, firefly_clamp=5.0, rr_depth=2)),
    "small_cap": dict(rows=lambda: scenes.random_scene(80, 8, smax=0.5), env="const", w=16, h=16, spp=2,
                      kw=dict(active_cap=2)),
    # shadow-ray roulette on (the build's option): the walks that stop at the
    # threshold must give the oracle's full-walk result bit for bit
    "mis_rr": dict(rows=lambda: scenes.random_scene(90, 21, smax=0.4, albedo=0.95), env="sky", w=24, h=20, spp=4,
                   kw=dict(shadow_rr=0.2, phase_g=0.3)),
    "ris_rr": dict(rows=lambda: scenes.random_scene(90, 22, smax=0.4, albedo=0.95), env="sky", w=20, h=20, spp=4,
                   kw=dict(shadow_rr=0.2, use_ris=1, ris_candidates=4)),
}


@pytest.mark.parametrize("case", sorted(RENDER_CASES))
def test_render_bitexact(ctx, case):
    c = RENDER_CASES[case]
    env = scenes.constant_env() if c["env"] == "const" else scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    cam = volprim.camera_perspective(**scenes.default_camera_args(c["w"], c["h"]))
    print(f"[case] render {case} {c['w']}x{c['h']} spp={c['spp']}", flush=True)
    img, ref, st, k = _render_both(ctx, c["rows"](), env, cam, width=c["w"], height=c["h"], spp=c["spp"], **c["kw"])
    assert_bitwise(img, ref, f"render {case}")
    assert st["active_overflow"] == k["active_overflow"]


@pytest.mark.parametrize("case", sorted(RENDER_CASES))
def test_render_wavefront_bitexact(ctx, case):
    # the wavefront-queue pipeline (pool of 4096 path slots, refilled every
    # iteration) writes the same per-sample radiance as the megakernel
    c = RENDER_CASES[case]
    env = scenes.constant_env() if c["env"] == "const" else scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    cam = volprim.camera_perspective(**scenes.default_camera_args(c["w"], c["h"]))
    print(f"[case] wavefront render {case} {c['w']}x{c['h']} spp={c['spp']}", flush=True)
    rows = c["rows"]()
    sc = ctx.scene(rows)
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    p = volprim.default_params(width=c["w"], height=c["h"], spp=c["spp"], **c["kw"])
    launch = volprim.LaunchConfig(256, 0, 0, 0, 1, volprim.PIPELINE_WAVEFRONT, 4096)
    imgs, st = ctx.render_jobs([(sc, ev, cam, p)], launch=launch)
    sc.close()
    ev.close()
    ref, _, k = vpo.render(rec, env, cam.pod(), p.width, p.height, p.spp, seed=p.seed, max_depth=p.max_depth,
                           rr_depth=p.rr_depth, rr_max_survival=p.rr_max_survival, firefly_clamp=p.firefly_clamp,
                           filter_stddev=p.filter_stddev, phase_g=p.phase_g, use_ris=bool(p.use_ris),
                           ris_candidates=p.ris_candidates, active_cap=p.active_cap, shadow_rr=p.shadow_rr)
    assert_bitwise(imgs[0], ref, f"wavefront render {case}")
    assert st["active_overflow"] == k["active_overflow"]


def test_render_orthographic_bitexact(ctx):
    rows = scenes.random_scene(30, 9)
    cam = volprim.camera_orthographic(20, 20, np.array([0, 0, -5], np.float32), np.zeros(3, np.float32),
                                      np.array([0, 1, 0], np.float32), 3.0)
    print("[case] render ortho 20x20", flush=True)
    img, ref, _, _ = _render_both(ctx, rows, scenes.constant_env(), cam, width=20, height=20, spp=2)
    assert_bitwise(img, ref, "render ortho")


def test_furnace_gpu(ctx):
    # albedo 1 medium in a constant white environment reproduces the field exactly:
    # |mean - 1| < max(5 SEM, 2e-3)  (experiments/mitsuba-reference/furnace_check.py:36-40)
    rows = scenes.random_scene(64, 11, albedo=1.0)
    rows[:, 10:13] = 1.0
    sc = ctx.scene(rows)
    ev = ctx.env(scenes.constant_env())
    cam = volprim.camera_perspective(**scenes.default_camera_args(64, 64))
    p = volprim.default_params(width=64, height=64, spp=64, max_depth=512, rr_depth=512, phase_g=0.3)
    print("[case] furnace 64x64 spp=64", flush=True)
    img, _ = ctx.render(sc, ev, cam, p)
    sc.close()
    ev.close()
    m = img[..., 0:3].mean()
    sem = img[..., 0:3].std() / np.sqrt(img[..., 0:3].size)
    assert abs(m - 1.0) < max(5 * sem, 2e-3), (m, sem)


@pytest.mark.parametrize("case", ["a", "b"])
def test_gpu_transmittance_vs_fp64_golden(ctx, case):
    import os
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_fp64.npz"))
    rows, rays, Tg = G[f"rows_{case}"], G[f"rays_{case}"], G[f"T_{case}"]
    print(f"[case] golden transmittance {case} n={rows.shape[0]} rays={rays.shape[0]}", flush=True)
    sc = ctx.scene(rows)
    T = ctx.transmittance_rays(sc, rays).astype(np.float64)
    sc.close()
    mask = Tg > 1e-30
    assert (np.abs(T - Tg) / Tg)[mask].max() < 1e-5


# ---- the optional host trees (SAH, median splits): same layout, same results for their stored order ----
@pytest.mark.parametrize("bvh", ["sah", "median"])
@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (5, 2), (300, 3), (4096, 4)])
def test_sah_scene_is_a_permutation_with_exact_records(ctx, bvh, n, seed):
    rows = scenes.c1_scattering() if n == 4096 else scenes.random_scene(n, seed)
    print(f"[case] {bvh} scene build n={n}", flush=True)
    sc = ctx.scene(rows, bvh=bvh)
    order, morton, rec, depth = sc.export()
    sc.close()
    assert np.array_equal(np.sort(order), np.arange(n, dtype=np.uint32))
    assert np.array_equal(morton, vpo.morton_codes(rows[:, 0:3]))
    assert_bitwise(rec, vpo.make_prims(rows)[order], "primitive records")


def _three_sigma_boxes(rows):
    # axis-aligned 3-sigma boxes of the rows' Gaussians (half extent along
    # world axis a: 3 sqrt(sum_j (R_aj s_j)^2)), in float64
    c = rows[:, 0:3].astype(np.float64)
    w, x, y, z = (rows[:, 3:7].astype(np.float64) / np.linalg.norm(rows[:, 3:7], axis=1, keepdims=True)).T
    s = rows[:, 7:10].astype(np.float64)
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
                  np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
                  np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)], 1)
    h = 3.0 * np.sqrt(((R * R) * (s * s)[:, None, :]).sum(2))
    return c - h, c + h


@pytest.mark.parametrize("flags,msg", [(volprim.SCENE_BVH_SAH | volprim.SCENE_BVH_MEDIAN, "exclude each other"),
                                       (volprim.SCENE_WALK_F32 | volprim.SCENE_WALK_F16, "exclude each other"),
                                       (16, "unknown scene flags")])
def test_scene_flags_rejected(ctx, flags, msg):
    import ctypes
    rows = scenes.random_scene(8, 1)
    prims = volprim.primitives_from_rows(volprim.f32(rows))
    h = ctypes.c_void_p()
    print(f"[case] scene flags {flags} rejected", flush=True)
    rc = volprim.load().vp_scene_create_ex(ctx.handle, ctypes.cast(prims, ctypes.c_void_p), 8, flags,
                                           ctypes.byref(h))
    assert rc != volprim.VP_OK and not h.value
    assert msg in volprim.load().vp_last_error().decode()


@pytest.mark.parametrize("n", [300, 1000])
def test_median_order_splits_on_record_boundaries(ctx, n):
    # every aligned group of 2^k stored leaves whose second half is not empty
    # is split in two by centre along some axis, ties going by input index:
    # the halves (and so the walk tree's 4-wide records, which group 4^k
    # aligned leaves) are spatial splits
    rows = scenes.random_scene(n, 21)
    print(f"[case] median order structure n={n}", flush=True)
    sc = ctx.scene(rows, bvh="median")
    order, _, _, _ = sc.export()
    sc.close()
    cen = rows[:, 0:3].astype(np.float32)
    size = 2
    while size // 2 < n:
        for start in range(0, n, size):
            left = order[start:start + size // 2]
            right = order[start + size // 2:start + size]
            if len(right) == 0:
                continue
            ok = False
            for ax in range(3):
                lk = sorted(zip(cen[left, ax].tolist(), left.tolist()))
                rk = sorted(zip(cen[right, ax].tolist(), right.tolist()))
                ok |= lk[-1] < rk[0]
            assert ok, (size, start)
        size *= 2


def test_median_order_has_smaller_walk_tree_surface_than_sah(ctx):
    # the surface-area cost of the 4-wide walk tree (sum over records of the
    # record box's surface relative to the root's) on the 4096-Gaussian
    # scattering scene: the median order's must be below the SAH order's
    rows = scenes.c1_scattering(n=4096, seed=1, sigma_multiplier=10.0, albedo=0.9)
    print("[case] median vs SAH walk-tree surface, 4096 Gaussians", flush=True)

    def cost(bvh):
        sc = ctx.scene(rows, bvh=bvh)
        order, _, _, _ = sc.export()
        sc.close()
        lo, hi = _three_sigma_boxes(rows)
        lo, hi = lo[order], hi[order]
        area = lambda a, b: 2.0 * ((b - a)[..., 0] * (b - a)[..., 1] + (b - a)[..., 1] * (b - a)[..., 2] +
                                   (b - a)[..., 2] * (b - a)[..., 0])
        total, root = 0.0, area(lo.min(0), hi.max(0))
        while lo.shape[0] > 4:
            lo, hi = lo.reshape(-1, 4, 3).min(1), hi.reshape(-1, 4, 3).max(1)
            total += area(lo, hi).sum() / root
        return total

    c_sah, c_med = cost("sah"), cost("median")
    print(f"walk-tree surface: sah {c_sah:.1f}, median {c_med:.1f}", flush=True)
    assert c_med < 0.9 * c_sah


@pytest.mark.parametrize("bvh", ["sah", "median"])
def test_sah_overlaps_and_transmittance_bitexact(ctx, bvh):
    rng = np.random.default_rng(77)
    rows = scenes.random_scene(400, 12)
    sc = ctx.scene(rows, bvh=bvh)
    _, _, rec, _ = sc.export()
    rays = _random_rays(rng, 3000, 1.5)
    print(f"[case] {bvh} overlaps + transmittance n=400", flush=True)
    cg, tg, ig = ctx.collect_overlaps(sc, rays, 512)
    T = ctx.transmittance_rays(sc, rays)
    sc.close()
    co, to, io = vpo.overlaps(rec, rays, 512)
    assert np.array_equal(cg, co)
    for r in range(rays.shape[0]):
        k = int(cg[r])
        assert np.array_equal(ig[r, :k], io[r, :k]) and np.array_equal(tg[r, :k], to[r, :k])
    assert_bitwise(T, vpo.transmittance_rays(rec, rays), "transmittance")


@pytest.mark.parametrize("bvh", ["sah", "median"])
@pytest.mark.parametrize("case", ["mis_sky", "small_cap", "ris_sky"])
def test_render_sah_bitexact(ctx, case, bvh):
    c = RENDER_CASES[case]
    env = scenes.constant_env() if c["env"] == "const" else scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    cam = volprim.camera_perspective(**scenes.default_camera_args(c["w"], c["h"]))
    print(f"[case] {bvh} render {case}", flush=True)
    sc = ctx.scene(c["rows"](), bvh=bvh)
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    p = volprim.default_params(width=c["w"], height=c["h"], spp=c["spp"], **c["kw"])
    img, st = ctx.render(sc, ev, cam, p)
    sc.close()
    ev.close()
    ref, _, k = vpo.render(rec, env, cam.pod(), p.width, p.height, p.spp, seed=p.seed, max_depth=p.max_depth,
                           rr_depth=p.rr_depth, rr_max_survival=p.rr_max_survival, firefly_clamp=p.firefly_clamp,
                           filter_stddev=p.filter_stddev, phase_g=p.phase_g, use_ris=bool(p.use_ris),
                           ris_candidates=p.ris_candidates, active_cap=p.active_cap, shadow_rr=p.shadow_rr)
    assert_bitwise(img, ref, f"{bvh} render {case}")
    assert st["active_overflow"] == k["active_overflow"]


def _oracle_render(rec, env, cam, p):
    return vpo.render(rec, env, cam.pod(), p.width, p.height, p.spp, seed=p.seed, max_depth=p.max_depth,
                      rr_depth=p.rr_depth, rr_max_survival=p.rr_max_survival, firefly_clamp=p.firefly_clamp,
                      filter_stddev=p.filter_stddev, phase_g=p.phase_g, use_ris=bool(p.use_ris),
                      ris_candidates=p.ris_candidates, active_cap=p.active_cap, shadow_rr=p.shadow_rr)


@pytest.mark.parametrize("walk", ["f16", "f32"])
@pytest.mark.parametrize("case", ["mis_rr", "ris_sky", "small_cap", "c0_scatter"])
def test_render_walk_formats_bitexact(ctx, case, walk):
    # the walk tree in binary16 records (bounds rounded outward) and in
    # binary32 records: both pipelines give the oracle's image bit for bit,
    # and the unit-sized test scenes pick the binary16 form by themselves
    c = RENDER_CASES[case]
    env = scenes.constant_env() if c["env"] == "const" else scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    cam = volprim.camera_perspective(**scenes.default_camera_args(c["w"], c["h"]))
    print(f"[case] walk {walk} render {case}", flush=True)
    rows = c["rows"]()
    auto = ctx.scene(rows)
    assert auto.walk_info()[0], "a unit-sized scene should take the binary16 records"
    growth = auto.walk_info()[1]
    auto.close()
    assert 0.0 <= growth < 1.0 / 16.0
    sc = ctx.scene(rows, walk=walk)
    assert sc.walk_info()[0] == (walk == "f16")
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    p = volprim.default_params(width=c["w"], height=c["h"], spp=c["spp"], **c["kw"])
    img, st = ctx.render(sc, ev, cam, p)
    launch = volprim.LaunchConfig(256, 0, 0, 0, 1, volprim.PIPELINE_WAVEFRONT, 4096)
    imgs_wf, _ = ctx.render_jobs([(sc, ev, cam, p)], launch=launch)
    sc.close()
    ev.close()
    ref, _, k = _oracle_render(rec, env, cam, p)
    assert_bitwise(img, ref, f"walk {walk} render {case}")
    assert_bitwise(imgs_wf[0], ref, f"walk {walk} wavefront render {case}")
    assert st["active_overflow"] == k["active_overflow"]


@pytest.mark.parametrize("walk", ["f16", "f32"])
def test_walk_formats_overlaps_and_transmittance_bitexact(ctx, walk):
    rng = np.random.default_rng(91)
    rows = scenes.random_scene(500, 31)
    sc = ctx.scene(rows, walk=walk)
    _, _, rec, _ = sc.export()
    rays = _random_rays(rng, 3000, 1.5)
    print(f"[case] walk {walk} overlaps + transmittance n=500", flush=True)
    cg, tg, ig = ctx.collect_overlaps(sc, rays, 512)
    T = ctx.transmittance_rays(sc, rays)
    sc.close()
    co, to, io = vpo.overlaps(rec, rays, 512)
    assert np.array_equal(cg, co)
    for r in range(rays.shape[0]):
        k = int(cg[r])
        assert np.array_equal(ig[r, :k], io[r, :k]) and np.array_equal(tg[r, :k], to[r, :k])
    assert_bitwise(T, vpo.transmittance_rays(rec, rays), "transmittance")


@pytest.mark.parametrize("walk", ["auto", "f16"])
def test_far_scene_walk_format(ctx, walk):
    # a scene 3000 units from the origin: binary16 bounds there are 2 apart,
    # so the automatic choice keeps binary32 records; forced binary16 records
    # (boxes grown to the binary16 grid, so many more leaf tests) still give
    # the oracle's image bit for bit
    off = np.array([3000.0, -1000.0, 500.0], np.float32)
    rows = scenes.random_scene(40, 33)
    rows[:, 0:3] += off
    env = scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    cam = volprim.camera_perspective(16, 16, off + np.array([0, 0, -3.5], np.float32), off,
                                     np.array([0, 1, 0], np.float32), 40.0)
    print(f"[case] far scene walk {walk}", flush=True)
    sc = ctx.scene(rows, walk=walk)
    half, growth = sc.walk_info()
    assert growth > 1.0 / 16.0
    assert half == (walk == "f16")
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    p = volprim.default_params(width=16, height=16, spp=2)
    img, st = ctx.render(sc, ev, cam, p)
    sc.close()
    ev.close()
    ref, _, k = _oracle_render(rec, env, cam, p)
    assert_bitwise(img, ref, f"far scene walk {walk}")
    assert img[..., :3].mean() > 0.0


def test_full_size_scattering_workload(ctx):
    # the measured workload at its full size: 4096 Gaussians (seed 1,
    # sigma multiplier 10, albedo 0.9, isotropic phase), 2048x1024 sky + sun,
    # 512x512 at 64 spp, depth 64, RR from 12, median-split tree.  The megakernel and
    # the wavefront pipeline must agree on every pixel; every 32nd row and
    # column is checked against the oracle (256 pixels x 64 samples).
    import os
    rows = scenes.c1_scattering(n=4096, seed=1, sigma_multiplier=10.0, albedo=0.9)
    env = scenes.sky_env(2048, 1024)
    cam = volprim.camera_perspective(**scenes.default_camera_args(512, 512, dist=3.5, fov=40.0))
    p = volprim.default_params(width=512, height=512, spp=64, seed=0, max_depth=64, rr_depth=12, phase_g=0.0)
    print("[case] full-size scattering 4096 Gaussians 512x512x64 megakernel + wavefront", flush=True)
    sc = ctx.scene(rows, bvh="median")
    ev = ctx.env(env)
    _, _, rec, _ = sc.export()
    img, st = ctx.render(sc, ev, cam, p)
    wf = volprim.LaunchConfig(256, 0, 0, 0, 0, volprim.PIPELINE_WAVEFRONT, 1 << 20)
    imgs_wf, st_wf = ctx.render_jobs([(sc, ev, cam, p)], launch=wf)
    sc.close()
    ev.close()
    assert st["active_overflow"] == 0 and st_wf["active_overflow"] == 0
    assert_bitwise(imgs_wf[0], img, "wavefront vs megakernel, full image")
    s = 32
    print(f"[case] oracle on every {s}th row and column", flush=True)
    ref, _, k = vpo.render(rec, env, cam.pod(), 512, 512, 64, seed=0, max_depth=64, rr_depth=12, phase_g=0.0,
                           stride=s, threads=min(os.cpu_count() or 1, 16), bvh=True)
    assert_bitwise(img[::s, ::s], ref[::s, ::s], "full-size render vs oracle (strided pixels)")
    assert np.isfinite(img).all() and img[..., :3].mean() > 0.0
