"""Pin the CPU oracle against what the mathematics fixes, independently of its
own code: FP64 golden vectors (quadrature / closed forms, tests/golden),
mpmath, published known-answer vectors, distribution identities, and
estimator invariants (furnace, closed-form absorption image)."""
import math
import os

import mpmath as mp
import numpy as np
import pytest
from scipy import integrate

import vpo
from volprim_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden_fp64.npz"))
mp.mp.dps = 30


# ---------------------------------------------------------------- math ----
def max_ulp(name, xs, ref_fn, ys=None):
    xs = np.asarray(xs, np.float32)
    got = vpo.math_eval(name, xs, ys).astype(np.float64)
    if ys is None:
        ref = np.array([float(ref_fn(mp.mpf(float(v)))) for v in xs])
    else:
        ref = np.array([float(ref_fn(mp.mpf(float(a)), mp.mpf(float(b)))) for a, b in zip(xs, ys)])
    ulp = np.maximum(np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64),
                     float(np.finfo(np.float32).smallest_subnormal))
    return float(np.max(np.abs(got - ref) / ulp))


@pytest.mark.parametrize("name,gen,ref,bound", [
    ("exp", lambda r: r.uniform(-103, 88.7, 3000), mp.exp, 1.0),
    ("log", lambda r: np.exp(r.uniform(-100, 88, 3000)), mp.log, 1.0),
    ("sin", lambda r: r.uniform(-20, 20, 3000), mp.sin, 2.0),
    ("cos", lambda r: r.uniform(-20, 20, 3000), mp.cos, 2.0),
    ("acos", lambda r: r.uniform(-1, 1, 3000), mp.acos, 3.0),
    ("erf", lambda r: r.uniform(-5, 5, 3000), mp.erf, 3.0),
    ("erfc", lambda r: r.uniform(-3, 10, 3000), mp.erfc, 8.0),
    ("erfinv", lambda r: r.uniform(-0.9999, 0.9999, 3000), mp.erfinv, 4.0),
])
def test_math_ulp_bounds(name, gen, ref, bound):
    assert max_ulp(name, gen(np.random.default_rng(5)), ref) <= bound


def test_atan2_ulp_bound():
    r = np.random.default_rng(6)
    y, x = r.normal(size=3000).astype(np.float32), r.normal(size=3000).astype(np.float32)
    assert max_ulp("atan2", y, mp.atan2, x) <= 2.0


def test_erf_diff_stable_on_tails():
    # erf(b) - erf(a) with both arguments deep in a tail keeps relative accuracy
    a = np.float32(np.linspace(2.0, 6.0, 200))
    b = (a + np.float32(0.01)).astype(np.float32)
    got = vpo.math_eval("erf_diff", a, b).astype(np.float64)
    ref = np.array([float(mp.erfc(mp.mpf(float(x))) - mp.erfc(mp.mpf(float(y)))) for x, y in zip(a, b)])
    assert np.max(np.abs(got - ref) / ref) < 5e-6


# ------------------------------------------------------- integer pieces ----
def test_morton_known_answers():
    # host/utils/math.h:100-102 documents expandBits(0b1111) = 0b1001001001
    assert vpo.lib().vpo_expand_bits(0b1111) == 0b1001001001
    assert vpo.lib().vpo_expand_bits(1023) == 0x09249249
    c = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.25, 1.0]], np.float32)
    codes = vpo.morton_codes(c)
    assert codes[0] == 0  # morton3D(scene_min) = 0
    # interleave z|y|x: (x=511, y=255, z=1023) from (0.5*1023, 0.25*1023, 1023)
    def interleave(x, y, z):
        out = 0
        for b in range(10):
            out |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
        return out
    assert codes[2] == interleave(int(0.5 * 1023.0), int(0.25 * 1023.0), 1023)


def test_sort_order_is_stable():
    codes = np.array([5, 3, 5, 1, 3, 5], np.uint32)
    assert vpo.sort_order(codes).tolist() == [3, 1, 4, 0, 2, 5]


def test_pcg32_known_answer_vector():
    # O'Neill's pcg32 minimal C demo, pcg32_srandom_r(&rng, 42u, 54u): the
    # reference's random::init performs exactly that seeding (random.cuh:41-49)
    out = vpo.pcg_stream(42, 54, 6)
    assert [hex(v) for v in out] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b",
                                     "0xcbed606e"]


def test_prim_uniform_is_uniform_and_order_free():
    lib = vpo.lib()
    u = np.array([lib.vpo_prim_uniform(0x1234, b, r, p) for b in range(4) for r in range(2) for p in range(2000)])
    assert 0.0 <= u.min() and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    assert len(set(np.round(u, 7))) > 0.99 * u.size  # distinct per (bounce, role, prim)


# ------------------------------------------------ primitive arithmetic ----
def _rays(rng, m, spread):
    o = rng.uniform(-spread, spread, (m, 3))
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


@pytest.mark.parametrize("case", ["a", "b"])
def test_transmittance_matches_fp64_golden(case):
    rows = G[f"rows_{case}"]
    order = vpo.scene_order(rows)
    rec = vpo.make_prims(rows)[order]
    T = vpo.transmittance_rays(rec, G[f"rays_{case}"]).astype(np.float64)
    Tg = G[f"T_{case}"]
    rel = np.abs(T - Tg) / np.maximum(Tg, 1e-30)
    # 1e-5 relative where exp(-tau) is representable well above float32's floor
    mask = Tg > 1e-30
    assert rel[mask].max() < 1e-5, rel[mask].max()


def test_single_gaussian_closed_form():
    rec = vpo.make_prims(G["c0_rows"])[0]
    d = G["c0_d"]
    rays = np.zeros((d.size, 6), np.float32)
    rays[:, 0] = d
    rays[:, 2] = -3.5
    rays[:, 5] = 1.0
    T = vpo.transmittance_rays(rec[None, :], rays).astype(np.float64)
    tau = -np.log(T)
    assert np.allclose(tau, G["c0_tau"], rtol=1e-5, atol=2e-7)
    assert abs(tau[0] - 2.0) < 2e-6  # central optical depth of the configuration


def test_optical_depth_matches_quadrature():
    rng = np.random.default_rng(9)
    rows = scenes.random_scene(1, 31, smin=0.1, smax=0.6)
    rec = vpo.make_prims(rows)[0]
    c = rows[0, 0:3].astype(np.float64)
    args = []
    for _ in range(300):
        o = c + rng.normal(size=3) * 0.8
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        args.append([*o, *d, 0.0, 0.0, 0.0])
    args = np.array(args, np.float32)
    # chord of each ray inside the bound, from the oracle's own exits
    ex_in = vpo.prim_eval("exit_inside", rec, args)
    t_en = vpo.prim_eval("entry_t", rec, args)
    ex_en = vpo.prim_eval("exit_entry", rec, args)
    inside = vpo.prim_eval("point_inside", rec, args) > 0
    t0 = np.where(inside, 0.0, t_en)
    t1 = np.where(inside, ex_in, ex_en)
    ok = inside | (t_en >= 0)
    args[:, 6], args[:, 7] = t0, t1
    od = vpo.prim_eval("optical_depth", rec, args)
    from golden.make_golden import density_along  # FP64 density along the ray
    for i in np.flatnonzero(ok)[:120]:
        q, _ = integrate.quad(lambda t: density_along(rows[0], args[i, :3].astype(np.float64),
                                                      args[i, 3:6].astype(np.float64), t),
                              float(t0[i]), float(t1[i]), epsrel=1e-12, epsabs=0)
        assert abs(od[i] - q) <= 2e-5 * q + 1e-7, (i, od[i], q)


def test_chord_poly_matches_mpmath():
    # S(d) = sqrt(pi/2) e^(d/2) erf(sqrt(d/2)) / sqrt(d) = sum d^n / (2n+1)!!
    # on [0, 9], the whole-chord depth factor: relative error below 5e-7
    d = np.concatenate([np.linspace(0.0, 9.0, 1001), np.random.default_rng(4).uniform(0.0, 9.0, 1000)])
    d = d.astype(np.float32)
    got = vpo.math_eval("chord_poly", d).astype(np.float64)

    def S(x):
        x = mp.mpf(float(x))
        return mp.mpf(1) if x == 0 else mp.sqrt(mp.pi / 2) * mp.exp(x / 2) * mp.erf(mp.sqrt(x / 2)) / mp.sqrt(x)
    ref = np.array([float(S(x)) for x in d])
    assert np.max(np.abs(got - ref) / ref) < 5e-7
    # the series' first terms: S(0) = 1, S'(0) = 1/3
    assert abs(got[0] - 1.0) < 3e-7
    h = np.float32(1e-2)
    s1 = vpo.math_eval("chord_poly", np.array([h], np.float32))[0]
    assert abs((s1 - 1.0) / h - (1.0 / 3.0 + h / 15.0)) < 1e-3


def test_chord_depth_matches_quadrature_and_erf_form():
    # The whole-chord optical depth of an entered primitive (one transform,
    # S(delta), no exp/erf) against FP64 quadrature of M N(x) between the
    # FP64 entry and exit, and against the erf-difference optical_depth over
    # the oracle's own entry and exit.  A dropped factor (2, e^(-9/2), M or
    # dnf) or delta confused with the closest distance fails both.
    rng = np.random.default_rng(19)
    from golden.make_golden import chord, density_along
    for seed in (31, 32, 33):
        rows = scenes.random_scene(1, seed, smin=0.1, smax=0.6)
        rec = vpo.make_prims(rows)[0]
        c = rows[0, 0:3].astype(np.float64)
        args = []
        for _ in range(200):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            o = c - 4.0 * d + rng.normal(size=3) * 0.3
            args.append([*o, *d, 0.0, 0.0, 0.0])
        args = np.array(args, np.float32)
        tau = vpo.prim_eval("chord_tau", rec, args).astype(np.float64)
        t_en = vpo.prim_eval("entry_t", rec, args)
        t_ex = vpo.prim_eval("exit_entry", rec, args)
        args2 = args.copy()
        args2[:, 6], args2[:, 7] = t_en, t_ex
        od = vpo.prim_eval("optical_depth", rec, args2).astype(np.float64)
        n = 0
        for i in range(args.shape[0]):
            ch, org_in = chord(rows[0], args[i, :3].astype(np.float64), args[i, 3:6].astype(np.float64))
            if org_in or ch is None:
                assert tau[i] == 0.0
                continue
            q, _ = integrate.quad(lambda t: density_along(rows[0], args[i, :3].astype(np.float64),
                                                          args[i, 3:6].astype(np.float64), t),
                                  ch[0], ch[1], epsrel=1e-12, epsabs=0)
            assert abs(tau[i] - q) <= 2e-5 * q + 1e-7, (seed, i, tau[i], q)
            assert abs(tau[i] - od[i]) <= 2e-5 * q + 1e-7, (seed, i, tau[i], od[i])
            n += 1
        assert n > 100


def test_inside_chord_matches_quadrature_and_erf_form():
    # Optical depth from an origin inside the 3-sigma bound to the exit (one
    # transform, one exp, two S polynomials) against FP64 quadrature of M N(x)
    # over [0, FP64 exit], and against the erf-difference optical_depth over
    # the oracle's own [0, exit_inside].  Origins near the centre, near the
    # bound, moving inwards and outwards.
    rng = np.random.default_rng(29)
    from golden.make_golden import chord, density_along
    n = 0
    for seed in (41, 42, 43):
        rows = scenes.random_scene(1, seed, smin=0.1, smax=0.6)
        rec = vpo.make_prims(rows)[0]
        c = rows[0, 0:3].astype(np.float64)
        s_max = float(np.max(rows[0, 7:10]))
        args = []
        for _ in range(150):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            o = c + rng.normal(size=3) * s_max * rng.uniform(0.1, 1.5)
            args.append([*o, *d, 0.0, 0.0, 0.0])
        args = np.array(args, np.float32)
        tau = vpo.prim_eval("inside_tau", rec, args).astype(np.float64)
        inside = vpo.prim_eval("point_inside", rec, args) > 0
        ex = vpo.prim_eval("exit_inside", rec, args)
        a2 = args.copy()
        a2[:, 6], a2[:, 7] = 0.0, ex
        od = vpo.prim_eval("optical_depth", rec, a2).astype(np.float64)
        for i in np.flatnonzero(inside):
            ch, org_in = chord(rows[0], args[i, :3].astype(np.float64), args[i, 3:6].astype(np.float64))
            if not org_in:
                continue
            dens = lambda t: density_along(rows[0], args[i, :3].astype(np.float64), args[i, 3:6].astype(np.float64), t)
            q, _ = integrate.quad(dens, 0.0, ch[1], epsrel=1e-12, epsabs=0)
            # the whole chord's depth scales the absolute bound: near the exit
            # the two terms of the closed form cancel to a few ulp of it
            from golden.make_golden import whiten
            pw, ww = whiten(rows[0], args[i, :3].astype(np.float64), args[i, 3:6].astype(np.float64))
            a_, b_, c_ = ww @ ww, pw @ ww, pw @ pw - 9.0
            t_back = (-b_ - math.sqrt(b_ * b_ - a_ * c_)) / a_  # the bound behind the origin
            full, _ = integrate.quad(dens, t_back, ch[1], epsrel=1e-12, epsabs=0)
            assert abs(tau[i] - q) <= 2e-5 * q + 2e-6 * full + 1e-7, (seed, i, tau[i], q)
            assert abs(tau[i] - od[i]) <= 2e-5 * q + 2e-6 * full + 1e-7, (seed, i, tau[i], od[i])
            n += 1
    assert n > 150


def test_origin_flight_inverts_the_inside_depth():
    # origin_flight(tau) lands where the depth from the origin equals tau
    # when tau is at most the depth to the exit, and passes (1e20) otherwise
    rng = np.random.default_rng(31)
    rows = scenes.random_scene(1, 47, smin=0.2, smax=0.5)
    rows[:, 13] = 5.0
    rec = vpo.make_prims(rows)[0]
    c = rows[0, 0:3].astype(np.float64)
    args = []
    for _ in range(600):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o = c + rng.normal(size=3) * 0.15
        args.append([*o, *d, 0.0, 0.0, 0.0])
    args = np.array(args, np.float32)
    inside = vpo.prim_eval("point_inside", rec, args) > 0
    full = vpo.prim_eval("inside_tau", rec, args)
    ex = vpo.prim_eval("exit_inside", rec, args)
    args[:, 8] = (rng.uniform(0.02, 1.3, args.shape[0]) * full).astype(np.float32)
    t = vpo.prim_eval("origin_flight", rec, args)
    hit = inside & (args[:, 8] <= full)
    assert hit.sum() > 300 and (inside & ~hit).sum() > 50
    assert np.all(t[inside & ~hit] >= 1e20)
    assert np.all((t[hit] >= 0.0) & (t[hit] <= ex[hit]))
    a2 = args.copy()
    a2[:, 6], a2[:, 7] = 0.0, t
    od = vpo.prim_eval("optical_depth", rec, a2)
    sel = hit & (args[:, 8] > 1e-3)
    assert np.allclose(od[sel], args[sel, 8], rtol=3e-4, atol=3e-6)


def test_entered_flight_inverts_the_chord_depth():
    # entered_flight(tau) lands where the optical depth from the entry equals
    # tau when tau is at most the chord's depth, and passes (1e20) otherwise
    rng = np.random.default_rng(23)
    rows = scenes.random_scene(1, 37, smin=0.2, smax=0.5)
    rows[:, 13] = 5.0  # mass: chords of depth ~0.5-5
    rec = vpo.make_prims(rows)[0]
    c = rows[0, 0:3].astype(np.float64)
    args = []
    for _ in range(600):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o = c - 4.0 * d + rng.normal(size=3) * 0.2
        args.append([*o, *d, 0.0, 0.0, 0.0])
    args = np.array(args, np.float32)
    full = vpo.prim_eval("chord_tau", rec, args)
    t_en = vpo.prim_eval("entry_t", rec, args)
    t_ex = vpo.prim_eval("exit_entry", rec, args)
    hit = t_en >= 0
    args[:, 8] = (rng.uniform(0.02, 1.3, args.shape[0]) * full).astype(np.float32)
    t = vpo.prim_eval("entered_flight", rec, args)
    inside = hit & (args[:, 8] <= full)
    assert inside.sum() > 300 and (hit & ~inside).sum() > 50
    assert np.all(t[hit & ~inside] >= 1e20)
    assert np.all((t[inside] >= t_en[inside]) & (t[inside] <= t_ex[inside]))
    a2 = args.copy()
    a2[:, 6], a2[:, 7] = t_en, t
    od = vpo.prim_eval("optical_depth", rec, a2)
    sel = inside & (args[:, 8] > 1e-3)
    assert np.allclose(od[sel], args[sel, 8], rtol=3e-4, atol=3e-6)


def test_exits_and_entry_match_fp64_quadratic():
    rng = np.random.default_rng(10)
    rows = scenes.random_scene(1, 41, smin=0.1, smax=0.6)
    rec = vpo.make_prims(rows)[0]
    from golden.make_golden import chord
    rays = _rays(rng, 2000, 2.0)
    args = np.concatenate([rays, np.zeros((2000, 3), np.float32)], 1)
    t_en = vpo.prim_eval("entry_t", rec, args)
    ex_en = vpo.prim_eval("exit_entry", rec, args)
    ex_in = vpo.prim_eval("exit_inside", rec, args)
    inside = vpo.prim_eval("point_inside", rec, args) > 0
    n_hit = 0
    for i in range(2000):
        ch, org_in = chord(rows[0], rays[i, :3].astype(np.float64), rays[i, 3:].astype(np.float64))
        if org_in:
            assert inside[i]
            assert abs(ex_in[i] - ch[1]) <= 1e-5 * max(1.0, ch[1])
        elif ch is None:
            assert t_en[i] < 0
        else:
            n_hit += 1
            assert abs(t_en[i] - ch[0]) <= 2e-5 * max(1.0, ch[0])
            assert abs(ex_en[i] - ch[1]) <= 2e-5 * max(1.0, ch[1])
    assert n_hit > 100


def test_inverse_cdf_matches_fp64_and_round_trips():
    rec = vpo.make_prims(G["inv_row"][None, :])[0]
    args = G["inv_args"]
    t = vpo.prim_eval("inv_cdf_span", rec, args).astype(np.float64)
    ref = G["inv_t"]
    fin = np.isfinite(ref) & (ref < 1e3)
    assert np.all(t[~np.isfinite(ref)] >= 1e20)  # saturated -> INF_F
    rel = np.abs(t[fin] - ref[fin]) / np.maximum(np.abs(ref[fin]), 1.0)
    assert np.median(rel) < 1e-5 and rel.max() < 2e-3
    # round trip on rays aimed through the primitive: from the entry point, the
    # optical depth over [t_entry, inv_cdf_span(t_entry, tau)] equals tau
    # whenever the scatter lands before the exit; otherwise tau exceeded the mass
    rng = np.random.default_rng(12)
    c = G["inv_row"][0:3].astype(np.float64)
    rows = []
    for _ in range(400):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o = c - 3.0 * d + rng.normal(size=3) * 0.15
        rows.append([*o, *d, 0.0, 0.0, 0.0])
    rt = np.array(rows, np.float32)
    te = vpo.prim_eval("entry_t", rec, rt)
    tx = vpo.prim_eval("exit_entry", rec, rt)
    hit = te >= 0
    rt[:, 6] = te
    full = vpo.prim_eval("optical_depth", rec, np.concatenate([rt[:, :6], te[:, None], tx[:, None], rt[:, 8:]], 1))
    rt[:, 8] = (rng.uniform(0.05, 1.2, 400) * full).astype(np.float32)
    ts = vpo.prim_eval("inv_cdf_span", rec, rt)
    rt[:, 7] = ts
    od = vpo.prim_eval("optical_depth", rec, rt)
    inside = hit & (ts <= tx) & (rt[:, 8] > 1e-3)
    beyond = hit & (rt[:, 8] > 1.001 * full)
    assert inside.sum() > 150
    assert np.allclose(od[inside], rt[inside, 8], rtol=2e-4, atol=2e-6)
    assert np.all(ts[beyond] > tx[beyond])  # more than the chord's mass: no scatter inside the bound


# ------------------------------------------------------------- phase -------
@pytest.mark.parametrize("g", [0.0, 0.85, -0.5, 0.3])
def test_hg_normalised_mean_cosine_and_pdf(g):
    wi = np.tile(np.array([0.0, 0.0, 1.0], np.float32), (1, 1))
    # integral over the sphere of the phase function = 1 (quadrature in cos theta)
    def f(c):
        wo = np.array([[math.sqrt(max(0.0, 1 - c * c)), 0.0, c]], np.float32)
        return 2 * math.pi * float(vpo.phase_eval(g, wi, wo)[0])
    val, _ = integrate.quad(f, -1, 1, epsrel=1e-8, limit=200)
    assert abs(val - 1.0) < 2e-5
    n = 200_000
    wis = np.random.default_rng(3).normal(size=(n, 3)).astype(np.float32)
    wis /= np.linalg.norm(wis, axis=1, keepdims=True)
    wo, pdf = vpo.phase_sample(g, wis, seed=5)
    cosv = np.sum(wo.astype(np.float64) * wis, axis=1)
    # mean cosine between the propagation direction and the scattered one is +g
    assert abs(cosv.mean() - g) < 5 * cosv.std() / math.sqrt(n) + 1e-4
    assert np.allclose(np.linalg.norm(wo, axis=1), 1.0, atol=2e-6)
    # cos recomputed from wo carries ~1e-7 rounding, amplified by d ln p/d cos <= 3g/(1-g)^2
    assert np.allclose(pdf, vpo.phase_eval(g, wis, wo), rtol=2e-4)


# -------------------------------------------------------- environment -----
def _env():
    return scenes.sky_env(128, 64, sun_radius_deg=3.0, sun_scale=200.0, supersample=4)


def test_alias_table_reproduces_reference_texel_distribution():
    env = _env()
    joint, prob, alias, total = vpo.env_tables(env)
    n = joint.size
    p_alias = prob.astype(np.float64) / n
    np.add.at(p_alias, alias, (1.0 - prob.astype(np.float64)) / n)
    marg, cond, joint2, total2 = vpo.env_ref_cdf(env)
    pm = np.diff(np.concatenate([[0.0], marg.astype(np.float64)]))
    pc = np.diff(np.concatenate([np.zeros((cond.shape[0], 1)), cond.astype(np.float64)], axis=1), axis=1)
    p_ref = (pm[:, None] * pc).reshape(-1)
    assert np.allclose(joint, joint2.reshape(-1)) and total == total2
    assert np.abs(p_alias - p_ref).max() < 1e-6
    assert abs(p_alias.sum() - 1.0) < 1e-9


def test_env_sampling_pdf_consistent_and_distribution():
    env = _env()
    tabs = vpo.env_tables(env)
    wo, pdf = vpo.env_sample(env, tabs, 200_000, seed=3)
    assert np.allclose(np.linalg.norm(wo, axis=1), 1.0, atol=3e-6)
    pdf2 = vpo.env_pdf(env, tabs, wo)
    agree = np.isclose(pdf, pdf2, rtol=1e-4)
    assert agree.mean() > 0.995  # disagreement only at texel boundaries
    # solid-angle integral of the pdf over the sphere = 1 (Monte Carlo with uniform directions)
    u = np.random.default_rng(4).normal(size=(400_000, 3)).astype(np.float32)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    est = vpo.env_pdf(env, tabs, u).astype(np.float64).mean() * 4 * math.pi
    assert abs(est - 1.0) < 0.02
    # chi-square of sampled texels vs joint/total
    h, w = env.shape[:2]
    theta = np.arctan2(wo[:, 2], wo[:, 0])
    polar = np.arccos(np.clip(wo[:, 1], -1, 1))
    ui = np.clip(((theta / (2 * math.pi) + 0.5) * w).astype(int), 0, w - 1)
    vi = np.clip((polar / math.pi * h).astype(int), 0, h - 1)
    counts = np.bincount(vi * w + ui, minlength=w * h).astype(np.float64)
    joint = tabs[0].astype(np.float64)
    expct = joint / joint.sum() * wo.shape[0]
    m = expct > 20
    chi2 = np.sum((counts[m] - expct[m]) ** 2 / expct[m])
    dof = m.sum() - 1
    assert chi2 < dof + 6 * math.sqrt(2 * dof)


def test_env_lookup_constant_and_texel_centres():
    const = scenes.constant_env((0.3, 0.6, 0.9))
    d = np.random.default_rng(8).normal(size=(1000, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rgb = vpo.env_lookup(const, d)
    assert np.all(rgb == np.array([0.3, 0.6, 0.9], np.float32))
    env = _env()
    dirs = scenes.texel_dirs(128, 64).reshape(-1, 3).astype(np.float32)
    rgb = vpo.env_lookup(env, dirs)
    assert np.allclose(rgb, env[..., :3].reshape(-1, 3), rtol=1e-4, atol=1e-5)


# ------------------------------------------------------------- camera ------
def test_perspective_camera_geometry():
    eye, tgt, up = np.array([0, 0, -3.5], np.float32), np.zeros(3, np.float32), np.array([0, 1, 0], np.float32)
    pod = vpo.camera_perspective(65, 33, eye, tgt, up, 40.0)
    o, d = vpo.camera_ray(pod, 32, 16, 0.0, 0.0)
    assert np.allclose(o, eye) and np.allclose(d, [0, 0, 1], atol=1e-6)
    # top-left pixel looks up-left: pixel row 0 is the top of the view (camera.h:55)
    _, d00 = vpo.camera_ray(pod, 0, 0, -0.5, -0.5)
    half_v = math.tan(math.radians(20.0))
    assert d00[1] > 0 and abs(d00[1] / d00[2] - half_v) < 1e-5
    # u = normalize(cross(up, eye - target)) = -x here, so column 0 lies at +x
    assert abs(d00[0] / d00[2] - half_v * 65 / 33) < 1e-5


# --------------------------------------------------- the estimator -------
def _cam(w, h):
    a = scenes.default_camera_args(w, h)
    return vpo.camera_perspective(a["w"], a["h"], a["eye"], a["target"], a["up"], a["vfov"])


def test_furnace_oracle():
    # conservative medium in a constant white field reproduces it: |mean-1| < max(5 SEM, 2e-3)
    # (experiments/mitsuba-reference/furnace_check.py:36-40)
    rows = scenes.random_scene(24, 12, albedo=1.0)
    rows[:, 10:13] = 1.0
    rec = vpo.make_prims(rows)[vpo.scene_order(rows)]
    img, _, k = vpo.render(rec, scenes.constant_env(), _cam(24, 24), 24, 24, 24, max_depth=1000, rr_depth=1000,
                           phase_g=0.5)
    m = img[..., 0].astype(np.float64)
    assert k["n_scatter"] > 5000
    assert abs(m.mean() - 1.0) < max(5 * m.std() / math.sqrt(m.size), 2e-3)


def test_absorption_image_matches_supersampled_closed_form():
    # C0 (one Gaussian, central optical depth 2, albedo 0): each pixel is the
    # box-filtered exp(-tau) of the closed form; the 16 jittered samples agree
    # within their Monte Carlo error.
    w = h = 24
    rows = scenes.c0_single_gaussian()
    rec = vpo.make_prims(rows)
    pod = _cam(w, h)
    img, smp, _ = vpo.render(rec, scenes.constant_env(), pod, w, h, 16, keep_samples=True)
    s, M = 0.25, float(rows[0, 13])
    ss = 24
    offs = (np.arange(ss) + 0.5) / ss - 0.5
    ref = np.zeros((h, w))
    var = np.zeros((h, w))
    eye = pod[0:3].astype(np.float64)
    p00, du, dv = pod[3:6].astype(np.float64), pod[6:9].astype(np.float64), pod[9:12].astype(np.float64)
    for y in range(h):
        for x in range(w):
            px = x + offs[None, :]
            py = y + offs[:, None]
            dirs = p00 + px[..., None] * du + py[..., None] * dv
            dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
            # impact parameter of the ray to the centre at the origin
            tc = -(eye[None, None, :] * dirs).sum(-1)
            dmin = np.linalg.norm(eye + tc[..., None] * dirs, axis=-1)
            from scipy import special
            tau = np.where(dmin < 3 * s, M / (2 * math.pi * s * s) * np.exp(-dmin ** 2 / (2 * s * s)) *
                           special.erf(np.sqrt(np.maximum(9 - dmin ** 2 / s ** 2, 0) / 2)), 0.0)
            T = np.exp(-tau)
            ref[y, x] = T.mean()
            var[y, x] = T.var()
    got = img[..., 0].astype(np.float64)
    sigma = np.sqrt(var / 16) + 2e-6
    z = np.abs(got - ref) / sigma
    assert np.mean(z < 3) > 0.99 and z.max() < 5
    # per-sample values are exp(-tau) of single rays: never above 1
    assert smp[..., 0].max() <= 1.0


def test_mis_and_ris_agree_in_expectation():
    rows = scenes.random_scene(24, 14, albedo=0.9)
    rec = vpo.make_prims(rows)[vpo.scene_order(rows)]
    env = _env()
    tabs = vpo.env_tables(env)
    kw = dict(max_depth=16, rr_depth=3, phase_g=0.5, env_tables_=tabs, keep_samples=True)
    a, sa, _ = vpo.render(rec, env, _cam(12, 12), 12, 12, 256, **kw)
    b, sb, _ = vpo.render(rec, env, _cam(12, 12), 12, 12, 256, use_ris=True, ris_candidates=6, **kw)
    ma, mb = sa[..., 1].astype(np.float64).mean(), sb[..., 1].astype(np.float64).mean()
    se = math.sqrt(sa[..., 1].var() / sa[..., 1].size + sb[..., 1].var() / sb[..., 1].size)
    assert abs(ma - mb) < 4 * se + 1e-4


def test_shadow_roulette_is_unbiased_and_stops_walks():
    # The shadow-ray roulette (shadow_rr = eps) keeps the estimator's
    # expectation: the image mean with eps = 0.3 matches eps = 0 within the
    # Monte Carlo error, on independent samples (different seeds), while the
    # shadow walks enter fewer primitives.  A roulette that dropped the 4x
    # survivor weight would darken the image by the killed share of the
    # direct light and fail the comparison.
    rows = scenes.random_scene(90, 21, smax=0.4, albedo=0.95)
    rec = vpo.make_prims(rows)[vpo.scene_order(rows)]
    env = _env()
    tabs = vpo.env_tables(env)
    kw = dict(max_depth=24, rr_depth=6, phase_g=0.3, env_tables_=tabs, keep_samples=True)
    a, sa, ka = vpo.render(rec, env, _cam(12, 12), 12, 12, 512, seed=1, **kw)
    b, sb, kb = vpo.render(rec, env, _cam(12, 12), 12, 12, 512, seed=2, shadow_rr=0.3, **kw)
    assert kb["n_sh_hit"] < 0.9 * ka["n_sh_hit"], (ka["n_sh_hit"], kb["n_sh_hit"])
    for c in range(3):
        xa, xb = sa[..., c].astype(np.float64), sb[..., c].astype(np.float64)
        se = math.sqrt(xa.var() / xa.size + xb.var() / xb.size)
        assert abs(xa.mean() - xb.mean()) < 4 * se, (c, xa.mean(), xb.mean(), se)
    # per pixel as well (3 sigma of the difference, allowing 1 % outliers)
    d = (sa[..., 1].mean(0) - sb[..., 1].mean(0)).astype(np.float64)
    s = np.sqrt((sa[..., 1].var(0) + sb[..., 1].var(0)) / sa.shape[0])
    assert np.mean(np.abs(d) > 3 * s + 1e-6) < 0.02


def test_shadow_roulette_levels_are_unbiased():
    # The roulette's estimate of T = e^-tau for one ray, over many paths
    # (keys): below tau_c = log(w / eps) it is T itself; past it the ray has
    # reached L = floor((tau - tau_c) / ln 4) + 1 levels and the estimate is
    # 4^L T with probability 4^-L (all L draws survive), else 0, so its mean
    # is T.  The first failing level is geometric with P(fail) = 3/4.
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 2**63, size=400_000, dtype=np.uint64)
    w, eps = 2.0, 1e-3
    T0, fail, tc = vpo.shadow_rr_T(w, eps, keys, 0.0)
    assert abs(tc - math.log(w / eps)) < 2.0 ** -15
    counts = np.bincount(np.minimum(fail, 6), minlength=7)[:4] / keys.size
    assert np.allclose(counts, [0.75, 0.1875, 0.046875, 0.01171875], atol=4e-3), counts
    step = math.log(4.0)
    for tau in [0.3, tc - 0.01, tc + 0.01, tc + 0.5 * step, tc + 1.2 * step, tc + 2.7 * step, tc + 4.5 * step]:
        T, _, _ = vpo.shadow_rr_T(w, eps, keys, tau)
        exact = math.exp(-tau)
        if tau < tc:
            assert np.allclose(T, np.float32(exact), rtol=2e-6, atol=0), tau
            continue
        L = int((tau - tc) // step) + 1
        alive = T > 0
        assert np.allclose(T[alive], np.float32(exact) * 4.0 ** L, rtol=2e-6), (tau, L)
        p = 4.0 ** -L
        assert abs(alive.mean() - p) < 4.0 * math.sqrt(p * (1 - p) / keys.size), (tau, alive.mean(), p)
        se = math.sqrt(exact * exact * (4.0 ** L - 1.0) / keys.size)
        assert abs(float(T.astype(np.float64).mean()) - exact) < 4.0 * se + 1e-12, (tau, T.mean(), exact, se)
    # the roulette off: T for every key
    Toff, foff, tcoff = vpo.shadow_rr_T(w, 0.0, keys[:1000], 7.0)
    assert (Toff == np.float32(math.exp(-7.0))).all() or np.allclose(Toff, math.exp(-7.0), rtol=2e-6)
    assert tcoff == 88.0


def test_shadow_roulette_keeps_the_furnace():
    rows = scenes.random_scene(24, 12, albedo=1.0)
    rows[:, 10:13] = 1.0
    rec = vpo.make_prims(rows)[vpo.scene_order(rows)]
    img, _, k = vpo.render(rec, scenes.constant_env(), _cam(24, 24), 24, 24, 24, max_depth=1000, rr_depth=1000,
                           phase_g=0.5, shadow_rr=0.05)
    m = img[..., 0].astype(np.float64)
    assert k["n_scatter"] > 5000
    assert abs(m.mean() - 1.0) < max(5 * m.std() / math.sqrt(m.size), 2e-3)


def test_argmin_free_flight_distribution():
    # The argmin of per-primitive closed-form free flights must sample the
    # first collision of the whole mixture: P[t* > t] = exp(-tau(0, t)) with
    # tau the summed FP64 optical depth, and P[escape] = exp(-tau_total).
    # Four overlapping Gaussians, the ray starting inside the first one.
    from golden.make_golden import chord, od_closed
    rng = np.random.default_rng(21)
    rows = scenes.random_scene(4, 77, smin=0.25, smax=0.45)
    rows[:, 0:3] = [[0.0, 0.0, 0.0], [0.1, -0.05, 0.9], [-0.1, 0.1, 1.7], [0.05, 0.0, 2.6]]
    rows[:, 13] = [0.35, 0.3, 0.25, 0.3]
    rec = vpo.make_prims(rows)
    o = np.array([0.05, 0.02, -0.2])
    d = np.array([0.02, -0.01, 1.0])
    d /= np.linalg.norm(d)
    ray = np.concatenate([o, d]).astype(np.float32)
    o64, d64 = ray[:3].astype(np.float64), ray[3:].astype(np.float64)
    chords = [chord(r, o64, d64)[0] for r in rows]
    assert chord(rows[0], o64, d64)[1] and all(c is not None for c in chords)

    def tau(t):
        return sum(od_closed(r, o64, d64, c[0], min(max(t, c[0]), c[1])) for r, c in zip(rows, chords))

    m = 200_000
    keys = rng.integers(0, 2**63, size=m, dtype=np.uint64)
    ts = vpo.first_scatter(rec, ray, keys).astype(np.float64)
    esc = ts >= 1e20
    tau_tot = tau(1e9)
    p_esc = math.exp(-tau_tot)
    assert 0.1 < p_esc < 0.5
    assert abs(esc.mean() - p_esc) < 4 * math.sqrt(p_esc * (1 - p_esc) / m)
    grid = np.linspace(0.0, max(c[1] for c in chords), 400)
    F = np.array([1.0 - math.exp(-tau(t)) for t in grid])
    F_emp = np.searchsorted(np.sort(ts[~esc]), grid, side="right") / m
    # Kolmogorov-Smirnov distance on the grid, 99.9 % critical value 1.95/sqrt(m)
    assert np.abs(F_emp - F).max() < 1.95 / math.sqrt(m), np.abs(F_emp - F).max()
    # every sampled collision lies inside some primitive's chord
    inside_any = np.zeros(m, bool)
    for c in chords:
        inside_any |= (ts >= c[0] - 1e-5) & (ts <= c[1] + 1e-5)
    assert inside_any[~esc].all()


@pytest.mark.parametrize("name", ["scatter", "ris", "small_cap", "measure", "dense"])
def test_bvh_queries_match_brute_force(name):
    # the CPU BVH (used for the timed CPU baseline) changes no image bit and
    # no overflow count: every query result is consumed order-independently
    rows = {"scatter": lambda: scenes.random_scene(200, 3), "ris": lambda: scenes.random_scene(120, 4),
            "small_cap": lambda: scenes.random_scene(80, 8, smax=0.5), "measure": lambda: scenes.random_scene(90, 5),
            "dense": lambda: scenes.c1_scattering(n=1024, seed=2)}[name]()
    kw = {"ris": dict(use_ris=True), "small_cap": dict(active_cap=2), "measure": dict(measure=True),
          "dense": dict(phase_g=0.0, max_depth=64)}.get(name, {})
    rec = vpo.make_prims(rows)[vpo.scene_order(rows)]
    env = scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    cam = vpo.camera_perspective(20, 20, np.array([0, 0, -3.5], np.float32), np.zeros(3, np.float32),
                                 np.array([0, 1, 0], np.float32), 40.0)
    a, _, ka = vpo.render(rec, env, cam, 20, 20, 2, **kw)
    b, _, kb = vpo.render(rec, env, cam, 20, 20, 2, bvh=True, **kw)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert ka["active_overflow"] == kb["active_overflow"]
    if name == "measure":
        assert ka["max_overlap"] == kb["max_overlap"] and ka["max_hits"] == kb["max_hits"]
