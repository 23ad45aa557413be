"""Write the FP64 golden vectors of tests/golden/golden_fp64.npz.

Independent of the oracle's code: every value is computed here in float64
from the primitive definition (centre, forward quaternion, scale, mass) with
numpy/scipy -- the 3-sigma ellipsoid chord from the exact quadratic, the
optical depth of the Gaussian density along the chord by adaptive quadrature
(scipy.integrate.quad) and, as a cross-check, by the closed form
M/((2pi)^(3/2) prod s) * sqrt(2pi)/|w| * exp(-perp^2/2) * (erf(b2)-erf(b1))/2.

Contents
  rows_*, rays_*, tau_*, T_*  : scenes of N primitives, fixed rays, FP64 total
                                optical depth and transmittance with the
                                reference's semantics (primitives containing the
                                origin integrate from 0 to their exit, entered
                                ones from entry to exit; sampling.cuh:548-583)
  c0_d, c0_tau                : single isotropic Gaussian (s = 0.25, central
                                optical depth 2): tau at impact parameter d
  inv_*                       : free-flight inversions (untruncated Gaussian CDF,
                                primitive.h:140-246) solved in FP64

Run from the repository root:  python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np
from scipy import integrate, special

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def load_scenes():
    import importlib.util
    pkg = os.path.join(ROOT, "cuda-volprim_amd")
    spec = importlib.util.spec_from_file_location("volprim_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["volprim_amd"] = mod
    spec.loader.exec_module(mod)
    from volprim_amd import scenes
    return scenes


def rot_matrix(q):
    w, x, y, z = np.asarray(q, np.float64) / np.linalg.norm(np.asarray(q, np.float64))
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def whiten(row, o, d):
    c = row[0:3].astype(np.float64)
    R = rot_matrix(row[3:7])
    s = row[7:10].astype(np.float64)
    p = (R.T @ (o - c)) / s
    w = (R.T @ d) / s
    return p, w


def density_along(row, o, d, t):
    p, w = whiten(row, o, d)
    s = row[7:10].astype(np.float64)
    x = p + t * w
    return float(row[13]) * math.exp(-0.5 * float(x @ x)) / ((2 * math.pi) ** 1.5 * float(np.prod(s)))


def chord(row, o, d):
    """[t0, t1] of the ray inside the 3-sigma ellipsoid, or None; origin-inside flag."""
    p, w = whiten(row, o, d)
    a, b, c = w @ w, p @ w, p @ p - 9.0
    disc = b * b - a * c
    if disc <= 0:
        return None, False
    r = math.sqrt(disc)
    t_in, t_out = (-b - r) / a, (-b + r) / a
    if c <= 0:  # origin inside
        return (0.0, t_out), True
    if t_in < 0:
        return None, False
    return (t_in, t_out), False


def od_closed(row, o, d, t0, t1):
    p, w = whiten(row, o, d)
    s = row[7:10].astype(np.float64)
    wl = math.sqrt(w @ w)
    wn = w / wl
    ps = p + t0 * w
    B = wn @ ps
    perp = ps @ ps - B * B
    L = (t1 - t0) * wl
    pref = float(row[13]) / ((2 * math.pi) ** 1.5 * float(np.prod(s))) * math.sqrt(2 * math.pi) / wl
    return pref * math.exp(-0.5 * perp) * 0.5 * (special.erf((B + L) / math.sqrt(2)) - special.erf(B / math.sqrt(2)))


def tau_ray(rows, o, d, check_quad):
    tot = 0.0
    for row in rows:
        ch, _ = chord(row, o, d)
        if ch is None:
            continue
        t0, t1 = ch
        v = od_closed(row, o, d, t0, t1)
        if check_quad:
            q, _ = integrate.quad(lambda t: density_along(row, o, d, t), t0, t1, epsabs=0, epsrel=1e-12, limit=200)
            assert abs(q - v) <= 1e-9 * max(abs(v), 1e-300) + 1e-14, (q, v)
        tot += v
    return tot


def main():
    scenes = load_scenes()
    rng = np.random.default_rng(20261202)
    out = {}
    cases = {"a": scenes.random_scene(40, 123, extent=0.8, smin=0.05, smax=0.35, rho_scale=2.0),
             "b": scenes.c1_scattering(n=4096, seed=1)[:256]}
    for name, rows in cases.items():
        m = 400
        o = rng.uniform(-1.2, 1.2, (m, 3))
        d = rng.normal(size=(m, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.concatenate([o, d], 1).astype(np.float32)
        tau = np.array([tau_ray(rows, rays[i, :3].astype(np.float64), rays[i, 3:].astype(np.float64), i < 40)
                        for i in range(m)])
        out[f"rows_{name}"] = rows
        out[f"rays_{name}"] = rays
        out[f"tau_{name}"] = tau
        out[f"T_{name}"] = np.exp(-tau)
    # single Gaussian: tau(d) = M/(2 pi s^2) exp(-d^2/(2 s^2)) erf(sqrt((9 - d^2/s^2)/2))
    c0 = scenes.c0_single_gaussian()
    s, M = float(c0[0, 7]), float(c0[0, 13])
    dd = np.linspace(0.0, 0.74, 38)
    out["c0_rows"] = c0
    out["c0_d"] = dd
    out["c0_tau"] = np.array([M / (2 * math.pi * s * s) * math.exp(-x * x / (2 * s * s)) *
                              special.erf(math.sqrt(max(9.0 - x * x / (s * s), 0.0) / 2.0)) for x in dd])
    # free-flight inversions from t_off: solve tau = integral_{t_off}^{t} sigma dt (untruncated)
    row = scenes.random_scene(1, 77, smin=0.2, smax=0.5, rho_scale=4.0)[0]
    inv_args, inv_t = [], []
    for k in range(200):
        o = rng.uniform(-2, 2, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o32, d32 = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
        p, w = whiten(row, o32, d32)
        wl = math.sqrt(w @ w)
        wp = (w @ p) / wl
        K = float(row[13]) / ((2 * math.pi) ** 1.5 * float(np.prod(row[7:10].astype(np.float64)))) \
            * math.sqrt(math.pi / 2) / wl * math.exp(-0.5 * (p @ p - wp * wp))
        t_off = float(np.float32(rng.uniform(0, 1.5)))
        tau = float(np.float32(rng.exponential(0.5)))
        arg = special.erf((wp + t_off * wl) / math.sqrt(2)) + tau / K
        t = (math.sqrt(2) * special.erfinv(arg) - wp) / wl if arg < 1.0 else math.inf
        inv_args.append([*o32, *d32, t_off, 0.0, tau])
        inv_t.append(t)
    out["inv_row"] = row
    out["inv_args"] = np.array(inv_args, np.float32)
    out["inv_t"] = np.array(inv_t)
    np.savez_compressed(os.path.join(HERE, "golden_fp64.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_fp64.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
