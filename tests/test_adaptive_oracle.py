"""The oracle's adaptive sampling (raygen.cuh:37-69,214-233 with
ENABLE_ADAPTIVE_SAMPLING on; batches of renderer.cpp:341-372), pinned
against a float64 restatement of the stopping rule over the oracle's own
per-sample radiance: a pixel runs batches of 16 samples and stops before a
batch once it has at least the minimum and the largest per-channel relative
sample deviation sqrt(M2/(n-1)) / max(mean, 1e-6) is below the threshold.
The float64 side uses two-pass variances, not the Welford recurrence, so a
dropped term, a standard error instead of a deviation or an off-by-one
batch shows up as a count mismatch away from the threshold."""
import numpy as np
import pytest

import vpo
from volprim_amd import scenes

W = H = 20
SPP = 96


def _setup(name):
    if name == "c0":
        rows, env, g = scenes.c0_single_gaussian(), scenes.constant_env(), 0.85
    else:
        rows, env, g = scenes.random_scene(48, 5), scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0), 0.85
    rec = vpo.make_prims(rows)
    cam = vpo.camera_perspective(W, H, np.array([0, 0, -3.5], np.float32), np.zeros(3, np.float32),
                                 np.array([0, 1, 0], np.float32), 40.0)
    return rec, env, cam, g


def _f64_counts(smp, thr, min_samples, batch=16):
    """Stopping count and mean per pixel from per-sample values, in float64."""
    x = smp[..., :3].astype(np.float64)  # [s, h, w, 3]
    spp = x.shape[0]
    counts = np.zeros(x.shape[1:3], np.int64)
    margin = np.full(x.shape[1:3], np.inf)
    for yy in range(x.shape[1]):
        for xx in range(x.shape[2]):
            n = 0
            while n < spp:
                if n >= min_samples:
                    v = x[:n, yy, xx]
                    mean = v.mean(0)
                    std = np.sqrt(((v - mean) ** 2).sum(0) / (n - 1))
                    err = (std / np.maximum(mean, 1e-6)).max()
                    margin[yy, xx] = min(margin[yy, xx], abs(err - thr) / max(thr, 1e-30))
                    if err < thr:
                        break
                n += min(batch, spp - n)
            counts[yy, xx] = n
    return counts, margin


@pytest.mark.parametrize("name,thr,min_samples", [("mis", 0.2, 32), ("mis", 1.0, 32), ("c0", 0.02, 32),
                                                  ("mis", 0.5, 40)])
def test_adaptive_counts_and_means_follow_the_rule(name, thr, min_samples):
    rec, env, cam, g = _setup(name)
    _, smp, _ = vpo.render(rec, env, cam, W, H, SPP, seed=3, phase_g=g, keep_samples=True)
    mean, _, k = vpo.render(rec, env, cam, W, H, SPP, seed=3, phase_g=g, adaptive_threshold=thr,
                            adaptive_min_samples=min_samples)
    got = k["samples_per_pixel"].astype(np.int64)
    want, margin = _f64_counts(smp, thr, min_samples)
    # decisions may differ only where float32 and float64 land on either side
    # of the threshold
    differ = got != want
    assert not (differ & (margin > 1e-3)).any(), (got[differ][:5], want[differ][:5])
    assert differ.sum() <= 2
    assert 0 < (got < SPP).sum() < got.size, "the case should stop some pixels and not others"
    # the Welford mean equals the mean of the samples each pixel took
    x = smp[..., :3].astype(np.float64)
    for yy in range(H):
        for xx in range(W):
            n = got[yy, xx]
            ref = x[:n, yy, xx].mean(0)
            assert np.allclose(mean[yy, xx, :3], ref, rtol=2e-5, atol=1e-6), (yy, xx, n)
    assert (mean[..., 3] == 1.0).all()


def test_adaptive_threshold_extremes():
    rec, env, cam, g = _setup("mis")
    # every pixel passes at the first test: the minimum rounded up to whole batches
    _, _, k = vpo.render(rec, env, cam, W, H, SPP, seed=3, phase_g=g, adaptive_threshold=1e30,
                         adaptive_min_samples=40)
    assert (k["samples_per_pixel"] == 48).all()
    _, _, k = vpo.render(rec, env, cam, W, H, SPP, seed=3, phase_g=g, adaptive_threshold=1e30,
                         adaptive_min_samples=1)
    assert (k["samples_per_pixel"] == 16).all()
    # a threshold nothing reaches: every pixel takes every sample, and the
    # Welford mean agrees with the exact-sum mean of the plain render
    plain, _, _ = vpo.render(rec, env, cam, W, H, SPP, seed=3, phase_g=g)
    m, _, k = vpo.render(rec, env, cam, W, H, SPP, seed=3, phase_g=g, adaptive_threshold=1e-30)
    c = k["samples_per_pixel"]
    assert (c == SPP).sum() >= c.size - 4
    assert np.allclose(m[c == SPP], plain[c == SPP], rtol=2e-5, atol=1e-6)
    # a smaller batch: counts are multiples of it
    _, _, k = vpo.render(rec, env, cam, W, H, SPP, seed=3, phase_g=g, adaptive_threshold=1.0,
                         adaptive_min_samples=8, adaptive_batch=4)
    c = k["samples_per_pixel"]
    assert (c % 4 == 0).all() and c.min() >= 8


def test_adaptive_rejects_partial_renders():
    rec, env, cam, g = _setup("mis")
    with pytest.raises(ValueError):
        vpo.render(rec, env, cam, W, H, SPP, adaptive_threshold=0.5, samples=(0, 32))
    with pytest.raises(ValueError):
        vpo.render(rec, env, cam, W, H, SPP, adaptive_threshold=0.5, keep_samples=True)
    with pytest.raises(ValueError):
        vpo.render(rec, env, cam, W, H, SPP, adaptive_threshold=0.5, adaptive_min_samples=0)
