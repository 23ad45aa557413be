"""vp_render_device_noalloc, the entry point that allocates nothing: its
image equals vp_render_device's bit for bit on two calls in a row (the second
smaller, so stale scratch contents from the first call would show), and
sizes beyond its scratch are rejected on the host."""
import numpy as np
import pytest
import torch

from volprim_amd import scenes, volprim

pytestmark = pytest.mark.gpu


def _render_both(ctx, rows, env, w, h, spp, seed, **kw):
    sc = ctx.scene(rows)
    ev = ctx.env(env)
    cam = volprim.camera_perspective(**scenes.default_camera_args(w, h))
    p = volprim.default_params(width=w, height=h, spp=spp, seed=seed, **kw)
    ref = torch.empty((h, w, 4), dtype=torch.float32, device="cuda:0")
    got = torch.full((h, w, 4), -1.0, dtype=torch.float32, device="cuda:0")
    print(f"[case] noalloc {w}x{h}x{spp} seed {seed} {kw}", flush=True)
    st_ref = ctx.render_device(sc, ev, cam, p, ref.data_ptr())
    st = ctx.render_device_noalloc(sc, ev, cam, p, got.data_ptr())
    torch.cuda.synchronize()
    sc.close()
    ev.close()
    return got.cpu().numpy(), ref.cpu().numpy(), st, st_ref


def test_noalloc_matches_render_device_twice(ctx):
    env = scenes.sky_env(64, 32, sun_radius_deg=4.0, sun_scale=50.0)
    # first call: 96x80 at 6 spp over 300 Gaussians, MIS
    got, ref, st, st_ref = _render_both(ctx, scenes.random_scene(300, 3, albedo=0.9), env, 96, 80, 6, 11,
                                        phase_g=0.3)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), np.abs(got - ref).max()
    assert st["active_overflow"] == st_ref["active_overflow"]
    # second call, smaller image and another scene, RIS with a small active-set
    # capacity: the scratch still holds the first call's sums and counters
    got2, ref2, st2, st_ref2 = _render_both(ctx, scenes.random_scene(40, 8, smax=0.4, albedo=0.8), env, 40, 24, 3,
                                            5, use_ris=1, active_cap=16)
    assert np.array_equal(got2.view(np.uint32), ref2.view(np.uint32)), np.abs(got2 - ref2).max()
    assert st2["active_overflow"] == st_ref2["active_overflow"]
    assert np.isfinite(got2).all() and got2[..., :3].mean() > 0.0


def test_noalloc_rejects_sizes_beyond_its_scratch(ctx):
    sc = ctx.scene(scenes.random_scene(8, 2))
    ev = ctx.env(scenes.constant_env())
    cam = volprim.camera_perspective(**scenes.default_camera_args(513, 512))
    out = torch.empty((512, 513, 4), dtype=torch.float32, device="cuda:0")
    try:
        with pytest.raises(volprim.VpError, match="512x512"):
            ctx.render_device_noalloc(sc, ev, cam, volprim.default_params(width=513, height=512, spp=1),
                                      out.data_ptr())
        with pytest.raises(volprim.VpError, match="no room"):
            ctx.render_device_noalloc(sc, ev, cam, volprim.default_params(width=8, height=8, spp=1,
                                                                          active_cap=60000), out.data_ptr())
    finally:
        sc.close()
        ev.close()
