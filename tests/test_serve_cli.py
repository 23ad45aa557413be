"""vp_serve: the JSON-lines request/response front end of the server."""
import json
import os
import subprocess

import numpy as np
import pytest

from volprim_amd import scenes, volprim, vpio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVE = os.path.join(ROOT, "cuda-volprim_amd", "vp_serve")


def _run(lines, cwd, timeout=120):
    if not os.path.exists(SERVE):
        pytest.fail(f"{SERVE} is not built")
    r = subprocess.run([SERVE], input="\n".join(lines) + "\n", cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stderr
    return [json.loads(l) for l in r.stdout.splitlines() if l.strip()]


def test_malformed_requests_get_error_responses(tmp_path):
    out = _run(['{"id": "a"', "not json", '[1, 2]', '{"id": 3, "output": "x.exr"}',
                '{"id": "b", "scene": "s.ply"}', '{"id": "c", "scene": "s.ply", "output": "o.exr", "spp": -1}',
                '{"id": "d", "scene": "s.ply", "output": "o.exr", "camera": {"eye": [1, 2]}}'], tmp_path)
    assert [o["status"] for o in out] == ["error"] * 7
    assert "bad request" in out[0]["message"] and "bad request" in out[1]["message"]
    assert "JSON object" in out[2]["message"]
    assert out[3]["id"] == "3" and "scene" in out[3]["message"]
    assert "output" in out[4]["message"]
    assert "spp" in out[5]["message"] and "eye" in out[6]["message"]


def test_options(tmp_path):
    # an unknown option is refused before any input is read
    r = subprocess.run([SERVE, "--bogus"], input="", cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--reserve-cus" in r.stderr
    # the CU reservation is accepted; malformed requests are answered without a GPU
    r = subprocess.run([SERVE, "--reserve-cus"], input='{"id": "a"\n', cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = [json.loads(l) for l in r.stdout.splitlines() if l.strip()]
    assert out and out[0]["status"] == "error"


@pytest.mark.gpu
def test_served_requests_match_direct_renders(tmp_path, ctx):
    print("case: vp_serve 3 requests", flush=True)
    rows = scenes.random_scene(60, 31, albedo=0.9)
    ply = str(tmp_path / "scene.ply")
    vpio.save_ply(ply, rows, sigma_multiplier=1.0)
    reqs = [dict(id="r0", scene=ply, sigma_multiplier=1.0, width=24, height=16, spp=3, seed=5, phase_g=0.3,
                 camera=dict(eye=[0, 0, -3.5], target=[0, 0, 0], up=[0, 1, 0], fov=40), output=str(tmp_path / "0.exr")),
            dict(id="r1", scene=ply, sigma_multiplier=1.0, width=16, height=16, spp=2, seed=6, ris=True,
                 camera=dict(eye=[3.0, 0.5, 0], target=[0, 0, 0], up=[0, 1, 0], fov=50), output=str(tmp_path / "1.exr")),
            dict(id="r2", scene=ply, sigma_multiplier=1.0, width=20, height=20, spp=1, seed=7,
                 camera=dict(eye=[0, 0, -5], target=[0, 0, 0], up=[0, 1, 0], ortho_height=3.0),
                 output=str(tmp_path / "2.exr"))]
    out = _run([json.dumps(r) for r in reqs], tmp_path, timeout=100)
    resp = {o["id"]: o for o in out if "id" in o}
    assert all(resp[r["id"]]["status"] == "ok" for r in reqs), out
    assert out[-1]["stats"]["completed"] == 3
    # the same renders through the C ABI directly
    sc = ctx.scene(vpio.load_ply(ply, 1.0), bvh="median")
    env = ctx.env(np.ones((32, 64, 4), np.float32))
    for r in reqs:
        c = r["camera"]
        args = dict(w=r["width"], h=r["height"], eye=np.array(c["eye"], np.float32),
                    target=np.array(c["target"], np.float32), up=np.array(c["up"], np.float32))
        cam = (volprim.camera_orthographic(**args, ortho_h=c["ortho_height"]) if "ortho_height" in c
               else volprim.camera_perspective(**args, vfov=c["fov"]))
        p = volprim.default_params(width=r["width"], height=r["height"], spp=r["spp"], seed=r["seed"],
                                   phase_g=r.get("phase_g", 0.85), use_ris=int(r.get("ris", False)))
        ref, _ = ctx.render(sc, env, cam, p)
        got = vpio.load_exr(r["output"])
        assert np.array_equal(got[..., :3], ref[..., :3]), r["id"]
    sc.close()
    env.close()
