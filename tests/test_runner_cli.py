"""The test_runner harness: command line on the CPU, one validation scene on
the GPU checked against the closed-form transmittance."""
import os
import subprocess
import sys

import numpy as np
import pytest

from volprim_amd import vpio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "cuda-volprim_amd", "test_runner")


def _run(args, cwd, timeout=60):
    if not os.path.exists(RUNNER):
        pytest.fail(f"{RUNNER} is not built")
    return subprocess.run([RUNNER] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)


def test_list_scenes(tmp_path):
    r = _run(["--list"], tmp_path)
    assert r.returncode == 0
    for n in ["single_gaussian_validation", "two_gaussian_validation", "cluster_validation", "asset_validation"]:
        assert n in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--scene", "nope"], "not found"),
    ([], "No action specified"),
    (["--spp", "0"], "positive"),
    (["--phase-g", "1.5"], "out of range"),
    (["--bogus"], "unknown argument"),
    (["--scene", "single_gaussian_validation", "--denoise"], "denoiser"),
    (["--validation"], "--scene"),
])
def test_command_line_errors(tmp_path, args, msg):
    r = _run(args, tmp_path)
    assert r.returncode != 0
    assert msg in r.stderr


def test_missing_asset_reports_loader_error(tmp_path):
    env = dict(os.environ, SG_PLY=str(tmp_path / "absent.ply"))
    r = subprocess.run([RUNNER, "--scene", "asset_validation"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60, env=env)
    assert r.returncode != 0 and "Failed to load asset" in r.stderr


@pytest.mark.gpu
def test_single_gaussian_validation_matches_closed_form(tmp_path):
    # Orthographic 256x256 view of [-3,3]^2, one unit Gaussian of mass 7.5,
    # albedo 0, constant white environment: each pixel is the box-filtered
    # exp(-tau(b)), tau(b) = M/(2 pi) exp(-b^2/2) erf(sqrt(9 - b^2)/sqrt(2)).
    from scipy import special
    print("case: test_runner single_gaussian_validation spp 16", flush=True)
    r = _run(["--scene", "single_gaussian_validation", "--spp", "16", "--output-dir", str(tmp_path)], tmp_path,
             timeout=100)
    assert r.returncode == 0, r.stderr
    assert "Rendering complete" in r.stdout
    img = vpio.load_exr(str(tmp_path / "single_gaussian_validation" / "0000.exr"))
    assert img.shape == (256, 256, 4)
    ss = 8
    offs = (np.arange(ss) + 0.5) / ss
    coord = ((np.arange(256)[:, None] + offs[None, :]) / 256.0 * 6.0 - 3.0).reshape(-1)
    b2 = coord[:, None] ** 2 + coord[None, :] ** 2
    inside = b2 < 9.0
    tau = np.where(inside, 7.5 / (2 * np.pi) * np.exp(-b2 / 2) * special.erf(np.sqrt(np.maximum(9 - b2, 0) / 2)), 0)
    ref = np.exp(-tau).reshape(256, ss, 256, ss).mean(axis=(1, 3))
    lum = img[..., :3]
    assert np.allclose(lum[..., 0], lum[..., 1]) and np.allclose(lum[..., 0], lum[..., 2])
    err = np.abs(lum[..., 0] - ref)
    # 16 jittered samples of a smooth field: only the box-filter sampling error remains
    assert err.mean() < 2e-3 and err.max() < 3e-2, (err.mean(), err.max())
    assert abs(lum[128, 128, 0] - np.exp(-7.5 * special.erf(3 / np.sqrt(2)) / (2 * np.pi))) < 5e-3
