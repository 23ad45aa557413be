"""Host check of the walk-tree traversal shared by every device walk.

Compiles tests/walk_tree_check.cpp against cuda-volprim_amd/csrc/vp_walk.h
with g++ and runs it: walks of the implicit 4-ary tree with the stackless
bit trail must meet exactly the leaves that brute force over the leaf boxes
finds, in ascending order, load every node at most once and never leave the
tree (ray walks, walks whose segment end shrinks while they run, point
walks; 1 to 20 000 leaves, including every size around a power of 4), and
the 16-bit form of the tree (bounds rounded outward to binary16) must
contain every 32-bit box and lose none of the 32-bit walks' leaves.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("seed", [1, 2])
def test_walk_tree_matches_brute_force(tmp_path, seed):
    exe = str(tmp_path / "walk_tree_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                    os.path.join(ROOT, "cuda-volprim_amd", "csrc"), "-o", exe,
                    os.path.join(ROOT, "tests", "walk_tree_check.cpp")], check=True)
    r = subprocess.run([exe, str(seed), "200"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_f16_directed_rounding_matches_numpy(tmp_path):
    """f16_down / f16_up (the 16-bit walk-tree records' outward rounding)
    against numpy's binary16: f16_down(x) is the largest binary16 <= x, found
    from numpy's round-to-nearest value and one np.nextafter step, and
    f16_up(x) the smallest >= x; f16_value decodes exactly as numpy does.
    Inputs cover every binary16 exponent, subnormals, values just past the
    largest finite 65504, infinities, signed zeros and every float32 that is
    itself a binary16 value."""
    import numpy as np
    exe = str(tmp_path / "walk_tree_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                    os.path.join(ROOT, "cuda-volprim_amd", "csrc"), "-o", exe,
                    os.path.join(ROOT, "tests", "walk_tree_check.cpp")], check=True)
    rng = np.random.default_rng(7)
    exact = np.arange(0, 0x10000, dtype=np.uint16).view(np.float16)
    exact = exact[np.isfinite(exact)].astype(np.float32)
    mags = np.exp2(rng.uniform(-30.0, 17.0, 200_000)).astype(np.float32)
    signs = np.where(rng.random(mags.size) < 0.5, -1.0, 1.0).astype(np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, 65504.0, -65504.0, 65505.0, -65505.0, 65519.0, 65520.0,
                        1e30, -1e30, 2.0 ** -24, 2.0 ** -25, -(2.0 ** -25), 2.0 ** -14, 2.0 ** -14 * (1 - 2 ** -20),
                        1e-45, -1e-45], np.float32)
    x = np.concatenate([exact, mags * signs, special, np.nextafter(exact, np.float32(np.inf)),
                        np.nextafter(exact, np.float32(-np.inf))]).astype(np.float32)
    (tmp_path / "in.bin").write_bytes(x.tobytes())
    r = subprocess.run([exe, "f16", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rec = np.frombuffer((tmp_path / "out.bin").read_bytes(), dtype=np.dtype(
        [("down", "<u2"), ("up", "<u2"), ("vdown", "<f4"), ("vup", "<f4")]))
    assert rec.size == x.size
    with np.errstate(over="ignore"):
        rn = x.astype(np.float16)
        rn32 = rn.astype(np.float32)
        down = np.where(rn32 <= x, rn, np.nextafter(rn, np.float16(-np.inf)))
        up = np.where(rn32 >= x, rn, np.nextafter(rn, np.float16(np.inf)))
    # a zero result keeps the sign numpy gives it; compare values for zeros
    got_down = rec["down"].view(np.float16)
    got_up = rec["up"].view(np.float16)
    bad = ~((got_down == down) & (np.signbit(got_down) == np.signbit(down)) | ((got_down == 0) & (down == 0)))
    assert not bad.any(), (x[bad][:5], got_down[bad][:5], down[bad][:5])
    bad = ~((got_up == up) & (np.signbit(got_up) == np.signbit(up)) | ((got_up == 0) & (up == 0)))
    assert not bad.any(), (x[bad][:5], got_up[bad][:5], up[bad][:5])
    assert np.array_equal(rec["vdown"], got_down.astype(np.float32))
    assert np.array_equal(rec["vup"], got_up.astype(np.float32))
    assert (rec["vdown"] <= x).all() and (rec["vup"] >= x).all()
