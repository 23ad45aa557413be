"""Host check of the walk-tree traversal shared by every device walk.

Compiles tests/walk_tree_check.cpp against cuda-volprim_amd/csrc/vp_walk.h
with g++ and runs it: walks of the implicit 4-ary tree with the stackless
bit trail must meet exactly the leaves that brute force over the leaf boxes
finds, in ascending order, load every node at most once and never leave the
tree (ray walks, walks whose segment end shrinks while they run, point
walks; 1 to 20 000 leaves, including every size around a power of 4).
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
