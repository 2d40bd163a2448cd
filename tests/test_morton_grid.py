"""Quadrant keys with the exact levels read off a grid (csrc/morton_grid.h)
against the level-by-level FP64 descent of insertIntoChild
(BarnesHutTree.cpp:100-134), compiled and run on the CPU
(tests/cpp/morton_grid_check.cpp): tree-like and arbitrary roots, points on
every depth's cell bounds and one ulp either side, NaN and infinities."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_grid_keys_match_descent(tmp_path):
    exe = str(tmp_path / "morton_grid_check")
    # Host-only build: the header's functions are __host__ __device__.
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    "-I", os.path.join(ROOT, "neon-physics-fluxion_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "morton_grid_check.cpp"), "-o", exe],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe, "2000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " mismatches 0 " in r.stdout
