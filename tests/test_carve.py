"""The engine pool's bookkeeping (RangeCarver, anchorwave_amd/csrc/ksw2_carve.h),
compiled on the host by tests/c/carve_check.cpp: random allocations, early
releases of a block's parts and frees keep free and owned ranges an exact
tiling of the pool, and freeing a block never frees a later block carved from
a part it gave back."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_range_carver(tmp_path):
    exe = tmp_path / "carve_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", os.path.join(ROOT, "tests", "c", "carve_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "carve ok" in r.stdout
