"""ksw_extd2_hip under AnchorWave's kalloc arena contract (km != NULL),
from a compiled C caller: tests/c/kalloc_arena.c links the reference's
kalloc and ksw_extd2_avx512 (oracle/_ref) beside libksw2hip.so and compares
every ksw_extz_t field, m_cigar included, and the CIGAR, on fresh and reused
ez, then frees the CIGARs with kfree."""
import os
import subprocess

import pytest

from anchorwave_amd import ksw2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
LIB_DIR = os.path.dirname(ksw2.LIB_PATH)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "libksw2ref.so")),
                    reason="reference checker oracle/_ref not built")
def test_kalloc_arena_caller(tmp_path):
    exe = tmp_path / "kalloc_arena"
    cmd = ["gcc", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c", "kalloc_arena.c"),
           "-o", str(exe), "-L", REF_DIR, "-lksw2ref", "-L", LIB_DIR, "-lksw2hip",
           f"-Wl,-rpath,{REF_DIR}:{LIB_DIR}"]
    subprocess.run(cmd, check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(r.stdout + r.stderr[-4000:], flush=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "differ" not in r.stderr
