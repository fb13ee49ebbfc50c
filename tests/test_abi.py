"""The C-ABI library (CPU only: loading, exports, host-side logic)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

from anchorwave_amd import ksw2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ksw2_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(ksw\w+)\s*\(", src, flags=re.M))
    return sorted(n for n in names if not n.endswith("_cb"))


def test_library_exports_every_declared_symbol():
    lib = ksw2.lib()  # fails loudly if the library was not built
    decl = declared_functions()
    assert "ksw_extd2_hip" in decl and len(decl) >= 10
    nm = subprocess.run(["nm", "-D", "--defined-only", ksw2.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in nm.splitlines() if " T " in line)
    missing = [d for d in decl if d not in exported]
    assert not missing, missing
    for d in decl:
        getattr(lib, d)
    assert set(ksw2.EXPORTS) == set(decl)


def test_struct_layout():
    assert C.sizeof(ksw2.KswExtz) == 56
    assert ksw2.KswExtz.max_q.offset == 4 and ksw2.KswExtz.cigar.offset == 48
    assert C.sizeof(ksw2.KswPair) == 40


def test_error_strings_and_no_device_path():
    lib = ksw2.lib()
    for st in (0, -1, -2, -3, -4, -5):
        assert lib.ksw2_hip_error_string(st)
    if not os.path.exists("/dev/kfd"):
        assert lib.ksw2_hip_init(0) == -5  # no GPU: ENODEV, no crash


def test_release_cached_memory_with_nothing_held():
    """With no engine block cached the release is a no-op that needs no GPU."""
    assert ksw2.lib().ksw2_hip_release_cached_memory() == 0
    ksw2.release_cached_memory()


def test_early_returns_need_no_device():
    """Calls the reference answers before aligning return the reset ez
    without touching a GPU (minimap2/ksw2_extd2_sse.c:81,105)."""
    t = np.zeros(10, np.uint8)
    r = ksw2.extd2(t[:0], t)  # qlen == 0
    assert r.score == ksw2.KSW_NEG_INF and r.n_cigar == 0 and r.zdropped == 0
    r = ksw2.extd2(t, t, mat=ksw2.default_matrix(0, -30))  # -min_sc > 2(q+e)
    assert r.score == ksw2.KSW_NEG_INF and r.n_cigar == 0 and r.mqe == ksw2.KSW_NEG_INF
    assert r.max_t == -1 and r.max_q == -1 and r.mte_q == -1 and r.mqe_t == -1


def test_footprint_model():
    """ksw2_hip_footprint (host only): the admission footprint of a call,
    at least one direction byte per cell for a full traceback, smaller with
    a band, 0 when the reference returns before aligning."""
    rng = np.random.default_rng(5)
    q = rng.integers(0, 4, 3000).astype(np.uint8)
    t = rng.integers(0, 4, 3200).astype(np.uint8)
    full = ksw2.footprint(q, t)
    # (the pass layout pads the target to 1024-row passes and stores the
    # pipeline fill and drain steps)
    assert full >= len(q) * len(t) and full < 1.5 * len(q) * len(t) + (1 << 20)
    assert ksw2.footprint(q, t, flag=ksw2.KSW_EZ_SCORE_ONLY) < full // 50  # no direction bytes
    banded = ksw2.footprint(q, t, w=700)
    assert banded < full and banded >= (len(q) + len(t) - 1) * 700
    assert ksw2.footprint(q[:1000], t[:1000]) < full
    assert ksw2.footprint(q[:0], t) == 0
    assert ksw2.footprint(q, t, m=1) == 0


def test_length_overflow_is_tested_before_the_scoring():
    """ksw_extd2_avx512 sets zdropped for qlen > INT_MAX - tlen + 1 before it
    swaps the gap pieces or returns early on -min_sc > 2(q+e)
    (minimap2/ksw2_extd2_avx.c:136-140): such a call is zdropped even with a
    matrix that would return early, and needs no device.  The lengths are
    only compared, never read past the small buffers."""
    import ctypes as C
    lib = ksw2.lib()
    buf = np.zeros(16, np.uint8)
    mat = ksw2.default_matrix(0, -30)  # -min_sc = 30 > 2(q+e) = 20
    ez = ksw2.KswExtz()
    lib.ksw_extd2_hip(None, 0x7fffffff, buf.ctypes.data, 2, buf.ctypes.data, 5, mat.ctypes.data,
                      8, 2, 75, 1, -1, -1, 0, 0, C.byref(ez))
    assert ez.zdropped == 1 and ez.score == ksw2.KSW_NEG_INF and ez.n_cigar == 0
    assert lib.ksw2_hip_footprint(0x7fffffff, buf.ctypes.data, 2, buf.ctypes.data, 5, mat.ctypes.data,
                                  8, 2, 75, 1, -1, -1, 0) == 2**64 - 1
    ez = ksw2.KswExtz()
    lib.ksw_extd2_hip(None, 16, buf.ctypes.data, 16, buf.ctypes.data, 5, mat.ctypes.data,
                      8, 2, 75, 1, -1, -1, 0, 0, C.byref(ez))
    assert ez.zdropped == 0 and ez.score == ksw2.KSW_NEG_INF  # early return, no overflow


def test_noalloc_rejections_need_no_device():
    """ksw_extd2_hip_noalloc refuses, before any device work, a CIGAR buffer
    smaller than qlen + tlen words (it never grows the caller's buffer) and a
    call larger than its fixed scratch; early returns reset ez as
    ksw_extd2_hip does."""
    rng = np.random.default_rng(3)
    t = rng.integers(0, 4, 12000).astype(np.uint8)
    st, r = ksw2.extd2_noalloc(t[:600], t[:600], cigar_capacity=1199)
    assert st == -1 and r.n_cigar == 0
    st, r = ksw2.extd2_noalloc(t, t)  # ~150 MB of direction bytes > 64 MiB
    assert st == ksw2.ETOOBIG and r.zdropped == 1 and r.n_cigar == 0
    st, r = ksw2.extd2_noalloc(t[:0], t)  # qlen == 0: the reference returns at once
    assert st == 0 and r.score == ksw2.KSW_NEG_INF and r.zdropped == 0
    st, r = ksw2.extd2_noalloc(t[:10], t[:10], flag=ksw2.KSW_EZ_SCORE_ONLY, cigar_capacity=0, m=1)
    assert st == 0 and r.score == ksw2.KSW_NEG_INF
