"""ctypes binding of ``lib/libksw2hip.so`` (C ABI in ``include/ksw2_hip.h``).

Mirrors the reference call ``ksw_extd2_sse(km, qlen, query, tlen, target, m,
mat, q, e, q2, e2, w, zdrop, end_bonus, flag, ez)`` (minimap2/ksw2.h:91-107):
same argument meaning, same ``ksw_extz_t`` output, same error conventions
(``ez.zdropped`` / ``score == KSW_NEG_INF``), plus a batch entry point.

The library is the product path.  Loading fails loudly when it has not been
built (run ``__graft_entry__.build()``); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Iterable, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libksw2hip.so")

KSW_NEG_INF = -0x40000000
KSW_EZ_SCORE_ONLY = 0x01
KSW_EZ_RIGHT = 0x02
KSW_EZ_GENERIC_SC = 0x04
KSW_EZ_APPROX_MAX = 0x08
KSW_EZ_APPROX_DROP = 0x10
KSW_EZ_EXTZ_ONLY = 0x40
KSW_EZ_REV_CIGAR = 0x80
KSW_EZ_SEMIGLOBAL_END = 0x800
KSW_EZ_SINGLETRACK = 0x1000  # two-track (Singletrack) traceback

STATUS = {0: "ok", -1: "invalid argument", -2: "unsupported", -3: "out of memory",
          -4: "HIP runtime error", -5: "no usable gfx950 device", -6: "deferred: device memory committed",
          -7: "larger than the no-allocation scratch"}
EAGAIN = -6
ETOOBIG = -7
NOALLOC_SCRATCH_BYTES = 64 << 20  # KSW2_HIP_NOALLOC_SCRATCH_BYTES

# Symbols include/ksw2_hip.h declares; tests check the library exports each.
EXPORTS = (
    "ksw_extd2_hip",
    "ksw_extd2_singletrack_hip",
    "ksw_extd2_hip_batch",
    "ksw_extd2_hip_noalloc",
    "ksw2_hip_release_cached_memory",
    "ksw2_hip_last_error",
    "ksw2_hip_error_string",
    "ksw2_hip_init",
    "ksw2_hip_route",
    "ksw2_hip_set_progress_callback",
    "ksw2_hip_clear_progress_callback",
    "ksw2_hip_server_create",
    "ksw2_hip_server_submit",
    "ksw2_hip_server_poll",
    "ksw2_hip_server_poll_status",
    "ksw2_hip_server_destroy",
    "ksw2_hip_plan_create",
    "ksw2_hip_plan_run",
    "ksw2_hip_plan_fetch",
    "ksw2_hip_plan_timing",
    "ksw2_hip_plan_graph",
    "ksw2_hip_plan_destroy",
    "ksw2_hip_plan_dirs",
    "ksw2_hip_certify_batch",
    "ksw2_hip_certify_free",
    "ksw2_hip_window_batch",
    "ksw2_hip_window_free",
    "ksw2_hip_footprint",
    "ksw2_hip_server_try_submit",
    "ksw2_hip_server_load",
)


class KswExtz(C.Structure):
    """``ksw_extz_t``, layout of minimap2/ksw2.h:40-50."""

    _fields_ = [
        ("max", C.c_uint32, 31),
        ("zdropped", C.c_uint32, 1),
        ("max_q", C.c_int), ("max_t", C.c_int),
        ("mqe", C.c_int), ("mqe_t", C.c_int),
        ("mte", C.c_int), ("mte_q", C.c_int),
        ("score", C.c_int),
        ("m_cigar", C.c_int), ("n_cigar", C.c_int),
        ("reach_end", C.c_int),
        ("stopped", C.c_int),
        ("cigar", C.POINTER(C.c_uint32)),
    ]


assert C.sizeof(KswExtz) == 56


class KswPair(C.Structure):
    """``ksw2_pair_t`` of include/ksw2_hip.h."""

    _fields_ = [
        ("query", C.c_void_p), ("target", C.c_void_p),
        ("qlen", C.c_int32), ("tlen", C.c_int32),
        ("w", C.c_int32), ("zdrop", C.c_int32), ("end_bonus", C.c_int32), ("flag", C.c_int32),
    ]


@dataclass(frozen=True)
class Result:
    """Every output field of ``ksw_extz_t`` plus the CIGAR as a tuple."""

    max: int
    zdropped: int
    max_q: int
    max_t: int
    mqe: int
    mqe_t: int
    mte: int
    mte_q: int
    score: int
    n_cigar: int
    reach_end: int
    stopped: int
    cigar: tuple

    @staticmethod
    def from_ez(ez: KswExtz) -> "Result":
        cig = tuple(np.ctypeslib.as_array(ez.cigar, (ez.n_cigar,)).tolist()) if ez.n_cigar > 0 else ()
        return Result(ez.max, ez.zdropped, ez.max_q, ez.max_t, ez.mqe, ez.mqe_t, ez.mte,
                      ez.mte_q, ez.score, ez.n_cigar, ez.reach_end, ez.stopped, cig)


class KswError(RuntimeError):
    pass


_lib = None
_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


def release_cached_memory() -> None:
    """``ksw2_hip_release_cached_memory``: frees the device block that batch
    calls larger than the memory budget keep for the next such call."""
    st = lib().ksw2_hip_release_cached_memory()
    if st != 0:
        raise KswError(f"ksw2_hip_release_cached_memory: {STATUS.get(st, st)}")


def lib():
    """Loads the HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KswError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.ksw_extd2_hip.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int8,
                                    C.c_void_p, C.c_int8, C.c_int8, C.c_int8, C.c_int8, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.POINTER(KswExtz)]
        L.ksw_extd2_singletrack_hip.argtypes = L.ksw_extd2_hip.argtypes
        L.ksw_extd2_hip.restype = None
        L.ksw_extd2_hip_noalloc.argtypes = L.ksw_extd2_hip.argtypes[1:]
        L.ksw_extd2_hip_noalloc.restype = C.c_int
        L.ksw_extd2_hip_batch.argtypes = [C.c_int, C.POINTER(KswPair), C.c_int8, C.c_void_p, C.c_int8,
                                          C.c_int8, C.c_int8, C.c_int8, C.POINTER(KswExtz)]
        L.ksw_extd2_hip_batch.restype = C.c_int
        L.ksw2_hip_release_cached_memory.argtypes = []
        L.ksw2_hip_release_cached_memory.restype = C.c_int
        L.ksw2_hip_last_error.restype = C.c_int
        L.ksw2_hip_error_string.restype = C.c_char_p
        L.ksw2_hip_error_string.argtypes = [C.c_int]
        L.ksw2_hip_init.argtypes = [C.c_int]
        L.ksw2_hip_init.restype = C.c_int
        L.ksw2_hip_route.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int8, C.c_void_p,
                                     C.c_int8, C.c_int8, C.c_int8, C.c_int8, C.c_int, C.c_int, C.c_int]
        L.ksw2_hip_route.restype = C.c_int
        L.ksw2_hip_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(KswPair), C.c_int8,
                                           C.c_void_p, C.c_int8, C.c_int8, C.c_int8, C.c_int8]
        L.ksw2_hip_plan_create.restype = C.c_int
        L.ksw2_hip_plan_run.argtypes = [C.c_void_p]
        L.ksw2_hip_plan_run.restype = C.c_int
        L.ksw2_hip_plan_fetch.argtypes = [C.c_void_p, C.POINTER(KswExtz)]
        L.ksw2_hip_plan_fetch.restype = C.c_int
        L.ksw2_hip_plan_timing.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                           C.POINTER(C.c_double)]
        L.ksw2_hip_plan_timing.restype = C.c_int
        L.ksw2_hip_plan_graph.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ksw2_hip_plan_graph.restype = C.c_int
        L.ksw2_hip_plan_dirs.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
        L.ksw2_hip_plan_dirs.restype = C.c_int
        L.ksw2_hip_certify_batch.argtypes = [C.c_int, C.POINTER(CertifyTask), C.c_int32, C.c_int32, C.c_int32,
                                             C.c_int32, C.c_int32, C.c_double, C.POINTER(CertifyResult)]
        L.ksw2_hip_certify_batch.restype = C.c_int
        L.ksw2_hip_certify_free.argtypes = [C.c_int, C.POINTER(CertifyResult)]
        L.ksw2_hip_certify_free.restype = None
        L.ksw2_hip_window_batch.argtypes = [C.c_int, C.POINTER(WindowTask), C.c_int64, C.c_int32, C.c_int32,
                                            C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(WindowResult)]
        L.ksw2_hip_window_batch.restype = C.c_int
        L.ksw2_hip_window_free.argtypes = [C.c_int, C.POINTER(WindowResult)]
        L.ksw2_hip_window_free.restype = None
        L.ksw2_hip_footprint.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int8, C.c_void_p, C.c_int8,
                                         C.c_int8, C.c_int8, C.c_int8, C.c_int, C.c_int, C.c_int]
        L.ksw2_hip_footprint.restype = C.c_uint64
        L.ksw2_hip_server_try_submit.argtypes = [C.c_void_p, C.POINTER(KswPair), C.POINTER(KswExtz),
                                                 C.POINTER(C.c_uint64), C.POINTER(C.c_int)]
        L.ksw2_hip_server_try_submit.restype = C.c_int
        L.ksw2_hip_server_load.argtypes = [C.c_void_p, C.POINTER(HipLoad)]
        L.ksw2_hip_server_load.restype = C.c_int
        L.ksw2_hip_plan_destroy.argtypes = [C.c_void_p]
        L.ksw2_hip_plan_destroy.restype = None
        L.ksw2_hip_set_progress_callback.argtypes = [PROGRESS_CB, C.c_void_p, C.c_int]
        L.ksw2_hip_set_progress_callback.restype = None
        L.ksw2_hip_clear_progress_callback.argtypes = []
        L.ksw2_hip_clear_progress_callback.restype = None
        L.ksw2_hip_server_create.argtypes = [C.POINTER(C.c_void_p), C.c_int8, C.c_void_p, C.c_int8, C.c_int8,
                                             C.c_int8, C.c_int8, C.c_size_t]
        L.ksw2_hip_server_create.restype = C.c_int
        L.ksw2_hip_server_submit.argtypes = [C.c_void_p, C.POINTER(KswPair), C.POINTER(KswExtz),
                                             C.POINTER(C.c_uint64)]
        L.ksw2_hip_server_submit.restype = C.c_int
        L.ksw2_hip_server_poll.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.c_int,
                                           C.c_int]
        L.ksw2_hip_server_poll.restype = C.c_int
        L.ksw2_hip_server_poll_status.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                                  C.POINTER(C.c_int32), C.c_int, C.c_int]
        L.ksw2_hip_server_poll_status.restype = C.c_int
        L.ksw2_hip_server_destroy.argtypes = [C.c_void_p]
        L.ksw2_hip_server_destroy.restype = None
        _lib = L
    return _lib


# int (*)(void *arguments, int antidiagonal, int total_antidiagonals)
PROGRESS_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int)


class progress_callback:
    """Context manager installing ``fn(antidiagonal, total) -> bool`` as the
    calling thread's progress callback (``ksw2_hip_set_progress_callback``,
    same meaning as minimap2's ``ksw_set_progress_callback``): a False
    return stops the running call, which then reports ``stopped = 1``."""

    def __init__(self, fn, interval: int = 256):
        self.fn, self.interval, self.calls = fn, interval, 0

    def _call(self, _args, r, total):
        self.calls += 1
        return 1 if self.fn(r, total) else 0

    def __enter__(self):
        self._cb = PROGRESS_CB(self._call)
        lib().ksw2_hip_set_progress_callback(self._cb, None, self.interval)
        return self

    def __exit__(self, *exc):
        lib().ksw2_hip_clear_progress_callback()
        return False


def default_matrix(match: int = 0, mismatch: int = -6, m: int = 5) -> np.ndarray:
    """AnchorWave's matrix (alignSlidingWindow.cpp:247-256): ``match`` on the
    diagonal including N vs N, ``mismatch`` elsewhere."""
    mat = np.full((m, m), mismatch, dtype=np.int8)
    np.fill_diagonal(mat, match)
    return mat.reshape(-1)


def _u8(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8))


def extd2(query, target, mat=None, m=5, q=8, e=2, q2=75, e2=1, w=-1, zdrop=-1, end_bonus=0,
          flag=0, check=True) -> Result:
    """One call of ``ksw_extd2_hip`` (km = NULL)."""
    L = lib()
    qs, ts = _u8(query), _u8(target)
    mat = np.ascontiguousarray(np.asarray(default_matrix() if mat is None else mat, dtype=np.int8))
    ez = KswExtz()
    L.ksw_extd2_hip(None, len(qs), qs.ctypes.data, len(ts), ts.ctypes.data, m, mat.ctypes.data,
                    q, e, q2, e2, w, zdrop, end_bonus, flag, C.byref(ez))
    st = L.ksw2_hip_last_error()
    res = Result.from_ez(ez)
    if ez.cigar:
        _libc.free(C.cast(ez.cigar, C.c_void_p))
    if check and st != 0:
        raise KswError(f"ksw_extd2_hip: {STATUS.get(st, st)}")
    return res


def extd2_noalloc(query, target, mat=None, m=5, q=8, e=2, q2=75, e2=1, w=-1, zdrop=-1, end_bonus=0,
                  flag=0, cigar_capacity=None):
    """One call of ``ksw_extd2_hip_noalloc``: ``extd2`` with every device
    buffer in the library's fixed scratch array and the CIGAR written into a
    caller buffer of ``cigar_capacity`` words (default qlen + tlen).
    Returns ``(status, Result)``."""
    L = lib()
    qs, ts = _u8(query), _u8(target)
    mat = np.ascontiguousarray(np.asarray(default_matrix() if mat is None else mat, dtype=np.int8))
    cap = len(qs) + len(ts) if cigar_capacity is None else cigar_capacity
    buf = (C.c_uint32 * max(cap, 1))()
    ez = KswExtz()
    ez.cigar = C.cast(buf, C.POINTER(C.c_uint32))
    ez.m_cigar = cap
    st = L.ksw_extd2_hip_noalloc(len(qs), qs.ctypes.data, len(ts), ts.ctypes.data, m, mat.ctypes.data,
                                 q, e, q2, e2, w, zdrop, end_bonus, flag, C.byref(ez))
    assert C.addressof(ez.cigar.contents) == C.addressof(buf) and ez.m_cigar == cap, "CIGAR buffer replaced"
    return st, Result.from_ez(ez)


def extd2_singletrack(query, target, mat=None, m=5, q=8, e=2, q2=75, e2=1, w=-1, zdrop=-1, end_bonus=0,
                      flag=0, check=True) -> Result:
    """One call of ``ksw_extd2_singletrack_hip`` (km = NULL): ``extd2`` with
    KSW_EZ_SINGLETRACK, through its own entry point."""
    L = lib()
    qs, ts = _u8(query), _u8(target)
    mat = np.ascontiguousarray(np.asarray(default_matrix() if mat is None else mat, dtype=np.int8))
    ez = KswExtz()
    L.ksw_extd2_singletrack_hip(None, len(qs), qs.ctypes.data, len(ts), ts.ctypes.data, m, mat.ctypes.data,
                                q, e, q2, e2, w, zdrop, end_bonus, flag, C.byref(ez))
    st = L.ksw2_hip_last_error()
    res = Result.from_ez(ez)
    if ez.cigar:
        _libc.free(C.cast(ez.cigar, C.c_void_p))
    if check and st != 0:
        raise KswError(f"ksw_extd2_singletrack_hip: {STATUS.get(st, st)}")
    return res


class ChainAnchor(C.Structure):
    """``ksw2_chain_anchor_t`` (AlignmentChainAnchor)."""

    _fields_ = [("query_position", C.c_uint32), ("reference_position", C.c_uint32), ("length", C.c_uint16)]


class CertifyTask(C.Structure):
    """``ksw2_certify_task_t`` (include/ksw2_hip.h)."""

    _fields_ = [("query", C.c_char_p), ("target", C.c_char_p), ("qlen", C.c_int64), ("tlen", C.c_int64),
                ("initial_band", C.c_int64), ("maximum_band", C.c_int64),
                ("precomputed_optimal_score", C.POINTER(C.c_int64)),
                ("anchors", C.POINTER(ChainAnchor)), ("n_anchors", C.c_int64)]


class CertifyResult(C.Structure):
    """``ksw2_certify_result_t`` (include/ksw2_hip.h)."""

    _fields_ = [("status", C.c_int32), ("score", C.c_int64), ("optimal_score", C.c_int64),
                ("final_band_width", C.c_int64), ("traceback_attempts", C.c_uint64),
                ("cigar", C.POINTER(C.c_uint32)), ("n_cigar", C.c_int32),
                ("best_effort_cigar", C.POINTER(C.c_uint32)), ("n_best_effort", C.c_int32)]


CERT_COMPLETED, CERT_TIME_LIMIT, CERT_NOT_CERTIFIED, CERT_FAILED = 0, 1, 2, 3


@dataclass
class Certified:
    """One task's result of ``certify_batch`` (the fields of
    Ksw2CertifiedResult, alignments as CIGARs)."""

    status: int
    score: int
    optimal_score: int
    final_band_width: int
    traceback_attempts: int
    cigar: List[int]
    best_effort_cigar: List[int]


def certify_batch(tasks: Sequence[tuple], mismatch=-6, open1=-8, extend1=-2, open2=-75, extend2=-1,
                  max_seconds: float = 0.0) -> List[Certified]:
    """``ksw2_hip_certify_batch``: tasks are (query, target, initial_band,
    maximum_band[, precomputed_optimal_score[, anchors]]) with str or bytes
    sequences; anchors (None: no selection plan) are (query position,
    reference position, length) triples; the scoring is AnchorWave's
    penalties (negative), as alignScoreCertifiedKsw2 takes them."""
    L = lib()
    n = len(tasks)
    arr = (CertifyTask * max(n, 1))()
    keep = []
    for k, t in enumerate(tasks):
        qs = t[0].encode() if isinstance(t[0], str) else bytes(t[0])
        ts = t[1].encode() if isinstance(t[1], str) else bytes(t[1])
        keep += [qs, ts]
        arr[k].query, arr[k].target = qs, ts
        arr[k].qlen, arr[k].tlen = len(qs), len(ts)
        arr[k].initial_band, arr[k].maximum_band = int(t[2]), int(t[3])
        if len(t) > 4 and t[4] is not None:
            v = C.c_int64(int(t[4]))
            keep.append(v)
            arr[k].precomputed_optimal_score = C.pointer(v)
        if len(t) > 5 and t[5] is not None:
            an = (ChainAnchor * max(len(t[5]), 1))()
            for i, (aq, ar, al) in enumerate(t[5]):
                an[i].query_position, an[i].reference_position, an[i].length = aq, ar, al
            keep.append(an)
            arr[k].anchors, arr[k].n_anchors = C.cast(an, C.POINTER(ChainAnchor)), len(t[5])
    out = (CertifyResult * max(n, 1))()
    st = L.ksw2_hip_certify_batch(n, arr, mismatch, open1, extend1, open2, extend2, float(max_seconds), out)
    if st != 0:
        raise KswError(f"ksw2_hip_certify_batch: {STATUS.get(st, st)}")
    res = []
    for k in range(n):
        o = out[k]
        res.append(Certified(o.status, o.score, o.optimal_score, o.final_band_width, o.traceback_attempts,
                             [o.cigar[i] for i in range(o.n_cigar)],
                             [o.best_effort_cigar[i] for i in range(o.n_best_effort)]))
    L.ksw2_hip_certify_free(n, out)
    return res


class HipLoad(C.Structure):
    """``ksw2_hip_load_t`` (include/ksw2_hip.h)."""

    _fields_ = [("budget_bytes", C.c_uint64), ("held_bytes", C.c_uint64), ("queued_bytes", C.c_uint64),
                ("pending", C.c_int64), ("pending_cells", C.c_double), ("cells_per_s", C.c_double),
                ("seq_arena_bytes", C.c_uint64), ("resident_tasks", C.c_int64), ("staged_flights", C.c_int64)]


def footprint(query, target, mat=None, m=5, q=8, e=2, q2=75, e2=1, w=-1, zdrop=-1, flag=0) -> int:
    """``ksw2_hip_footprint``: device bytes the call holds while it runs."""
    qs, ts = _u8(query), _u8(target)
    mat = np.ascontiguousarray(np.asarray(default_matrix() if mat is None else mat, dtype=np.int8))
    return int(lib().ksw2_hip_footprint(len(qs), qs.ctypes.data, len(ts), ts.ctypes.data, m, mat.ctypes.data,
                                        q, e, q2, e2, w, zdrop, flag))


class WindowTask(C.Structure):
    """``ksw2_window_task_t`` (include/ksw2_hip.h)."""

    _fields_ = [("query", C.c_char_p), ("target", C.c_char_p), ("qlen", C.c_int64), ("tlen", C.c_int64)]


class WindowResult(C.Structure):
    """``ksw2_window_result_t`` (include/ksw2_hip.h)."""

    _fields_ = [("status", C.c_int32), ("score", C.c_int64), ("windows", C.c_int32),
                ("cigar", C.POINTER(C.c_uint32)), ("n_cigar", C.c_int32)]


@dataclass
class Windowed:
    """One task's result of ``window_batch``."""

    status: int
    score: int
    windows: int
    cigar: List[int]


def window_batch(tasks: Sequence[tuple], window: int, matching=0, mismatch=-6, open1=-8, extend1=-2, open2=-75,
                 extend2=-1) -> List[Windowed]:
    """``ksw2_hip_window_batch``: tasks are (query, target) str or bytes;
    alignSlidingWindow's window size and penalties."""
    L = lib()
    n = len(tasks)
    arr = (WindowTask * max(n, 1))()
    keep = []
    for k, t in enumerate(tasks):
        qs = t[0].encode() if isinstance(t[0], str) else bytes(t[0])
        ts = t[1].encode() if isinstance(t[1], str) else bytes(t[1])
        keep += [qs, ts]
        arr[k].query, arr[k].target, arr[k].qlen, arr[k].tlen = qs, ts, len(qs), len(ts)
    out = (WindowResult * max(n, 1))()
    st = L.ksw2_hip_window_batch(n, arr, int(window), matching, mismatch, open1, extend1, open2, extend2, out)
    if st != 0:
        raise KswError(f"ksw2_hip_window_batch: {STATUS.get(st, st)}")
    res = [Windowed(out[k].status, out[k].score, out[k].windows, [out[k].cigar[i] for i in range(out[k].n_cigar)])
           for k in range(n)]
    L.ksw2_hip_window_free(n, out)
    return res


def route(query, target, mat=None, m=5, q=8, e=2, q2=75, e2=1, w=-1, zdrop=-1, flag=0) -> int:
    """``ksw2_hip_route``: 0 when the call takes the unbanded fast kernel,
    else the reason mask; -1 when the reference returns before aligning."""
    qs, ts = _u8(query), _u8(target)
    mat = np.ascontiguousarray(np.asarray(default_matrix() if mat is None else mat, dtype=np.int8))
    return lib().ksw2_hip_route(len(qs), qs.ctypes.data, len(ts), ts.ctypes.data, m, mat.ctypes.data,
                                q, e, q2, e2, w, zdrop, flag)


def extd2_batch(pairs: Sequence[tuple], mat=None, m=5, q=8, e=2, q2=75, e2=1, w=-1, zdrop=-1,
                end_bonus=0, flag=0, check=True, timing: dict = None) -> List[Result]:
    """``ksw_extd2_hip_batch`` over ``pairs``: tuples (query, target) or
    (query, target, flag[, w[, zdrop[, end_bonus]]]); the keyword values are
    the defaults for fields a tuple leaves out.  Scoring is shared.  When
    ``timing`` is a dict, its "seconds" receives the wall time of the C call
    alone (uploads, device work, CIGAR delivery)."""
    import time
    L = lib()
    n = len(pairs)
    arr, keep = _pair_array(pairs, w=w, zdrop=zdrop, end_bonus=end_bonus, flag=flag)
    mat = np.ascontiguousarray(np.asarray(default_matrix() if mat is None else mat, dtype=np.int8))
    ezs = (KswExtz * max(n, 1))()
    t0 = time.perf_counter()
    st = L.ksw_extd2_hip_batch(n, arr, m, mat.ctypes.data, q, e, q2, e2, ezs)
    if timing is not None:
        timing["seconds"] = time.perf_counter() - t0
    out = []
    for k in range(n):
        out.append(Result.from_ez(ezs[k]))
        if ezs[k].cigar:
            _libc.free(C.cast(ezs[k].cigar, C.c_void_p))
    del keep
    if check and st != 0:
        raise KswError(f"ksw_extd2_hip_batch: {STATUS.get(st, st)}")
    return out


def _pair_array(pairs, w=-1, zdrop=-1, end_bonus=0, flag=0):
    n = len(pairs)
    keep = []
    arr = (KswPair * max(n, 1))()
    for k, pr in enumerate(pairs):
        qs, ts = _u8(pr[0]), _u8(pr[1])
        keep.append((qs, ts))
        f = pr[2] if len(pr) > 2 else flag
        pw = pr[3] if len(pr) > 3 else w
        pz = pr[4] if len(pr) > 4 else zdrop
        pb = pr[5] if len(pr) > 5 else end_bonus
        arr[k] = KswPair(qs.ctypes.data, ts.ctypes.data, len(qs), len(ts), pw, pz, pb, f)
    return arr, keep


class Plan:
    """A device-resident batch (``ksw2_hip_plan_*``): upload once, run many
    times, fetch results.  Used by the benchmark to time device work with the
    inputs already in HBM."""

    def __init__(self, pairs, mat=None, m=5, q=8, e=2, q2=75, e2=1, w=-1, zdrop=-1, end_bonus=0,
                 flag=0):
        L = lib()
        self.n = len(pairs)
        arr, keep = _pair_array(pairs, w, zdrop, end_bonus, flag)
        mat = np.ascontiguousarray(np.asarray(default_matrix() if mat is None else mat, dtype=np.int8))
        h = C.c_void_p()
        st = L.ksw2_hip_plan_create(C.byref(h), self.n, arr, m, mat.ctypes.data, q, e, q2, e2)
        if st != 0:
            raise KswError(f"ksw2_hip_plan_create: {STATUS.get(st, st)}")
        self.h = h

    def run(self):
        st = lib().ksw2_hip_plan_run(self.h)
        if st != 0:
            raise KswError(f"ksw2_hip_plan_run: {STATUS.get(st, st)}")

    def timing(self):
        dp, bt, cells = C.c_float(), C.c_float(), C.c_double()
        lib().ksw2_hip_plan_timing(self.h, C.byref(dp), C.byref(bt), C.byref(cells))
        return dp.value, bt.value, cells.value

    def graph_stats(self) -> tuple[int, int]:
        """(runs so far, runs that replayed the captured hipGraph)."""
        runs, rep = C.c_int(), C.c_int()
        st = lib().ksw2_hip_plan_graph(self.h, C.byref(runs), C.byref(rep))
        if st != 0:
            raise KswError(f"ksw2_hip_plan_graph: {STATUS.get(st, st)}")
        return runs.value, rep.value

    def fetch(self, timing: dict = None) -> List[Result]:
        """``ksw2_hip_plan_fetch`` into ``ksw_extz_t`` structs, as Results.
        When ``timing`` is a dict, its "seconds" receives the wall time of the
        C call alone (scores and CIGARs delivered into the structs)."""
        import time
        ezs = (KswExtz * max(self.n, 1))()
        t0 = time.perf_counter()
        st = lib().ksw2_hip_plan_fetch(self.h, ezs)
        if timing is not None:
            timing["seconds"] = time.perf_counter() - t0
        out = []
        for k in range(self.n):
            out.append(Result.from_ez(ezs[k]))
            if ezs[k].cigar:
                _libc.free(C.cast(ezs[k].cigar, C.c_void_p))
        if st != 0:
            raise KswError(f"ksw2_hip_plan_fetch: {STATUS.get(st, st)}")
        return out

    def stored_geometry(self, index: int) -> tuple[int, bool]:
        """Rows per lane of the stripe kernel that ran pair ``index`` and
        whether the pair is stored transposed (its query in the stripes and
        its target swept), from ``ksw2_hip_plan_dirs``."""
        r, tr = C.c_int(), C.c_int()
        st = lib().ksw2_hip_plan_dirs(self.h, index, None, 0, C.byref(r), C.byref(tr))
        if st != 0:
            raise KswError(f"ksw2_hip_plan_dirs: {STATUS.get(st, st)}")
        return r.value, bool(tr.value)

    def stored_transposed(self, index: int) -> bool:
        return self.stored_geometry(index)[1]

    def _stored(self, index: int, qlen: int, tlen: int, tracks: int):
        """Raw pass-layout bytes of a pair, the byte index of every stored
        cell (striped row i, swept column j) as a (kt, kq) array, and the
        bytes per lane of one track."""
        R, tr = self.stored_geometry(index)
        kq, kt = (tlen, qlen) if tr else (qlen, tlen)
        SR = 64 * R
        npass = ((kt + SR - 1) // SR + 1) // 2
        nsteps = kq + 127
        per_lane = 2 * R * tracks
        nbytes = npass * nsteps * 64 * per_lane
        buf = np.zeros(nbytes, dtype=np.uint8)
        r, t2 = C.c_int(), C.c_int()
        st = lib().ksw2_hip_plan_dirs(self.h, index, buf.ctypes.data, nbytes, C.byref(r), C.byref(t2))
        if st != 0:
            raise KswError(f"ksw2_hip_plan_dirs: {STATUS.get(st, st)}")
        i = np.arange(kt)[:, None]
        j = np.arange(kq)[None, :]
        s, rem = i // SR, i % SR
        p, h = s // 2, s % 2
        k, rho = rem // R, rem % R
        c = j + k + 64 * h
        idx = ((p * nsteps + c) * 64 + k) * per_lane + 4 * (rho // 2) + 2 * h + rho % 2
        return tr, buf, idx, 2 * R

    def dirs(self, index: int, qlen: int, tlen: int, raw: bool = False) -> np.ndarray:
        """Direction bytes of pair ``index`` as a (tlen, qlen) matrix (or the
        raw pass-layout buffer when ``raw``).  Stored layout, with i the
        striped index, j the swept one (the target and the query, or the
        query and the target for a transposed pair) and R the rows per lane
        (8 or 4; a stripe is 64R rows): stripe s = i // 64R is half h = s & 1
        of pass p = s >> 1; lane k = (i % 64R) // R, row rho = i % R; column
        j of that half is written at step c = j + k + 64h; byte
        ((p * (swept + 127) + c) * 64 + k) * 2R + 4 * (rho // 2) + 2h +
        rho % 2.  The bytes are the reference's direction bytes either way."""
        tr, buf, idx, _ = self._stored(index, qlen, tlen, 1)
        if raw:
            return buf
        m = buf[idx]
        return m.T if tr else m

    def tracks(self, index: int, qlen: int, tlen: int) -> np.ndarray:
        """Singletrack tracks of pair ``index`` (flag KSW_EZ_SINGLETRACK) as a
        (2, tlen, qlen) int8 array: p = H(i,j)-H(i,j-1) and p2 =
        H(i,j)-H(i-1,j).  Same cell order as ``dirs``, 4R bytes per lane and
        step (the negated stored-horizontal track, then the negated
        stored-vertical one, 2R bytes each; a transposed pair's are the
        reference's p2 and p)."""
        tr, buf, idx, half = self._stored(index, qlen, tlen, 2)
        raw = buf.view(np.int8)
        t0, t1 = -raw[idx].astype(np.int16), -raw[idx + half].astype(np.int16)
        return np.stack([t1.T, t0.T]) if tr else np.stack([t0, t1])

    def close(self):
        if self.h:
            lib().ksw2_hip_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving API (``ksw2_hip_server_*``): submit pairs as they arrive,
    collect results as they complete.  ``submit`` returns a ticket;
    ``poll`` returns ``[(ticket, Result), ...]`` for completed pairs."""

    def __init__(self, mat=None, m=5, q=8, e=2, q2=75, e2=1, device_bytes=0):
        self._L = lib()
        self._mat = np.ascontiguousarray(np.asarray(default_matrix() if mat is None else mat, dtype=np.int8))
        h = C.c_void_p()
        st = self._L.ksw2_hip_server_create(C.byref(h), m, self._mat.ctypes.data, q, e, q2, e2, device_bytes)
        if st != 0:
            raise KswError(f"ksw2_hip_server_create: {STATUS.get(st, st)}")
        self._h = h
        self._ez = {}   # ticket -> KswExtz (kept alive until completion)
        self.status = {}  # ticket -> status of the completed pair (0: the device path ran it)
        self.last_error = 0  # ksw2_hip_last_error() after the last poll: the server's first failure

    def submit(self, query, target, w=-1, zdrop=-1, end_bonus=0, flag=0) -> int:
        qs, ts = _u8(query), _u8(target)
        pr = KswPair(qs.ctypes.data, ts.ctypes.data, len(qs), len(ts), w, zdrop, end_bonus, flag)
        ez = KswExtz()
        tk = C.c_uint64()
        st = self._L.ksw2_hip_server_submit(self._h, C.byref(pr), C.byref(ez), C.byref(tk))
        if st != 0:
            raise KswError(f"ksw2_hip_server_submit: {STATUS.get(st, st)}")
        self._ez[tk.value] = ez
        return tk.value

    def try_submit(self, query, target, w=-1, zdrop=-1, end_bonus=0, flag=0):
        """``ksw2_hip_server_try_submit``: (ticket, 0) when queued, (None,
        retry_ms) when deferred."""
        qs, ts = _u8(query), _u8(target)
        pr = KswPair(qs.ctypes.data, ts.ctypes.data, len(qs), len(ts), w, zdrop, end_bonus, flag)
        ez = KswExtz()
        tk = C.c_uint64()
        retry = C.c_int()
        st = self._L.ksw2_hip_server_try_submit(self._h, C.byref(pr), C.byref(ez), C.byref(tk), C.byref(retry))
        if st == EAGAIN:
            return None, retry.value
        if st != 0:
            raise KswError(f"ksw2_hip_server_try_submit: {STATUS.get(st, st)}")
        self._ez[tk.value] = ez
        return tk.value, 0

    def load(self) -> dict:
        """``ksw2_hip_server_load``."""
        ld = HipLoad()
        st = self._L.ksw2_hip_server_load(self._h, C.byref(ld))
        if st != 0:
            raise KswError(f"ksw2_hip_server_load: {STATUS.get(st, st)}")
        return {k: getattr(ld, k) for k, _ in HipLoad._fields_}

    def poll(self, max_n: int = 1024, wait_ms: int = 0, results: bool = True):
        """Completed pairs: [(ticket, Result or None, latency_s), ...]."""
        if getattr(self, "_pbuf_n", 0) < max_n:
            self._pbuf = (C.c_uint64 * max_n)()
            self._plat = (C.c_double * max_n)()
            self._pst = (C.c_int32 * max_n)()
            self._pbuf_n = max_n
        buf, lat, pst = self._pbuf, self._plat, self._pst
        n = self._L.ksw2_hip_server_poll_status(self._h, buf, lat, pst, max_n, wait_ms)
        if n < 0:
            raise KswError(f"ksw2_hip_server_poll_status: {STATUS.get(n, n)}")
        self.last_error = self._L.ksw2_hip_last_error()
        out = []
        for k in range(n):
            self.status[buf[k]] = pst[k]
            ez = self._ez.pop(buf[k])
            out.append((buf[k], Result.from_ez(ez) if results else None, lat[k]))
            if ez.cigar:
                _libc.free(C.cast(ez.cigar, C.c_void_p))
        return out

    def pending(self) -> int:
        return len(self._ez)

    def close(self):
        if self._h:
            self._L.ksw2_hip_server_destroy(self._h)
            self._h = None
            for ez in self._ez.values():
                if ez.cigar:
                    _libc.free(C.cast(ez.cigar, C.c_void_p))
            self._ez.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
