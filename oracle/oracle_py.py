"""TEST INFRASTRUCTURE ONLY: ctypes access to the checkers.

* ``liboracle.so``        -- the scalar C restatement in ``ksw2_oracle.c``;
* ``_ref/libksw2ref.so``  -- the reference ``ksw_extd2_{sse,avx2,avx512}``
                             (minimap2/ksw2_extd2_sse.c, ksw2_extd2_avx.c)
                             compiled in place by ``oracle/Makefile``.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module.  The product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libksw2ref.so")

ISA_SSE = 16
ISA_AVX512 = 64


class KswExtz(C.Structure):
    """``ksw_extz_t`` (minimap2/ksw2.h:40-50)."""

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
        cig = tuple(ez.cigar[i] for i in range(ez.n_cigar)) if ez.n_cigar > 0 else ()
        return Result(ez.max, ez.zdropped, ez.max_q, ez.max_t, ez.mqe, ez.mqe_t, ez.mte,
                      ez.mte_q, ez.score, ez.n_cigar, ez.reach_end, ez.stopped, cig)

    def as_dict(self) -> dict:
        d = dict(self.__dict__)
        d["cigar"] = list(self.cigar)
        return d


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]

_SIG = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int8, C.c_void_p,
        C.c_int8, C.c_int8, C.c_int8, C.c_int8, C.c_int, C.c_int, C.c_int, C.c_int,
        C.POINTER(KswExtz)]

_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE_SO)
        lib.oracle_ksw_extd2.argtypes = _SIG[1:] + [C.c_int, C.c_void_p, C.c_int]
        lib.oracle_ksw_extd2.restype = None
        lib.oracle_cigar_score.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                           C.c_void_p, C.c_int8, C.c_void_p, C.c_int8,
                                           C.c_int8, C.c_int8, C.c_int8]
        lib.oracle_gotoh2_global.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                             C.c_int8, C.c_void_p, C.c_int8, C.c_int8,
                                             C.c_int8, C.c_int8]
        lib.oracle_range_enable.argtypes = [C.c_int]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        for name in ("ksw_extd2_sse", "ksw_extd2_avx2", "ksw_extd2_avx512"):
            fn = getattr(lib, name)
            fn.argtypes = _SIG
            fn.restype = None
        _ref = lib
    return _ref


def _u8(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8))


def _i8(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int8))


def default_matrix(match: int = 0, mismatch: int = -6, m: int = 5) -> np.ndarray:
    """AnchorWave's 5x5 matrix (alignSlidingWindow.cpp:247-256): match on the
    diagonal (N vs N included), mismatch elsewhere."""
    mat = np.full((m, m), mismatch, dtype=np.int8)
    np.fill_diagonal(mat, match)
    return mat.reshape(-1)


def run_oracle(query, target, mat=None, m=5, q=8, e=2, q2=75, e2=1, w=-1, zdrop=-1,
               end_bonus=0, flag=0, isa=ISA_AVX512, want_dirs=False, stop_r=-1):
    """Runs the C restatement; returns ``Result`` (and the (tlen, qlen)
    direction-byte matrix of in-band cells when ``want_dirs``; with
    KSW_EZ_SINGLETRACK a (2, tlen, qlen) int8 array of the two difference
    tracks instead, p = H(i,j)-H(i,j-1) and p2 = H(i,j)-H(i-1,j))."""
    lib = oracle_lib()
    qs, ts = _u8(query), _u8(target)
    mat = _i8(default_matrix() if mat is None else mat)
    ez = KswExtz()
    dirs = None
    dptr = None
    if want_dirs and (flag & 0x1000) and isa == ISA_AVX512:
        dirs = np.zeros((2, len(ts), len(qs)), dtype=np.int8)
        dptr = dirs.ctypes.data
    elif want_dirs:
        dirs = np.zeros((len(ts), len(qs)), dtype=np.uint8)
        dptr = dirs.ctypes.data
    lib.oracle_ksw_extd2(len(qs), qs.ctypes.data, len(ts), ts.ctypes.data, m, mat.ctypes.data,
                         q, e, q2, e2, w, zdrop, end_bonus, flag, C.byref(ez), isa, dptr, stop_r)
    res = Result.from_ez(ez)
    if ez.cigar:
        _libc.free(C.cast(ez.cigar, C.c_void_p))
    return (res, dirs) if want_dirs else res


def run_ref(query, target, mat=None, m=5, q=8, e=2, q2=75, e2=1, w=-1, zdrop=-1,
            end_bonus=0, flag=0, isa="avx512"):
    """Runs the compiled reference ``ksw_extd2_<isa>`` with km == NULL."""
    lib = ref_lib()
    fn = getattr(lib, "ksw_extd2_" + isa)
    qs, ts = _u8(query), _u8(target)
    mat = _i8(default_matrix() if mat is None else mat)
    ez = KswExtz()
    fn(None, len(qs), qs.ctypes.data, len(ts), ts.ctypes.data, m, mat.ctypes.data, q, e, q2,
       e2, w, zdrop, end_bonus, flag, C.byref(ez))
    res = Result.from_ez(ez)
    if ez.cigar:
        _libc.free(C.cast(ez.cigar, C.c_void_p))
    return res


def cigar_score(cigar: Sequence[int], query, target, mat=None, m=5, q=8, e=2, q2=75, e2=1) -> int:
    lib = oracle_lib()
    qs, ts = _u8(query), _u8(target)
    mat = _i8(default_matrix() if mat is None else mat)
    cg = np.ascontiguousarray(np.asarray(cigar, dtype=np.uint32))
    return lib.oracle_cigar_score(len(cg), cg.ctypes.data, len(qs), qs.ctypes.data, len(ts),
                                  ts.ctypes.data, m, mat.ctypes.data, q, e, q2, e2)


def gotoh2_global(query, target, mat=None, m=5, q=8, e=2, q2=75, e2=1) -> int:
    lib = oracle_lib()
    qs, ts = _u8(query), _u8(target)
    mat = _i8(default_matrix() if mat is None else mat)
    return lib.oracle_gotoh2_global(len(qs), qs.ctypes.data, len(ts), ts.ctypes.data, m,
                                    mat.ctypes.data, q, e, q2, e2)


class RangeProbe:
    """Context manager recording the int32 extremes of every int8 intermediate."""

    def __enter__(self):
        oracle_lib().oracle_range_enable(1)
        return self

    def __exit__(self, *exc):
        lib = oracle_lib()

        class R(C.Structure):
            _fields_ = [("lo", C.c_int32 * 16), ("hi", C.c_int32 * 16), ("wraps", C.c_int64),
                        ("cap_binds", C.c_int64)]

        r = R()
        lib.oracle_range_get(C.byref(r))
        lib.oracle_range_enable(0)
        self.lo = list(r.lo)
        self.hi = list(r.hi)
        self.wraps = r.wraps
        self.cap_binds = r.cap_binds
        return False


def cigar_str(cigar: Sequence[int]) -> str:
    return "".join(f"{c >> 4}{'MID'[c & 0xf] if (c & 0xf) < 3 else '?'}" for c in cigar)


def cigar_lengths(cigar: Sequence[int]):
    """(query bases, target bases) consumed by a CIGAR."""
    ql = sum(c >> 4 for c in cigar if (c & 0xf) in (0, 1))
    tl = sum(c >> 4 for c in cigar if (c & 0xf) in (0, 2))
    return ql, tl
