"""Test infrastructure: a CPU restatement of AnchorWave's sliding-window
alignment, the checker for ``ksw2_hip_window_batch``.  Only ``tests/`` may
import it.

Every ksw_extd2 call goes through the oracle's C restatement (oracle_py),
which golden vectors from the reference pin (SEMIGLOBAL_END included).
The host logic restates src/myImportandFunction/alignSlidingWindow.cpp:
  * minimap2AlignmentWithArena :241-332 (whole interval, global)
  * alignmentMinimap2WithArena :334-431 (a window, SEMIGLOBAL_END, the
    better end at :392-398)
  * alignSlidingWindow :457-590 (the window chain, shifted penalties at
    :535-539, the tails at :559-580, the rescored result at :587-589)
  * getSubsequence src/impl/getSubsequence.cpp:191-211 (1-based, inclusive,
    clipped)
Pinned by tests/test_window_oracle.py (rows rebuild the sequences, a window
covering the interval is the global alignment, the score is the rescoring
of the rows, closed forms for empty sides).
"""
from dataclasses import dataclass, field
from typing import List

import certify_py as CP
import oracle_py as O

SEMIGLOBAL_END = 0x800


@dataclass
class Windowed:
    status: int = 0
    score: int = 0
    windows: int = 0
    cigar: List[int] = field(default_factory=list)
    rows: tuple = ("", "")


def _push(ops, op, length):
    if length <= 0:
        return
    if ops and (ops[-1] & 0xF) == op:
        ops[-1] += length << 4
    else:
        ops.append(length << 4 | op)


def _int8(v):
    return ((v + 128) & 0xFF) - 128


def _call(q: str, d: str, mis, o1, e1, o2, e2, flag):
    qs, gq = CP.encode(q)
    ts, gt = CP.encode(d)
    flag |= CP.GENERIC_SC if gq or gt else 0
    b = _int8(mis if mis < 0 else -mis)  # int8_t b = mismatch < 0 ? mismatch : -mismatch
    return O.run_oracle(qs, ts, mat=O.default_matrix(0, b), q=_int8(-o1), e=_int8(-e1), q2=_int8(-o2),
                        e2=_int8(-e2), w=-1, zdrop=-1, flag=flag)


def subsequence(s: str, start: int, end: int) -> str:
    """getSubsequence (src/impl/getSubsequence.cpp:191-211)."""
    if not s:
        return ""
    if start > end:
        start, end = end, start
    start = max(1, start)
    if start > len(s):
        start = end = len(s)
    elif end > len(s):
        end = len(s)
    return s[start - 1:end]


def align(dq: str, dd: str, window: int, ms=0, mis=-6, o1=-8, e1=-2, o2=-75, e2=-1) -> Windowed:
    """alignSlidingWindow (:457-590); alignments as rows and M/I/D runs."""
    res = Windowed()
    lq, ld = len(dq), len(dd)
    aq, ad, ops = [], [], []
    qstart = dstart = 1
    ratio = (ld / window if window else (float("inf") if ld else float("nan"))) * \
            (lq / window if window else (float("inf") if lq else float("nan")))
    if ratio <= 1:  # :489-496
        res.windows = 1
        if lq and ld:
            r = _call(dq, dd, mis, o1, e1, o2, e2, 0)
            a, b = CP.rows(r.cigar, dq, dd)
            aq.append(a), ad.append(b)
            for c in r.cigar:
                _push(ops, c & 0xF, c >> 4)
        dstart, qstart = ld + 1, lq + 1
    else:
        if e2 + ms < 0:  # :535-539
            wp = (mis + ms, o1 + ms, e1 + ms, o2 + ms, e2 + ms)
        else:
            wp = (mis + ms - 1, o1 + ms - 1, e1 + ms - 1, o2 + ms - 1, -1)
        while dstart <= ld and qstart <= lq:
            if window > 1073741824:
                res.status = 2
                return res
            q = subsequence(dq, qstart, min(qstart + window, lq))
            d = subsequence(dd, dstart, min(dstart + window, ld))
            r = _call(q, d, *wp, SEMIGLOBAL_END)
            res.windows += 1
            if r.mqe > r.mte:  # :392-398
                endt, endq = r.mqe_t + 1, len(q)
            else:
                endt, endq = len(d), r.mte_q + 1
            a, b = CP.rows(r.cigar, q, d)
            if len(a) - a.count("-") != endq or len(b) - b.count("-") != endt:  # :547-553
                res.status = 1
                return res
            aq.append(a), ad.append(b)
            for c in r.cigar:
                _push(ops, c & 0xF, c >> 4)
            qstart += endq
            dstart += endt
    c1 = ld - dstart  # :559-572
    if c1 >= 0:
        aq.append("-" * (c1 + 1)), ad.append(dd[dstart - 1:dstart + c1])
        _push(ops, 2, c1 + 1)
    c2 = lq - qstart
    if c2 >= 0:
        aq.append(dq[qstart - 1:qstart + c2]), ad.append("-" * (c2 + 1))
        _push(ops, 1, c2 + 1)
    a, b = "".join(aq), "".join(ad)
    res.rows = (a, b)
    res.score = CP.rescore(a, b, mis, o1, e1, o2, e2)
    res.cigar = ops
    return res
