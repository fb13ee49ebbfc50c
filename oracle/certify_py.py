"""Test infrastructure: a CPU restatement of AnchorWave's score-certified
banded KSW2 alignment, the checker for ``ksw2_hip_certify_batch``.  Only
``tests/`` may import it.

Every ksw_extd2 call goes through the oracle's C restatement (oracle_py),
which golden vectors from the reference pin.  The host logic restates
src/myImportandFunction/alignSlidingWindow.cpp:
  * encodeKswSequence :644-660 and the adapter's table :820-838
  * gappedAlignmentReconstructs :662-675
  * scoreGappedTwoPieceAffineAlignment :677-721
  * ksw2ScoreOnly :723-796
  * alignSlidingWindow_minimap2 :800-890 (CIGAR -> rows, rescoring)
  * alignScoreCertifiedKsw2 :1197-1400 with selectionPlan == NULL
    (no chained traceback) and no deadline.
Pinned by tests/test_certify_oracle.py: a certified alignment's rescored
score equals an independent Gotoh optimum, the band schedule matches its
closed form, and the rows rebuild both sequences.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

import oracle_py as O

COMPLETED, TIME_LIMIT, NOT_CERTIFIED, FAILED = 0, 1, 2, 3
KSW_NEG_INF = -0x40000000
GENERIC_SC, SCORE_ONLY = 0x04, 0x01


def encode(seq: str):
    """encodeKswSequence (alignSlidingWindow.cpp:644-660): codes and whether
    a wildcard (code 4) occurs."""
    table = {"A": 0, "a": 0, "C": 1, "c": 1, "G": 2, "g": 2, "T": 3, "t": 3}
    codes = np.array([table.get(c, 4) for c in seq], dtype=np.uint8)
    return codes, bool((codes == 4).any())


def matrix(mismatching_penalty: int) -> np.ndarray:
    b = mismatching_penalty if mismatching_penalty < 0 else -mismatching_penalty
    return O.default_matrix(0, b)


def reconstructs(row: str, seq: str) -> bool:
    """gappedAlignmentReconstructs (:662-675)."""
    s = row.replace("-", "")
    return s == seq


def rescore(aq: str, ad: str, mis, o1, e1, o2, e2) -> int:
    """scoreGappedTwoPieceAffineAlignment (:677-721)."""
    assert len(aq) == len(ad)
    score, col = 0, 0
    while col < len(aq):
        qg, dg = aq[col] == "-", ad[col] == "-"
        assert not (qg and dg)
        if qg or dg:
            end = col + 1
            while end < len(aq) and (aq[end] == "-") == qg and (ad[end] == "-") != qg:
                end += 1
            length = end - col
            score += max(o1 + e1 * length, o2 + e2 * length)
            col = end
            continue
        if aq[col] != ad[col]:
            score += mis
        col += 1
    return score


def rows(cigar, dq: str, dd: str):
    """The gapped rows alignSlidingWindow_minimap2 builds (:852-870)."""
    aq, ad, qp, tp = [], [], 0, 0
    for c in cigar:
        ln, op = c >> 4, c & 0xF
        if op in (0, 7, 8):
            aq.append(dq[qp:qp + ln]), ad.append(dd[tp:tp + ln])
            qp += ln
            tp += ln
        elif op == 1:
            aq.append(dq[qp:qp + ln]), ad.append("-" * ln)
            qp += ln
        elif op == 2:
            aq.append("-" * ln), ad.append(dd[tp:tp + ln])
            tp += ln
    return "".join(aq), "".join(ad)


def score_only(dq: str, dd: str, mis, o1, e1, o2, e2) -> int:
    """ksw2ScoreOnly (:723-796) on non-empty inputs."""
    q, gq = encode(dq)
    t, gt = encode(dd)
    flag = SCORE_ONLY | (GENERIC_SC if gq or gt else 0)
    r = O.run_oracle(q, t, mat=matrix(mis), q=-o1, e=-e1, q2=-o2, e2=-e2, w=-1, zdrop=-1, flag=flag)
    return r.score


def banded(dq: str, dd: str, band: int, mis, o1, e1, o2, e2):
    """alignSlidingWindow_minimap2 (:800-890): (rescored score, cigar, valid)."""
    q, gq = encode(dq)
    t, gt = encode(dd)
    flag = GENERIC_SC if gq or gt else 0
    r = O.run_oracle(q, t, mat=matrix(mis), q=-o1, e=-e1, q2=-o2, e2=-e2, w=band, zdrop=-1, flag=flag)
    aq, ad = rows(r.cigar, dq, dd)
    valid = len(aq) == len(ad) and reconstructs(aq, dq) and reconstructs(ad, dd)
    return rescore(aq, ad, mis, o1, e1, o2, e2), list(r.cigar), valid


@dataclass
class Certified:
    status: int = FAILED
    score: int = 0
    optimal_score: int = 0
    final_band_width: int = -1
    traceback_attempts: int = 0
    cigar: List[int] = field(default_factory=list)
    best_effort_cigar: List[int] = field(default_factory=list)


def certify(dq: str, dd: str, initial_band: int, maximum_band: int, mis=-6, o1=-8, e1=-2, o2=-75, e2=-1,
            precomputed: Optional[int] = None) -> Certified:
    """alignScoreCertifiedKsw2 (:1197-1400), selectionPlan == NULL, no
    deadline; alignments as CIGARs."""
    res = Certified()
    minlen = min(len(dq), len(dd))
    diff = abs(len(dq) - len(dd))
    if minlen == 0:  # :1216-1232
        length = max(len(dq), len(dd))
        res.score = max(o1 + e1 * length, o2 + e2 * length)
        res.optimal_score, res.final_band_width, res.traceback_attempts = res.score, 0, 1
        res.status = COMPLETED
        if length:
            res.cigar = [length << 4 | (1 if dq else 2)]
        return res
    if initial_band < 0 or maximum_band < 0 or maximum_band < diff:  # :1233-1238
        res.status = NOT_CERTIFIED
        return res
    res.optimal_score = precomputed if precomputed is not None else score_only(dq, dd, mis, o1, e1, o2, e2)
    if res.optimal_score == KSW_NEG_INF:
        res.status = FAILED
        return res
    max_band = min(max(len(dq), len(dd)), maximum_band)  # :1257-1259
    requested = max(diff, max(1, initial_band))  # :1292-1300
    doubled = requested * 2
    band = min(max_band, max(requested, doubled))
    have_best = False
    best: List[int] = []

    def attempt(b):
        nonlocal have_best, best
        res.traceback_attempts += 1
        score, cigar, valid = banded(dq, dd, b, mis, o1, e1, o2, e2)
        if score == res.optimal_score and valid:
            res.score, res.final_band_width, res.cigar, res.status = score, b, cigar, COMPLETED
            return True
        if valid and (not have_best or score > res.score):
            res.score, best, have_best = score, cigar, True
        return False

    while diff <= band <= max_band:  # :1301-1345
        if attempt(band):
            return res
        if band == max_band:
            break
        nxt = min(max_band, max(band + 1, band * 2))
        if nxt <= band:
            break
        band = nxt
    if requested < min(max_band, doubled):  # :1347-1394
        if attempt(requested):
            res.final_band_width = requested
            return res
    res.final_band_width = band
    res.status = NOT_CERTIFIED
    res.best_effort_cigar = best if have_best else []
    return res
