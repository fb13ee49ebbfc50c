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
  * tryChainedKsw2Traceback :1056-1195 (the selection plan's
    certification anchors)
  * alignScoreCertifiedKsw2 :1197-1400, with or without a selection plan,
    and no deadline.
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


def cigar_of_rows(aq: str, ad: str) -> List[int]:
    """Runs of aligned (M), query-only (I) and reference-only (D) columns."""
    out: List[int] = []
    for a, b in zip(aq, ad):
        op = 1 if b == "-" else 2 if a == "-" else 0
        if out and (out[-1] & 0xF) == op:
            out[-1] += 1 << 4
        else:
            out.append(1 << 4 | op)
    return out


def chained(dq: str, dd: str, anchors, maximum_band: int, mis, o1, e1, o2, e2):
    """tryChainedKsw2Traceback (:1056-1195) without a deadline: the
    alignment through the plan's certification anchors (query position,
    reference position, length), each gap between them aligned by
    alignSlidingWindow_minimap2 at band min(max(lengths), maximum_band).
    Returns (completed, rows, score)."""
    if len(anchors) < 2 or maximum_band == 0:  # :1071-1074
        return False, None, 0
    aq, ad = [], []

    def segment(qs, qe, rs, re_):  # appendSegment :1081-1131
        ql, rl = qe - qs, re_ - rs
        if ql == 0 or rl == 0:
            if ql:
                aq.append(dq[qs:qe]), ad.append("-" * ql)
            elif rl:
                aq.append("-" * rl), ad.append(dd[rs:re_])
            return True
        if abs(ql - rl) > maximum_band or ql > 0x7FFFFFFF or rl > 0x7FFFFFFF:
            return False
        band = min(max(ql, rl), maximum_band)
        _score, cigar, valid = banded(dq[qs:qe], dd[rs:re_], band, mis, o1, e1, o2, e2)
        if not valid:
            return False
        a, b = rows(cigar, dq[qs:qe], dd[rs:re_])
        aq.append(a), ad.append(b)
        return True

    qpos = rpos = used = 0
    for aqp, arp, aln in anchors:  # :1133-1170
        if aln == 0 or aqp < qpos or arp < rpos or aqp + aln > len(dq) or arp + aln > len(dd):
            continue
        if dq[aqp:aqp + aln].upper() != dd[arp:arp + aln].upper():
            continue
        if not segment(qpos, aqp, rpos, arp):
            return False, None, 0
        aq.append(dq[aqp:aqp + aln]), ad.append(dd[arp:arp + aln])
        qpos, rpos, used = aqp + aln, arp + aln, used + 1
    if used < 2 or not segment(qpos, len(dq), rpos, len(dd)):  # :1171-1179
        return False, None, 0
    a, b = "".join(aq), "".join(ad)
    if not reconstructs(a, dq) or not reconstructs(b, dd):  # :1180-1185
        return False, None, 0
    return True, (a, b), rescore(a, b, mis, o1, e1, o2, e2)


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
            precomputed: Optional[int] = None, anchors=None) -> Certified:
    """alignScoreCertifiedKsw2 (:1197-1400) without a deadline; ``anchors``
    (None: no selection plan) are the plan's certification anchors;
    alignments as CIGARs."""
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
    have_best = False
    best: List[int] = []
    if anchors is not None:  # :1260-1289
        res.traceback_attempts += 1
        ok, rws, score = chained(dq, dd, anchors, max_band, mis, o1, e1, o2, e2)
        if ok:
            cig = cigar_of_rows(*rws)
            res.score = score
            if score == res.optimal_score:
                res.final_band_width, res.cigar, res.status = -2, cig, COMPLETED
                return res
            best, have_best = cig, True
    requested = max(diff, max(1, initial_band))  # :1292-1300
    doubled = requested * 2
    band = min(max_band, max(requested, doubled))

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
