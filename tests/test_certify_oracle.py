"""Pins the restatement of AnchorWave's score-certified banding
(oracle/certify_py.py), CPU only.  Its ksw_extd2 calls are the oracle's
(pinned by the golden vectors); these checks pin the host logic around them:
- a certified alignment rescores to the optimum of an independent textbook
  two-piece Gotoh DP (oracle_gotoh2_global), and its rows rebuild both
  sequences;
- the band schedule (twice the first band, doubling to the cap, the skipped
  first band last) by its closed form, on tasks that can never certify;
- the reference's edge cases: an empty side, a cap below the length
  difference, a KSW_NEG_INF optimum, case-sensitive rescoring.
"""
from __future__ import annotations

import numpy as np
import pytest

import certify_py as CP
import oracle_py as O

PEN = dict(mis=-6, o1=-8, e1=-2, o2=-75, e2=-1)
KS = dict(q=8, e=2, q2=75, e2=1)


def _dna(rng, n):
    return "".join("ACGT"[i] for i in rng.integers(0, 4, n))


def _mutate(rng, s, sub=0.08, indel=0.03):
    out = []
    for c in s:
        r = rng.random()
        if r < sub:
            out.append("ACGT"[(("ACGT".index(c)) + int(rng.integers(1, 4))) % 4])
        elif r < sub + indel / 2:
            continue
        elif r < sub + indel:
            out.append(c)
            out.append("ACGT"[int(rng.integers(0, 4))])
        else:
            out.append(c)
    return "".join(out)


def test_certified_alignments_are_optimal():
    rng = np.random.default_rng(71)
    done = 0
    for k in range(40):
        t = _dna(rng, int(rng.integers(20, 160)))
        q = _mutate(rng, t)
        r = CP.certify(q, t, int(rng.integers(0, 20)), 10_000, **PEN)
        assert r.status == CP.COMPLETED, (k, r)
        qs, _ = CP.encode(q)
        ts, _ = CP.encode(t)
        opt = O.gotoh2_global(qs, ts, **KS)
        assert r.optimal_score == opt and r.score == opt, (k, r.score, opt)
        aq, ad = CP.rows(r.cigar, q, t)
        assert CP.reconstructs(aq, q) and CP.reconstructs(ad, t)
        assert CP.rescore(aq, ad, **PEN) == opt
        done += 1
    assert done == 40


def _bands_tried(monkeypatch, q, t, initial, maximum):
    """Runs a certification that never certifies (a precomputed optimum
    above any alignment's score) and returns the bands of its banded calls,
    in order."""
    seen = []
    original = CP.banded

    def record(dq, dd, band, *a):
        seen.append(band)
        return original(dq, dd, band, *a)

    monkeypatch.setattr(CP, "banded", record)
    r = CP.certify(q, t, initial, maximum, precomputed=1, **PEN)
    monkeypatch.setattr(CP, "banded", original)
    return r, seen


# (initial band, maximum band, target length, query length) -> the bands
# alignScoreCertifiedKsw2 tries (alignSlidingWindow.cpp:1292-1394), worked
# out by hand: cap = min(max(qlen, tlen), maximum); requested =
# max(|qlen - tlen|, max(1, initial)); the first band is min(cap, 2 *
# requested), then min(cap, max(b + 1, 2b)) until the cap, then the skipped
# requested band when requested < min(cap, 2 * requested).
SCHEDULES = [
    ((3, 64, 120, 120), [6, 12, 24, 48, 64, 3], 64),
    ((0, 200, 120, 120), [2, 4, 8, 16, 32, 64, 120, 1], 120),
    ((10, 10_000, 120, 120), [20, 40, 80, 120, 10], 120),
    ((7, 7, 120, 120), [7], 7),
    ((50, 51, 120, 120), [51, 50], 51),
    ((0, 1000, 100, 130), [60, 120, 130, 30], 130),   # |qlen - tlen| = 30 drives the first band
    ((40, 45, 100, 130), [45, 40], 45),               # cap between requested and 2x: the cap, then requested
    ((45, 45, 100, 130), [45], 45),                   # requested == cap: no skipped band
]


@pytest.mark.parametrize("args,bands,final", SCHEDULES)
def test_band_schedule_literal(args, bands, final, monkeypatch):
    initial, maximum, tl, ql = args
    rng = np.random.default_rng(72)
    t = _dna(rng, tl)
    q = _dna(rng, ql)
    r, seen = _bands_tried(monkeypatch, q, t, initial, maximum)
    assert seen == bands, (args, seen)
    assert r.status == CP.NOT_CERTIFIED and r.traceback_attempts == len(bands)
    assert r.final_band_width == final
    assert r.best_effort_cigar, "a valid banded alignment is kept"
    aq, ad = CP.rows(r.best_effort_cigar, q, t)
    assert CP.rescore(aq, ad, **PEN) == r.score


def test_edge_cases():
    # one side empty: all gap, closed form
    r = CP.certify("", "ACGTA", 5, 5, **PEN)
    assert (r.status, r.score, r.final_band_width, r.traceback_attempts) == (CP.COMPLETED, -18, 0, 1)
    assert r.cigar == [5 << 4 | 2]
    r = CP.certify("ACG", "", 5, 5, **PEN)
    assert r.score == max(-8 - 2 * 3, -75 - 3) and r.cigar == [3 << 4 | 1]
    r = CP.certify("", "", 5, 5, **PEN)
    assert r.status == CP.COMPLETED and r.cigar == [] and r.score == -8
    # a cap below the length difference, or negative bands: not certified, nothing run
    for a, b in [(5, 2), (-1, 50), (5, -1)]:
        r = CP.certify("ACGTACGTAC", "ACGTA", a, b, **PEN)
        assert (r.status, r.traceback_attempts, r.final_band_width) == (CP.NOT_CERTIFIED, 0, -1)
    # a KSW_NEG_INF optimum fails
    r = CP.certify("ACGT", "ACGT", 1, 10, precomputed=CP.KSW_NEG_INF, **PEN)
    assert r.status == CP.FAILED
    # rescoring compares characters: lower case against upper case is a
    # mismatch there (ksw_extd2 sees equal codes), so nothing certifies
    r = CP.certify("acgtacgtac", "ACGTACGTAC", 1, 100, **PEN)
    assert r.optimal_score == 0 and r.status == CP.NOT_CERTIFIED and r.score == -60
    # a wildcard takes the generic matrix (N on its diagonal scores 0, N
    # against a base the mismatch), and the rescoring agrees
    r = CP.certify("ACGNNACGT", "ACGNNACGT", 1, 100, **PEN)
    assert r.status == CP.COMPLETED and r.optimal_score == 0
    r = CP.certify("ACGNNACGT", "ACGTTACGT", 1, 100, **PEN)
    assert r.status == CP.COMPLETED and r.optimal_score == -12


def test_planted_mistakes_are_caught(monkeypatch):
    """A gap piece dropped from the rescoring stops certification (its
    score no longer reaches the Gotoh optimum); a schedule without the
    skipped first band attempts one band fewer than the reference."""
    original = CP.rescore

    def one_piece(aq, ad, mis, o1, e1, o2, e2):
        return original(aq, ad, mis, o1, e1, o1, e1)

    rng = np.random.default_rng(73)
    t = _dna(rng, 500)
    q = t[:150] + t[300:]  # a 150 bp deletion: the long-gap piece is the optimum
    assert CP.certify(q, t, 1, 10_000, **PEN).status == CP.COMPLETED
    monkeypatch.setattr(CP, "rescore", one_piece)
    assert CP.certify(q, t, 1, 10_000, **PEN).status == CP.NOT_CERTIFIED
    monkeypatch.setattr(CP, "rescore", original)
    # a schedule that stops at the cap without the skipped first band, or
    # doubles from the requested band instead of starting at twice it, is
    # not the literal list
    rng2 = np.random.default_rng(72)
    t2, q2 = _dna(rng2, 120), _dna(rng2, 120)
    _r, seen = _bands_tried(monkeypatch, q2, t2, 3, 64)
    assert seen[:-1] != [3, 6, 12, 24, 48, 64] and seen != [6, 12, 24, 48, 64]


def _runs_of_matches(cigar, q, t, min_len):
    """Exact-match runs of an alignment, as (query position, reference
    position, length) anchors."""
    out, qp, tp = [], 0, 0
    for c in cigar:
        ln, op = c >> 4, c & 0xF
        if op == 0:
            k = 0
            while k < ln:
                if q[qp + k] == t[tp + k]:
                    s = k
                    while k < ln and q[qp + k] == t[tp + k]:
                        k += 1
                    if k - s >= min_len:
                        out.append((qp + s, tp + s, k - s))
                else:
                    k += 1
            qp, tp = qp + ln, tp + ln
        elif op == 1:
            qp += ln
        else:
            tp += ln
    return out


def test_chained_traceback_through_optimal_anchors_is_optimal():
    """Anchors taken from an optimal alignment's exact-match runs lie on an
    optimal path and no gap crosses them, so the chained alignment (each gap
    between anchors aligned with a band covering it) scores the independent
    Gotoh optimum: the task completes at the chained attempt (band -2, one
    traceback), and its rows rebuild both sequences."""
    rng = np.random.default_rng(81)
    done = 0
    for k in range(24):
        t = _dna(rng, int(rng.integers(200, 700)))
        q = _mutate(rng, t)
        base = CP.certify(q, t, 0, 10_000, **PEN)
        assert base.status == CP.COMPLETED
        anchors = _runs_of_matches(base.cigar, q, t, 12)
        if len(anchors) < 2:
            continue
        qs, _ = CP.encode(q)
        ts, _ = CP.encode(t)
        opt = O.gotoh2_global(qs, ts, **KS)
        r = CP.certify(q, t, 0, 10_000, anchors=anchors, **PEN)
        assert r.status == CP.COMPLETED and r.final_band_width == -2 and r.traceback_attempts == 1, (k, r)
        assert r.score == opt == r.optimal_score
        aq, ad = CP.rows(r.cigar, q, t)
        assert CP.reconstructs(aq, q) and CP.reconstructs(ad, t) and CP.rescore(aq, ad, **PEN) == opt
        done += 1
    assert done >= 15


def test_chained_traceback_literal_cases():
    """Hand-checked cases of tryChainedKsw2Traceback (alignSlidingWindow.cpp:
    1056-1195): the anchors' runs merge with the gaps' M runs; an anchor that
    does not match (case-insensitively) or starts before the previous
    anchor's end is skipped; fewer than two usable anchors, or a gap whose
    length difference exceeds the band, makes the chain unavailable and the
    banded loop runs (one more traceback attempt)."""
    q = t = "AAAACCCCGGGGTTTT"
    r = CP.certify(q, t, 2, 64, anchors=[(0, 0, 4), (8, 8, 4)], **PEN)
    assert (r.status, r.final_band_width, r.cigar, r.traceback_attempts, r.score) == \
        (CP.COMPLETED, -2, [16 << 4 | 0], 1, 0)
    # the overlapping (2, 2, 4) and the mismatching (9, 10, 3) are skipped
    r = CP.certify(q, t, 2, 64, anchors=[(0, 0, 4), (2, 2, 4), (9, 10, 3), (12, 12, 4)], **PEN)
    assert (r.status, r.final_band_width, r.traceback_attempts) == (CP.COMPLETED, -2, 1)
    # anchors match case-insensitively, but the rescoring compares
    # characters: 16 mismatches against a lower-case copy
    ok, rws, sc = CP.chained(q, t.lower(), [(0, 0, 4), (12, 12, 4)], 64, **PEN)
    assert ok and sc == 16 * -6 and rws == (q, t.lower())
    # one usable anchor: unavailable, then the banded loop certifies at band 4
    r = CP.certify(q, t, 2, 64, anchors=[(0, 0, 4), (2, 3, 4)], **PEN)
    assert (r.status, r.final_band_width, r.traceback_attempts) == (CP.COMPLETED, 4, 2)
    # a gap of query 8 against reference 0 after (4, 4): difference 8 > maximum band 6
    q2 = "AAAACCCCGGGGTTTT"
    t2 = "AAAATTTT"
    ok, _rows, _s = CP.chained(q2, t2, [(0, 0, 4), (12, 4, 4)], 6, **PEN)
    assert ok  # one side empty: a pure gap, no DP
    ok, _rows, _s = CP.chained(q2 + "A", t2 + "CC", [(0, 0, 4), (12, 4, 4)], 6, **PEN)
    assert ok
    ok, _rows, _s = CP.chained("AAAA" + "CGCGCGCGCGCG" + "TTTT", "AAAA" + "CG" + "TTTT", [(0, 0, 4), (16, 6, 4)],
                               6, **PEN)
    assert not ok  # 12 against 2: difference 10 > 6
