"""Pins the restatement of AnchorWave's sliding-window alignment
(oracle/window_py.py), CPU only.  Its ksw_extd2 calls are the oracle's
(golden vectors pin them, SEMIGLOBAL_END included); these checks pin the
window chain around them:
- the rows rebuild both sequences and the score is their rescoring;
- a window that covers the interval gives the global alignment, whose
  score is the optimum of an independent textbook Gotoh DP;
- windows advance by the chosen end and the tails close the alignment;
- empty sides and getSubsequence's clipping.
"""
import numpy as np

import certify_py as CP
import oracle_py as O
import window_py as WP

PEN = dict(mis=-6, o1=-8, e1=-2, o2=-75, e2=-1)


def _dna(rng, n):
    return "".join("ACGT"[i] for i in rng.integers(0, 4, n))


def _mutate(rng, s, sub=0.06, indel=0.03):
    out = []
    for c in s:
        r = rng.random()
        if r < sub:
            out.append("ACGT"[int(rng.integers(0, 4))])
        elif r < sub + indel / 2:
            continue
        elif r < sub + indel:
            out.append(c + "ACGT"[int(rng.integers(0, 4))])
        else:
            out.append(c)
    return "".join(out)


def test_rows_rebuild_and_score_is_rescoring():
    rng = np.random.default_rng(91)
    for k in range(30):
        t = _dna(rng, int(rng.integers(40, 700)))
        q = _mutate(rng, t)
        r = WP.align(q, t, int(rng.choice([16, 50, 120])), **PEN)
        assert r.status == 0
        a, b = r.rows
        assert CP.reconstructs(a, q) and CP.reconstructs(b, t)
        assert r.score == CP.rescore(a, b, **PEN)
        assert (a, b) == CP.rows(r.cigar, q, t)  # the runs spell the same rows
        assert r.windows >= 1


def test_one_window_is_the_global_optimum():
    rng = np.random.default_rng(92)
    for k in range(20):
        t = _dna(rng, int(rng.integers(20, 200)))
        q = _mutate(rng, t)
        r = WP.align(q, t, 1000, **PEN)
        assert r.windows == 1
        qs, _ = CP.encode(q)
        ts, _ = CP.encode(t)
        assert r.score == O.gotoh2_global(qs, ts, q=8, e=2, q2=75, e2=1)


def test_windows_advance_and_tails():
    # identical sequences: every window ends on the diagonal, no tail
    s = "ACGT" * 50
    r = WP.align(s, s, 30, **PEN)
    assert r.cigar == [len(s) << 4] and r.score == 0
    assert r.windows == (len(s) + 30) // 31
    # a query much shorter than the target: the target's rest is one gap
    r = WP.align("ACGTACGTAC", "ACGTACGTAC" + "T" * 300, 20, **PEN)
    assert r.cigar[-1] & 0xF == 2 and r.status == 0
    assert CP.reconstructs(r.rows[1], "ACGTACGTAC" + "T" * 300)


def test_edge_cases():
    r = WP.align("", "ACGT", 2, **PEN)  # ratio 0: the whole-interval branch, no call result
    assert r.cigar == [] and r.score == 0 and r.windows == 1
    r = WP.align("", "ACGT" * 10, 1, **PEN)
    assert r.cigar == [] and r.score == 0
    assert WP.subsequence("ACGTA", 4, 9) == "TA" and WP.subsequence("ACGTA", 7, 9) == "A"
    assert WP.subsequence("ACGTA", 3, 1) == "ACG"


def test_tied_ends_take_the_target_end():
    """alignmentMinimap2WithArena continues from the query's end only when
    mqe > mte (alignSlidingWindow.cpp:392-398).  Worked by hand for the
    query TCG against the target TTACGGGG with windows of 3: the first
    window (TCG against TTAC) ties mqe = mte = -12 (query end at target row
    2, target end at query column 1), so it ends at the target's end (CIGAR
    1M2D1M, rows T--C / TTAC); the second window (G against GGGG) has mqe 0
    > mte -14 and ends at the query's end (1M); the target tail GGG is one
    deletion (-14).  Score -12 + 0 - 14 = -26."""
    r = WP._call("TCG", "TTAC", -6, -8, -2, -75, -1, WP.SEMIGLOBAL_END)
    assert (r.mqe, r.mqe_t, r.mte, r.mte_q) == (-12, 2, -12, 1)
    res = WP.align("TCG", "TTACGGGG", 3)
    assert (res.status, res.score, res.windows) == (0, -26, 2)
    assert res.rows == ("T--CG---", "TTACGGGG")
    assert res.cigar == [1 << 4 | 0, 2 << 4 | 2, 2 << 4 | 0, 3 << 4 | 2]
