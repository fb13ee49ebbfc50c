"""Progress callback / deadline (minimap2/ksw2_progress.c semantics):
the callback is polled on the host while a call runs; a zero return stops
it and the result reports stopped = 1."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu


def test_callback_that_never_stops_changes_nothing():
    pairs = random_pairs(71, 20, 1, 1200)
    for w in (-1, 90):
        with ksw2.progress_callback(lambda r, t: True) as pc:
            got = ksw2.extd2_batch(pairs, w=w)
        assert pc.calls >= 1
        for (q, t), g in zip(pairs, got):
            assert same(g, O.run_oracle(q, t, w=w))


def test_deadline_already_passed_stops_at_antidiagonal_zero():
    """The first poll is anti-diagonal 0: a callback that refuses it leaves
    every result as the reference's stop at r = 0 (reset values, stopped)."""
    pairs = random_pairs(72, 12, 1, 900)
    for w in (-1, 50):
        with ksw2.progress_callback(lambda r, t: False):
            got = ksw2.extd2_batch(pairs, w=w)
        for k, ((q, t), g) in enumerate(zip(pairs, got)):
            exp = O.run_oracle(q, t, w=w, stop_r=0)
            assert exp.stopped == 1
            assert same(g, exp), f"pair {k}: {diff(g, exp)}"


def test_stop_midway_banded_pair_backtracks_from_its_maximum():
    """A long banded pair (general kernel) stopped after its first poll: the
    reference's stopped result backtracks from (max_t, max_q), so the CIGAR
    spans exactly that prefix and rescores to ez->max.  A positive match
    score makes H grow along the diagonal, so the maximum is a real cell
    (with match 0, H never exceeds 0 and no maximum cell is recorded)."""
    rng = np.random.default_rng(73)
    t = W.random_seq(rng, 40000)
    q = W.mutate(rng, t, 0.05, 0.01)
    mat = O.default_matrix(2, -4)
    state = {"n": 0}

    def cb(r, total):
        state["n"] += 1
        return state["n"] <= 1   # allow anti-diagonal 0, stop at the next poll

    with ksw2.progress_callback(cb):
        g = ksw2.extd2(q, t, mat=mat, w=1500, check=False)
    assert g.stopped == 1 and g.zdropped == 0
    assert 0 <= g.max_t < len(t) and 0 <= g.max_q < len(q)
    ql, tl = O.cigar_lengths(g.cigar)
    assert (ql, tl) == (g.max_q + 1, g.max_t + 1)
    assert O.cigar_score(g.cigar, q[: g.max_q + 1], t[: g.max_t + 1], mat=mat) == g.max


def test_stop_midway_fast_pair_reports_stopped():
    """The stripe kernel has no stop point; a pair whose deadline passes
    while it runs is reported stopped with no path."""
    rng = np.random.default_rng(74)
    t = W.random_seq(rng, 30000)
    q = W.mutate(rng, t, 0.05, 0.01)
    state = {"n": 0}

    def cb(r, total):
        state["n"] += 1
        return state["n"] <= 1

    with ksw2.progress_callback(cb):
        g = ksw2.extd2(q, t, check=False)
    assert state["n"] >= 2, "the call finished before a second poll"
    assert g.stopped == 1 and g.n_cigar == 0
