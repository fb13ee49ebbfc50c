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


def _stop_after_first_poll():
    state = {"n": 0}

    def cb(r, total):
        state["n"] += 1
        return state["n"] <= 1   # allow anti-diagonal 0, stop at the next poll

    return state, cb


@pytest.mark.parametrize("kind", ["stripe", "band"])
def test_stop_midway_ends_the_fast_kernels_early(kind):
    """The stripe and band kernels sample the stop flag every 256 columns or
    anti-diagonals: a 30 kbp pair stopped after its first poll ends well
    before its full DP time, is reported stopped with no path, and the next
    call on the thread runs normally."""
    import time
    rng = np.random.default_rng(74)
    t = W.random_seq(rng, 30000)
    q = W.mutate(rng, t, 0.05, 0.01)
    w = -1 if kind == "stripe" else 4096
    assert ksw2.route(q, t, w=w) == (0 if kind == "stripe" else 1)
    ksw2.extd2(q[:2000], t[:2000], w=w)  # warm
    t0 = time.perf_counter()
    full = ksw2.extd2(q, t, w=w)
    t_full = time.perf_counter() - t0
    state, cb = _stop_after_first_poll()
    t0 = time.perf_counter()
    with ksw2.progress_callback(cb):
        g = ksw2.extd2(q, t, w=w, check=False)
    t_stop = time.perf_counter() - t0
    print(f"{kind}: full {t_full * 1e3:.1f} ms, stopped after {t_stop * 1e3:.1f} ms ({state['n']} polls)", flush=True)
    assert state["n"] >= 2, "the call finished before a second poll"
    assert g.stopped == 1 and g.n_cigar == 0
    assert t_stop < 0.5 * t_full, (t_stop, t_full)
    again = ksw2.extd2(q[:3000], t[:3000], w=w)
    assert same(again, O.run_oracle(q[:3000], t[:3000], w=w))
    assert full.stopped == 0 and full.n_cigar > 0


def test_stop_midway_batch_of_fast_pairs():
    """A batch of stripe-kernel pairs on teams of several waves stopped
    midway: every pair is reported stopped, and the plan's memory is fine
    for the next batch."""
    pairs = [p for p in W.uneven(seed=19, n=200) if len(p[0]) * len(p[1]) > 2e7][:24]
    state, cb = _stop_after_first_poll()
    with ksw2.progress_callback(cb):
        got = ksw2.extd2_batch(pairs, check=False)
    if state["n"] >= 2:
        assert all(g.stopped == 1 for g in got)
    small = random_pairs(75, 10, 1, 900)
    got = ksw2.extd2_batch(small)
    for (q, t), g in zip(small, got):
        assert same(g, O.run_oracle(q, t))
