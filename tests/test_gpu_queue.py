"""Persistent one-wave workers over a device work queue (KSW2_HIP_QUEUE=1, or
=2 for launches with more pairs than the device holds workers; the default
is one workgroup per pair): workers claim pairs from the launch's
longest-first list with a device-scope atomic add until it is exhausted.  More pairs than the
device holds waves (each worker takes several), every flag group, repeated
runs (graph replays), and the same pairs launched one workgroup per pair
(KSW2_HIP_QUEUE=0) all give the reference's results."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu


def _check(got, exp, tag):
    bad = [f"{tag} pair {k}: {diff(g, r)}" for k, (g, r) in enumerate(zip(got, exp)) if not same(g, r)]
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(exp)} differ)"


def test_more_pairs_than_workers(monkeypatch):
    """9,000 short pairs: more than the 4,096 one-wave workers the device
    holds, so workers claim again after their first pair."""
    print("queue: 9000 short pairs", flush=True)
    monkeypatch.setenv("KSW2_HIP_QUEUE", "2")
    pairs = random_pairs(3101, 9000, 20, 600)
    exp = [O.run_oracle(q, t) for q, t in pairs]
    plan = ksw2.Plan(pairs)
    for rep in range(3):
        plan.run()
        _check(plan.fetch(), exp, f"run {rep}")
    plan.close()


@pytest.mark.parametrize("flag", [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_SINGLETRACK,
                                  ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_EXTZ_ONLY | ksw2.KSW_EZ_REV_CIGAR])
def test_queue_and_grid_agree(flag, monkeypatch):
    """The fixed batch's kind of pairs plus short ragged ones, team 1 forced:
    the queue workers' results equal the oracle's and those of one
    workgroup per pair."""
    print(f"queue vs grid flag {flag:#x}", flush=True)
    monkeypatch.setenv("KSW2_HIP_TEAM", "1")
    monkeypatch.setenv("KSW2_HIP_QUEUE", "1")  # the queue even with one pair per worker
    pairs = W.fixed(seed=7, n=300) + random_pairs(3200 + flag, 500, 1, 700)
    exp = [O.run_oracle(q, t, flag=flag) for q, t in pairs]
    got_q = ksw2.extd2_batch(pairs, flag=flag)
    _check(got_q, exp, "queue")
    monkeypatch.setenv("KSW2_HIP_QUEUE", "0")
    got_g = ksw2.extd2_batch(pairs, flag=flag)
    _check(got_g, exp, "grid")


def test_mixed_groups_in_one_plan(monkeypatch):
    """Four flag groups in one plan: four queue launches, each with its own
    head."""
    monkeypatch.setenv("KSW2_HIP_TEAM", "1")
    monkeypatch.setenv("KSW2_HIP_QUEUE", "1")
    flags = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_SINGLETRACK]
    pairs = random_pairs(3300, 1600, 1, 900)
    rng = np.random.default_rng(5)
    fl = [int(flags[k]) for k in rng.integers(0, 4, len(pairs))]
    exp = [O.run_oracle(q, t, flag=f) for (q, t), f in zip(pairs, fl)]
    got = ksw2.extd2_batch([(q, t, f) for (q, t), f in zip(pairs, fl)])
    _check(got, exp, "mixed")


def test_autotuned_lone_pair_layouts():
    """Single calls on pairs whose bins the autotuner's table covers (short
    swept, long striped sequence: eight rows, a few waves) give the
    reference's results; the table is read by the library itself."""
    rng = np.random.default_rng(41)
    pairs = []
    for qn, tn in ((1000, 9000), (1200, 20000), (2500, 16000), (2000, 30000), (4000, 30000)):
        t = W.random_seq(rng, tn)
        q = W.mutate(rng, t[tn // 3: tn // 3 + qn], 0.06, 0.015)
        pairs.append((q, t))
    for k, (q, t) in enumerate(pairs):
        print(f"tuned lone pair {len(q)} x {len(t)}", flush=True)
        g = ksw2.extd2(q, t)
        r = O.run_oracle(q, t)
        assert same(g, r), f"pair {k}: {diff(g, r)}"
