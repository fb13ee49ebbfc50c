"""The fetch's dense copy of the CIGAR runs (a compaction kernel moves each
delivered pair's runs out of the arena; the host copies only those): forced
on small batches, every flag, band, z-drop, score-only and Singletrack pair
gives the oracle's result, in batch calls and in a resident plan fetched
after each of several runs."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu

FLAGS = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_EXTZ_ONLY,
         ksw2.KSW_EZ_REV_CIGAR, ksw2.KSW_EZ_SINGLETRACK]


def _mixed(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for k, (q, t) in enumerate(random_pairs(seed, n, 1, 2500)):
        flag = FLAGS[k % len(FLAGS)]
        w = -1 if k % 3 else int(rng.integers(10, 700))
        zdrop = -1 if k % 4 else int(rng.integers(30, 200))
        if flag == ksw2.KSW_EZ_SINGLETRACK:
            w, zdrop = -1, -1  # (the Singletrack combinations have tests of their own)
        out.append((q, t, flag, w, zdrop, int(rng.integers(0, 20))))
    return out


def _oracle(p):
    q, t, flag, w, zdrop, eb = p
    return O.run_oracle(q, t, flag=flag, w=w, zdrop=zdrop, end_bonus=eb)


@pytest.mark.parametrize("dense", ["1", None])
def test_batch_fetch_dense_and_whole(dense, monkeypatch):
    if dense:
        monkeypatch.setenv("KSW2_HIP_DENSE_FETCH", dense)
    pairs = _mixed(11, 70)
    got = ksw2.extd2_batch(pairs)
    bad = [f"pair {k} flag {p[2]:#x} w {p[3]} zdrop {p[4]}: {diff(g, _oracle(p))}"
           for k, (p, g) in enumerate(zip(pairs, got)) if not same(g, _oracle(p))]
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(pairs)} differ)"


def test_plan_fetch_dense_after_each_run(monkeypatch):
    monkeypatch.setenv("KSW2_HIP_DENSE_FETCH", "1")
    pairs = random_pairs(12, 40, 200, 3000)
    exp = [O.run_oracle(q, t) for q, t in pairs]
    plan = ksw2.Plan(pairs)
    for rep in range(3):
        plan.run()
        got = plan.fetch()
        for k, (g, e) in enumerate(zip(got, exp)):
            assert same(g, e), f"run {rep} pair {k}: {diff(g, e)}"
    plan.close()
