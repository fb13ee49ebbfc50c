"""Wide stripe teams: a pair's passes spread over several workgroups on
different CUs, handed on through device-scope counters, with the epilogue
in a launch of its own.  Every result (and every direction byte or track
value) must be what the one-workgroup teams and the oracle give."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu

FLAGS = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_REV_CIGAR, ksw2.KSW_EZ_SEMIGLOBAL_END,
         ksw2.KSW_EZ_EXTZ_ONLY, ksw2.KSW_EZ_SINGLETRACK, ksw2.KSW_EZ_SINGLETRACK | ksw2.KSW_EZ_RIGHT]


def _check(pairs, flag=0):
    from concurrent.futures import ThreadPoolExecutor
    fn = (lambda p: O.run_ref(p[0], p[1], flag=flag)) if O.ref_available() and not (
        flag & ksw2.KSW_EZ_SINGLETRACK) else (lambda p: O.run_oracle(p[0], p[1], flag=flag))
    got = ksw2.extd2_batch(pairs, flag=flag)
    with ThreadPoolExecutor(16) as ex:
        exp = list(ex.map(fn, pairs))
    bad = [f"pair {k} qlen={len(q)} tlen={len(t)} flag={flag:#x}: {diff(g, r)}"
           for k, ((q, t), g, r) in enumerate(zip(pairs, got, exp)) if not same(g, r)]
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(pairs)} differ)"


def _long_target_pairs(rng, shapes):
    """A query from a window of a long target (mutated): many passes, few
    columns."""
    out = []
    for tl, ql in shapes:
        t = W.random_seq(rng, tl)
        a = int(rng.integers(0, tl - ql))
        out.append((W.mutate(rng, t[a:a + ql], 0.05, 0.01), t))
    return out


def _pairs(seed):
    # 1..6 passes (1024 target rows each), some uneven-batch intervals and
    # 13-20 pass pairs: teams wider than, equal to and narrower than a
    # pair's pass count
    rng = np.random.default_rng(seed)
    return random_pairs(seed, 14, 300, 5000) + \
        [p for p in W.uneven(seed=seed, n=120) if len(p[0]) * len(p[1]) < 3e7][:8] + \
        _long_target_pairs(rng, ((12500, 700), (16400, 1300), (20000, 400)))


@pytest.mark.parametrize("wgs", [1, 2, 3])
def test_forced_wide_teams_every_flag(wgs, monkeypatch):
    monkeypatch.setenv("KSW2_HIP_WIDE", str(wgs))
    pairs = _pairs(600 + wgs)
    assert max(len(t) for _, t in pairs) > 12 * 1024
    for flag in FLAGS:
        _check(pairs, flag=flag)


def test_wide_direction_bytes_every_cell(monkeypatch):
    monkeypatch.setenv("KSW2_HIP_WIDE", "2")
    pairs = random_pairs(611, 4, 2500, 6000)
    for flag in (0, ksw2.KSW_EZ_RIGHT):
        plan = ksw2.Plan(pairs, flag=flag)
        plan.run()
        got = plan.fetch()
        for k, (q, t) in enumerate(pairs):
            ref, rd = O.run_oracle(q, t, flag=flag, want_dirs=True)
            gd = plan.dirs(k, len(q), len(t))
            nbad = int((gd != rd).sum())
            assert nbad == 0, f"pair {k} flag {flag}: {nbad} cells differ"
            assert same(got[k], ref)
        plan.close()


def test_wide_repeat_runs_identical(monkeypatch):
    """A plan with wide teams run repeatedly (control blocks cleared by a
    fill kernel each run) gives identical, exact results."""
    monkeypatch.setenv("KSW2_HIP_WIDE", "3")
    pairs = random_pairs(612, 10, 2000, 9000)
    refs = [O.run_oracle(q, t) for q, t in pairs]
    plan = ksw2.Plan(pairs)
    for _ in range(3):
        plan.run()
        got = plan.fetch()
        bad = [k for k in range(len(pairs)) if not same(got[k], refs[k])]
        assert not bad, bad
    plan.close()


def test_long_pairs_take_wide_teams_on_their_own():
    """Two intervals of 20-24 passes beside short pairs: the plan gives the
    long ones teams above one workgroup (the batch would otherwise wait on
    them); results exact."""
    rng = np.random.default_rng(613)
    pairs = _long_target_pairs(rng, ((20500, 1500), (24000, 900))) + random_pairs(614, 30, 50, 600)
    _check(pairs)


def test_engine_flights_with_wide_teams(monkeypatch):
    """Batches streamed through the engine (per-pair completion through the
    host mirror, written by the wide epilogue)."""
    monkeypatch.setenv("KSW2_HIP_WIDE", "2")
    monkeypatch.setenv("KSW2_HIP_MEM_BUDGET", str(48 << 20))
    pairs = random_pairs(615, 40, 500, 4000) + W.fixed(seed=2, n=8)
    got = ksw2.extd2_batch(pairs)
    for k, ((q, t), g) in enumerate(zip(pairs, got)):
        exp = O.run_oracle(q, t)
        assert same(g, exp), f"pair {k}: {diff(g, exp)}"


def test_spread_pairs_exact_wide_or_one_workgroup(monkeypatch):
    """A lone call with 5-16 passes runs its team spread over two or more
    workgroups when waiting-launch slots are free, else as one workgroup:
    both give the oracle's results."""
    pairs = W.single(seed=1) + random_pairs(616, 3, 4200, 9000)
    exp = [O.run_oracle(q, t) for q, t in pairs]
    for k, (q, t) in enumerate(pairs):
        g = ksw2.extd2(q, t)
        assert same(g, exp[k]), f"spread pair {k}: {diff(g, exp[k])}"
    monkeypatch.setenv("KSW2_HIP_SPREAD_FALLBACK", "1")
    got = ksw2.extd2_batch(pairs)  # a small batch: spread pairs, one-workgroup launch
    for k, (g, e) in enumerate(zip(got, exp)):
        assert same(g, e), f"fallback pair {k}: {diff(g, e)}"


def test_concurrent_single_calls_from_threads():
    """Calls from several host threads at once (AnchorWave's per-interval
    pattern): some take the waiting-launch slots and run spread, the others
    run as one workgroup; every result exact."""
    from concurrent.futures import ThreadPoolExecutor
    pairs = random_pairs(617, 24, 4200, 7000)
    exp = [O.run_oracle(q, t) for q, t in pairs]
    with ThreadPoolExecutor(8) as ex:
        got = list(ex.map(lambda p: ksw2.extd2(p[0], p[1]), pairs))
    for k, (g, e) in enumerate(zip(got, exp)):
        assert same(g, e), f"pair {k}: {diff(g, e)}"


def test_wide_give_up_recomputes_the_pair(monkeypatch):
    """A wide pair whose team gave up a wait is recomputed by its epilogue
    workgroup as a one-workgroup team (forced here from the host for every
    wide pair, after its team ran normally): results stay exact and the
    batch succeeds, for every flag."""
    monkeypatch.setenv("KSW2_HIP_WIDE", "2")
    monkeypatch.setenv("KSW2_HIP_WIDE_REDO", "1")
    pairs = _pairs(611)
    for flag in FLAGS:
        _check(pairs, flag=flag)
