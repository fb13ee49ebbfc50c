"""Pairs stored transposed: the query in the stripes and the target swept,
which the library picks when the query gives more 1024-row passes than the
target (a long query no longer sets a long sweep).  Tie order, direction
bytes, Singletrack tracks, mqe/mte and every result must be the reference's
in either orientation; KSW2_HIP_TRANSPOSE=1 forces it on every pair."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu

FLAGS = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_REV_CIGAR,
         ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_EXTZ_ONLY, ksw2.KSW_EZ_SEMIGLOBAL_END | ksw2.KSW_EZ_RIGHT,
         ksw2.KSW_EZ_SINGLETRACK]


def _check(pairs, flag=0, ref=False, **kw):
    got = ksw2.extd2_batch(pairs, flag=flag, **kw)
    run = O.run_ref if ref and O.ref_available() else O.run_oracle
    bad = []
    for k, ((q, t), g) in enumerate(zip(pairs, got)):
        r = run(q, t, flag=flag, **kw)
        if not same(g, r):
            bad.append(f"pair {k} qlen={len(q)} tlen={len(t)} flag={flag:#x}: {diff(g, r)}")
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(pairs)} differ)"


@pytest.mark.parametrize("flag", FLAGS)
def test_forced_transposition_every_flag(flag, monkeypatch):
    monkeypatch.setenv("KSW2_HIP_TRANSPOSE", "1")
    _check(random_pairs(700 + flag, 50, 1, 1500), flag=flag)


def test_forced_transposition_edge_lengths(monkeypatch):
    monkeypatch.setenv("KSW2_HIP_TRANSPOSE", "1")
    rng = np.random.default_rng(15)
    pairs = [(W.random_seq(rng, ql), W.random_seq(rng, tl))
             for ql, tl in [(1, 1), (1, 7), (7, 1), (1, 513), (513, 1), (63, 64), (64, 63), (511, 512),
                            (512, 513), (1025, 2), (2, 1024), (3000, 40)]]
    for flag in (0, ksw2.KSW_EZ_SEMIGLOBAL_END):
        _check(pairs, flag=flag)


@pytest.mark.parametrize("flag", [0, ksw2.KSW_EZ_RIGHT])
def test_transposed_direction_bytes_every_cell(flag, monkeypatch):
    monkeypatch.setenv("KSW2_HIP_TRANSPOSE", "1")
    pairs = random_pairs(93 + flag, 5, 300, 1800) + [W.fixed(seed=2, n=1)[0]]
    plan = ksw2.Plan(pairs, flag=flag)
    plan.run()
    got = plan.fetch()
    for k, (q, t) in enumerate(pairs):
        assert plan.stored_transposed(k)
        ref, rd = O.run_oracle(q, t, flag=flag, want_dirs=True)
        gd = plan.dirs(k, len(q), len(t))
        nbad = int((gd != rd).sum())
        assert nbad == 0, f"pair {k}: {nbad} cells differ, first {np.argwhere(gd != rd)[:4].tolist()}"
        assert same(got[k], ref), diff(got[k], ref)
    plan.close()


def test_transposed_singletrack_tracks_every_cell(monkeypatch):
    monkeypatch.setenv("KSW2_HIP_TRANSPOSE", "1")
    pairs = random_pairs(94, 4, 200, 1300)
    plan = ksw2.Plan(pairs, flag=ksw2.KSW_EZ_SINGLETRACK)
    plan.run()
    got = plan.fetch()
    for k, (q, t) in enumerate(pairs):
        ref, rt = O.run_oracle(q, t, flag=ksw2.KSW_EZ_SINGLETRACK, want_dirs=True)
        gt = plan.tracks(k, len(q), len(t))
        nbad = int((gt != rt).sum())
        assert nbad == 0, f"pair {k}: {nbad} track values differ"
        assert same(got[k], ref), diff(got[k], ref)
    plan.close()


@pytest.mark.parametrize("team,wide", [(3, 0), (16, 0), (0, 2)])
def test_transposed_teams(team, wide, monkeypatch):
    """Transposed pairs handed between the waves of a team, and over a wide
    team of workgroups on different CUs."""
    monkeypatch.setenv("KSW2_HIP_TRANSPOSE", "1")
    if team:
        monkeypatch.setenv("KSW2_HIP_TEAM", str(team))
    if wide:
        monkeypatch.setenv("KSW2_HIP_WIDE", str(wide))
    pairs = random_pairs(500 + team + wide, 8, 2000, 6000)
    for flag in (0, ksw2.KSW_EZ_RIGHT):
        _check(pairs, flag=flag)


def test_long_query_pairs_are_transposed():
    """Pairs whose query has more 1024-row passes than the target are stored
    transposed by default, and results equal the reference's: uneven-batch
    pairs with a long inserted query segment."""
    un = W.uneven(seed=17, n=3000)
    pairs = [p for p in un if len(p[0]) > 2 * len(p[1]) + 2048 and len(p[0]) * len(p[1]) < 1.2e8][:10]
    assert len(pairs) >= 5
    plan = ksw2.Plan(pairs)
    for k in range(len(pairs)):
        assert plan.stored_transposed(k)
    plan.close()
    even = [p for p in un if (len(p[1]) + 1023) // 1024 >= (len(p[0]) + 1023) // 1024][:3]
    plan = ksw2.Plan(even)
    for k in range(len(even)):
        assert not plan.stored_transposed(k)
    plan.close()
    _check(pairs, ref=True)
    _check(pairs, flag=ksw2.KSW_EZ_SEMIGLOBAL_END, ref=True)
