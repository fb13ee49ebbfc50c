"""Stripe kernel with four rows per lane (256-row stripes) instead of eight:
the library picks it for plans that would leave most of the device idle
(a lone pair, a few short ones), so that twice the waves share a pair's
sweep; KSW2_HIP_ROWS=4 forces it on every plan.  Tie order, direction bytes,
Singletrack tracks, teams, wide teams and transposed storage must give the
reference's results at either row count."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu

FLAGS = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_REV_CIGAR,
         ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_EXTZ_ONLY, ksw2.KSW_EZ_SEMIGLOBAL_END | ksw2.KSW_EZ_RIGHT,
         ksw2.KSW_EZ_SINGLETRACK]


def _check(pairs, flag=0, **kw):
    got = ksw2.extd2_batch(pairs, flag=flag, **kw)
    bad = []
    for k, ((q, t), g) in enumerate(zip(pairs, got)):
        r = O.run_oracle(q, t, flag=flag, **kw)
        if not same(g, r):
            bad.append(f"pair {k} qlen={len(q)} tlen={len(t)} flag={flag:#x}: {diff(g, r)}")
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(pairs)} differ)"


@pytest.mark.parametrize("flag", FLAGS)
def test_four_rows_every_flag(flag, monkeypatch):
    monkeypatch.setenv("KSW2_HIP_ROWS", "4")
    print(f"rows 4 flag {flag:#x}", flush=True)
    _check(random_pairs(800 + flag, 50, 1, 1500), flag=flag)


def test_four_rows_edge_lengths(monkeypatch):
    """Lengths at the 256-row stripe and 512-row pass boundaries."""
    monkeypatch.setenv("KSW2_HIP_ROWS", "4")
    rng = np.random.default_rng(16)
    pairs = [(W.random_seq(rng, ql), W.random_seq(rng, tl))
             for ql, tl in [(1, 1), (1, 255), (255, 1), (300, 256), (256, 257), (63, 511), (511, 512),
                            (512, 513), (40, 769), (769, 40), (1, 1025), (2000, 1023)]]
    for flag in (0, ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_EXTZ_ONLY):
        _check(pairs, flag=flag)


@pytest.mark.parametrize("transpose", ["0", "1"])
@pytest.mark.parametrize("flag", [0, ksw2.KSW_EZ_RIGHT])
def test_four_rows_direction_bytes_every_cell(flag, transpose, monkeypatch):
    monkeypatch.setenv("KSW2_HIP_ROWS", "4")
    monkeypatch.setenv("KSW2_HIP_TRANSPOSE", transpose)
    pairs = random_pairs(95 + flag, 5, 200, 1800) + [W.fixed(seed=3, n=1)[0]]
    plan = ksw2.Plan(pairs, flag=flag)
    plan.run()
    got = plan.fetch()
    for k, (q, t) in enumerate(pairs):
        assert plan.stored_geometry(k) == (4, transpose == "1")
        ref, rd = O.run_oracle(q, t, flag=flag, want_dirs=True)
        gd = plan.dirs(k, len(q), len(t))
        nbad = int((gd != rd).sum())
        assert nbad == 0, f"pair {k}: {nbad} cells differ, first {np.argwhere(gd != rd)[:4].tolist()}"
        assert same(got[k], ref), diff(got[k], ref)
    plan.close()


@pytest.mark.parametrize("transpose", ["0", "1"])
def test_four_rows_singletrack_tracks_every_cell(transpose, monkeypatch):
    monkeypatch.setenv("KSW2_HIP_ROWS", "4")
    monkeypatch.setenv("KSW2_HIP_TRANSPOSE", transpose)
    pairs = random_pairs(96, 4, 200, 1300)
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


@pytest.mark.parametrize("team,wide", [(3, 0), (16, 0), (0, 2), (0, 6)])
def test_four_rows_teams(team, wide, monkeypatch):
    """Four-row passes handed between the waves of a team, and over wide
    teams of workgroups on different CUs."""
    monkeypatch.setenv("KSW2_HIP_ROWS", "4")
    if team:
        monkeypatch.setenv("KSW2_HIP_TEAM", str(team))
    if wide:
        monkeypatch.setenv("KSW2_HIP_WIDE", str(wide))
    print(f"rows 4 team {team} wide {wide}", flush=True)
    pairs = random_pairs(520 + team + wide, 8, 2000, 7000)
    for flag in (0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY):
        _check(pairs, flag=flag)


def test_rows_chosen_by_occupancy():
    """A lone pair runs on four rows per lane; a batch that fills the device
    at eight rows keeps eight.  Results equal the reference's either way."""
    rng = np.random.default_rng(97)
    t = W.random_seq(rng, 6000)
    q = W.mutate(rng, t, 0.05, 0.01)
    plan = ksw2.Plan([(q, t)])
    assert plan.stored_geometry(0)[0] == 4
    plan.run()
    g = plan.fetch()[0]
    plan.close()
    assert same(g, O.run_oracle(q, t)), diff(g, O.run_oracle(q, t))
    big = W.fixed(seed=4, n=600)
    plan = ksw2.Plan(big)
    assert plan.stored_geometry(0)[0] == 8
    plan.close()
