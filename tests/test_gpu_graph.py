"""hipGraph replay of a resident plan: from the second run on, a plan of
stripe-kernel pairs replays its launches as one captured graph.  Every
replay is checked against the reference's results, with the results
cleared before each replay (KSW2_HIP_GRAPH_CHECK: fill kernels outside the
graph) and the program's own eager work and synchronisations between
replays."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu


def _expect(pairs, flag):
    return [O.run_oracle(q, t, flag=flag) for q, t in pairs]


def _check(got, exp, tag):
    bad = [f"{tag} pair {k}: {diff(g, r)}" for k, (g, r) in enumerate(zip(got, exp)) if not same(g, r)]
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(exp)} differ)"


def _replays(pairs, flag, exp, tag):
    other = random_pairs(901 + flag, 30, 100, 1500)
    other_exp = _expect(other, 0)
    plan = ksw2.Plan(pairs, flag=flag)
    for rep in range(5):
        plan.run()
        _check(plan.fetch(), exp, f"{tag} run {rep}")
        # eager work between replays: another plan and single calls
        op = ksw2.Plan(other)
        op.run()
        _check(op.fetch(), other_exp, f"other after run {rep}")
        op.close()
        q, t = other[rep]
        assert same(ksw2.extd2(q, t), other_exp[rep])
    runs, replays = plan.graph_stats()
    plan.close()
    assert runs == 5 and replays == 4, (runs, replays)


@pytest.mark.parametrize("flag,team", [(0, "-1"), (ksw2.KSW_EZ_RIGHT, "3"), (ksw2.KSW_EZ_SCORE_ONLY, "16"),
                                       (ksw2.KSW_EZ_SINGLETRACK, "1"),
                                       (ksw2.KSW_EZ_EXTZ_ONLY | ksw2.KSW_EZ_REV_CIGAR, "4")])
def test_graph_replays_match_the_reference(flag, team, monkeypatch):
    """Teams forced (one-workgroup teams: a plan that leaves the device idle
    would otherwise spread its teams over CUs, which is never captured)."""
    monkeypatch.setenv("KSW2_HIP_GRAPH_CHECK", "1")
    monkeypatch.setenv("KSW2_HIP_TEAM", team)
    print(f"graph flag {flag:#x} team {team}", flush=True)
    pairs = random_pairs(900 + flag, 120, 1, 2500) + W.fixed(seed=6, n=64)
    _replays(pairs, flag, _expect(pairs, flag), f"flag {flag:#x}")


def test_graph_replays_of_a_device_filling_batch(monkeypatch):
    """A batch that fills the device (eight rows per lane, the planner's
    own teams), as the benchmark runs it."""
    monkeypatch.setenv("KSW2_HIP_GRAPH_CHECK", "1")
    pairs = W.fixed(seed=7, n=1024)
    plan = ksw2.Plan(pairs)
    assert plan.stored_geometry(0)[0] == 8
    plan.close()
    _replays(pairs, 0, _expect(pairs, 0), "fixed 1024")


def test_graph_off_and_ineligible_plans(monkeypatch):
    """KSW2_HIP_GRAPH=0 keeps direct launches; a plan with general-kernel
    pairs (here a narrow band) is never captured.  Results unchanged."""
    pairs = random_pairs(910, 40, 200, 1500)
    exp = _expect(pairs, 0)
    monkeypatch.setenv("KSW2_HIP_GRAPH", "0")
    plan = ksw2.Plan(pairs)
    for _ in range(3):
        plan.run()
        _check(plan.fetch(), exp, "graph off")
    assert plan.graph_stats() == (3, 0)
    plan.close()
    monkeypatch.delenv("KSW2_HIP_GRAPH")
    banded = random_pairs(911, 10, 300, 900)
    bexp = [O.run_oracle(q, t, w=40) for q, t in banded]
    plan = ksw2.Plan(banded, w=40)
    for _ in range(3):
        plan.run()
        _check(plan.fetch(), bexp, "banded")
    assert plan.graph_stats() == (3, 0)
    plan.close()
