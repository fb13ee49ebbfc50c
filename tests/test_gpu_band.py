"""The band kernel (ksw2_band.hip): banded calls inside the stripe kernel's
exactness domain, bit-exact against the oracle (every ez field and the
CIGAR), including band edges, band exits and its fallback to the general
kernel."""
from collections import defaultdict

import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import as_result, diff, golden_cases, random_pairs, same

pytestmark = pytest.mark.gpu

FLAGS = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_REV_CIGAR, ksw2.KSW_EZ_SEMIGLOBAL_END,
         ksw2.KSW_EZ_EXTZ_ONLY]


def check(tuples, **kw):
    got = ksw2.extd2_batch(tuples, **kw)
    bad = []
    for k, ((q, t, f, w), g) in enumerate(zip(tuples, got)):
        r = O.run_oracle(q, t, flag=f, w=w, **kw)
        if not same(g, r):
            bad.append(f"pair {k} qlen={len(q)} tlen={len(t)} w={w} flag={f:#x}: {diff(g, r)}")
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(tuples)} differ)"
    return got


def test_band_kernel_takes_wide_bands():
    """Pairs 1-6 kbp under bands 511-2000 wide, all flags: the band kernel's
    default routing (band at least 512 wide)."""
    rng = np.random.default_rng(71)
    pairs = random_pairs(72, 36, 1000, 6000)
    tuples = [(q, t, FLAGS[k % len(FLAGS)], int(rng.integers(511, 2000))) for k, (q, t) in enumerate(pairs)]
    check(tuples)


@pytest.mark.parametrize("wlo,whi", [(0, 20), (20, 130), (130, 600)])
def test_band_kernel_narrow_bands_forced(wlo, whi, monkeypatch):
    """Every banded pair in the domain forced onto the band kernel, with
    narrow bands: the 16/64 roundings, the injection at st16, the AVX-512
    lane wrap and stale positions at both edges on every anti-diagonal."""
    monkeypatch.setenv("KSW2_HIP_BAND_MIN", "0")
    rng = np.random.default_rng(73 + wlo)
    pairs = random_pairs(74 + wlo, 40, 1, 1500)
    tuples = [(q, t, FLAGS[k % len(FLAGS)], int(rng.integers(wlo, whi))) for k, (q, t) in enumerate(pairs)]
    check(tuples)


def test_band_kernel_golden_vectors_forced(monkeypatch):
    """The reference's own vectors, batched by scoring, with every banded
    in-domain pair on the band kernel."""
    monkeypatch.setenv("KSW2_HIP_BAND_MIN", "0")
    groups = defaultdict(list)
    for c, q, t, kw in golden_cases():
        key = (kw["m"], tuple(kw["mat"].tolist()), kw["q"], kw["e"], kw["q2"], kw["e2"])
        groups[key].append((c, q, t, kw))
    bad, n = [], 0
    for (m, mat, q_, e, q2, e2), items in groups.items():
        tuples = [(q, t, kw["flag"], kw["w"], kw["zdrop"], kw["end_bonus"]) for _, q, t, kw in items]
        got = ksw2.extd2_batch(tuples, mat=np.array(mat, np.int8), m=m, q=q_, e=e, q2=q2, e2=e2)
        for (c, q, t, kw), g in zip(items, got):
            n += 1
            if not same(g, as_result(c["expect"])):
                bad.append(f"{c['name']}: {diff(g, as_result(c['expect']))}")
    assert not bad, "\n".join(bad[:6]) + f"\n({len(bad)} of {n} differ)"


def test_band_exits():
    """Length differences beyond the band: the reference leaves the band
    (zdropped) and, with match score 0, reports no maximum and no path."""
    rng = np.random.default_rng(75)
    tuples = []
    for k in range(10):
        t = W.random_seq(rng, int(rng.integers(3000, 9000)))
        cut = int(rng.integers(800, 2500))
        q = np.concatenate([t[:len(t) // 3], t[len(t) // 3 + cut:]])
        tuples.append((q, t, FLAGS[k % len(FLAGS)], int(rng.integers(512, cut))))
    got = check(tuples)
    assert sum(g.zdropped for g in got) >= 5


def test_band_kernel_long_pair_w4096():
    """A 40 kbp pair with a 1-3 kbp insertion under w = 4096 (the banded
    long-pair workload's band), both gap alignments."""
    rng = np.random.default_rng(61)
    W.mutate(rng, W.random_seq(rng, 20000), 0.08, 0.02)  # same stream as test_long_banded_pairs
    t2 = W.random_seq(rng, 40000)
    q2 = W.big_indel(rng, W.mutate(rng, t2, 0.1, 0.02), 1000, 3000)
    check([(q2, t2, 0, 4096), (q2, t2, ksw2.KSW_EZ_RIGHT, 4096)])


@pytest.mark.parametrize("team", ["1", "3"])
def test_band_kernel_forced_team_sizes(team, monkeypatch):
    """Stripes on one wave in turn (team 1) and spread over three waves."""
    monkeypatch.setenv("KSW2_HIP_BAND_MIN", "0")
    monkeypatch.setenv("KSW2_HIP_BAND_TEAM", team)
    rng = np.random.default_rng(77)
    pairs = random_pairs(78, 12, 1500, 5000)
    check([(q, t, FLAGS[k % len(FLAGS)], int(rng.integers(100, 1500))) for k, (q, t) in enumerate(pairs)])


def test_band_redo_falls_back_to_the_general_kernel(monkeypatch):
    """Pairs flagged by the band kernel are recomputed by the general kernel
    on the same stream (forced here for every pair): results stay exact."""
    monkeypatch.setenv("KSW2_HIP_BAND_REDO", "1")
    rng = np.random.default_rng(79)
    pairs = random_pairs(80, 16, 800, 3000)
    check([(q, t, FLAGS[k % len(FLAGS)], int(rng.integers(512, 1500))) for k, (q, t) in enumerate(pairs)])
