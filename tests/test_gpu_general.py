"""The general (anti-diagonal) kernel: every call outside the stripe
kernel's exactness domain -- bands, z-drop, APPROX_MAX / APPROX_DROP, generic
matrices, the wildcard code, e <= e2, positive match scores, wrapping int8
intermediates -- checked bit-exact (every ez field and the CIGAR) against
the vectors the compiled reference wrote (tests/golden) and against the
oracle on seeded random inputs."""
import zlib
from collections import defaultdict

import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import as_result, diff, golden_cases, random_pairs, same

pytestmark = pytest.mark.gpu

FLAGS = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_REV_CIGAR,
         ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_EXTZ_ONLY,
         ksw2.KSW_EZ_SEMIGLOBAL_END | ksw2.KSW_EZ_RIGHT]


def test_golden_vectors_one_call_each():
    """Each reference-written vector through ksw_extd2_hip (one call)."""
    bad, n = [], 0
    for c, q, t, kw in golden_cases():
        got = ksw2.extd2(q, t, **kw)
        n += 1
        if not same(got, as_result(c["expect"])):
            bad.append(f"{c['name']} qlen={len(q)} tlen={len(t)} {kw['flag']:#x} w={kw['w']} "
                       f"zdrop={kw['zdrop']}: {diff(got, as_result(c['expect']))}")
    assert n >= 380
    assert not bad, "\n".join(bad[:6]) + f"\n({len(bad)} of {n} differ)"


def test_golden_vectors_batched():
    """The same vectors, grouped by scoring into batches with per-pair band,
    z-drop, end bonus and flags (both kernels inside one plan)."""
    groups = defaultdict(list)
    for c, q, t, kw in golden_cases():
        key = (kw["m"], tuple(kw["mat"].tolist()), kw["q"], kw["e"], kw["q2"], kw["e2"])
        groups[key].append((c, q, t, kw))
    bad, n = [], 0
    for (m, mat, q_, e, q2, e2), items in groups.items():
        pairs = [(q, t, kw["flag"], kw["w"], kw["zdrop"], kw["end_bonus"]) for _, q, t, kw in items]
        got = ksw2.extd2_batch(pairs, mat=np.array(mat, np.int8), m=m, q=q_, e=e, q2=q2, e2=e2)
        for (c, q, t, kw), g in zip(items, got):
            n += 1
            if not same(g, as_result(c["expect"])):
                bad.append(f"{c['name']}: {diff(g, as_result(c['expect']))}")
    assert not bad, "\n".join(bad[:6]) + f"\n({len(bad)} of {n} differ)"


SCENARIOS = {
    # name: (scoring kwargs, per-call kwargs)
    "band_narrow": (dict(), dict(w=17)),
    "band_medium": (dict(), dict(w=150)),
    "zdrop": (dict(), dict(zdrop=40)),
    "zdrop_band": (dict(), dict(zdrop=100, w=60)),
    "e_equal_e2": (dict(q=6, e=2, q2=30, e2=2), dict()),
    "e_below_e2": (dict(q=4, e=1, q2=24, e2=3), dict()),
    "positive_match": (dict(mat=O.default_matrix(2, -4)), dict()),
    "wrapping": (dict(q=40, e=12, q2=90, e2=5, mat=O.default_matrix(0, -30)), dict()),
    "big_match": (dict(mat=O.default_matrix(9, -9), q=20, e=3, q2=60, e2=1), dict(zdrop=200)),
}


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_scenarios_vs_oracle(name):
    skw, ckw = SCENARIOS[name]
    pairs = random_pairs(zlib.crc32(name.encode()) % 1000, 24, 1, 900)
    for flag in FLAGS + [ksw2.KSW_EZ_APPROX_MAX, ksw2.KSW_EZ_APPROX_MAX | ksw2.KSW_EZ_APPROX_DROP]:
        kw = dict(skw, **ckw, flag=flag)
        got = ksw2.extd2_batch(pairs, **kw)
        bad = [f"pair {k} ({len(q)}x{len(t)}): {diff(g, O.run_oracle(q, t, **kw))}"
               for k, ((q, t), g) in enumerate(zip(pairs, got)) if not same(g, O.run_oracle(q, t, **kw))]
        assert not bad, f"{name} flag {flag:#x}\n" + "\n".join(bad[:5])


def test_wildcards_and_generic_matrix():
    """Codes m-1 (N) in match/mismatch mode (scored sc_N), and a non-uniform
    KSW_EZ_GENERIC_SC matrix."""
    rng = np.random.default_rng(44)
    pairs = []
    for q, t in random_pairs(45, 30, 1, 800):
        t = t.copy()
        t[rng.integers(0, len(t), max(1, len(t) // 30))] = 4
        pairs.append((q, t))
    gmat = rng.integers(-8, 3, 25).astype(np.int8)
    for mat, flag in ((O.default_matrix(), 0), (O.default_matrix(), ksw2.KSW_EZ_RIGHT),
                      (gmat, ksw2.KSW_EZ_GENERIC_SC), (gmat, ksw2.KSW_EZ_GENERIC_SC | ksw2.KSW_EZ_SCORE_ONLY)):
        got = ksw2.extd2_batch(pairs, mat=mat, flag=flag)
        for k, ((q, t), g) in enumerate(zip(pairs, got)):
            r = O.run_oracle(q, t, mat=mat, flag=flag)
            assert same(g, r), f"pair {k} flag {flag:#x}: {diff(g, r)}"


@pytest.mark.parametrize("threads", [64, 256, 1024])
def test_forced_general_on_fast_inputs(threads, monkeypatch):
    """Inputs the stripe kernel would take, forced through the general
    kernel at several workgroup sizes (ownership t mod threads)."""
    monkeypatch.setenv("KSW2_HIP_FORCE_GENERAL", "1")
    monkeypatch.setenv("KSW2_HIP_GEN_THREADS", str(threads))
    pairs = random_pairs(500 + threads, 16, 1, 1500) + W.fixed(seed=2, n=2)
    for flag in FLAGS:
        got = ksw2.extd2_batch(pairs, flag=flag)
        for k, ((q, t), g) in enumerate(zip(pairs, got)):
            r = O.run_oracle(q, t, flag=flag)
            assert same(g, r), f"pair {k} flag {flag:#x} threads {threads}: {diff(g, r)}"


def test_band_exit_long_pair():
    """A pair whose length difference exceeds the band: the reference leaves
    the band (zdropped) and backtracks from its maximum."""
    rng = np.random.default_rng(8)
    t = W.random_seq(rng, 3000)
    q = np.concatenate([t[:1200], t[2000:]])
    for flag in (0, ksw2.KSW_EZ_EXTZ_ONLY, ksw2.KSW_EZ_SEMIGLOBAL_END):
        for w in (100, 500, 999, 1001):
            r = O.run_oracle(q, t, w=w, flag=flag)
            g = ksw2.extd2(q, t, w=w, flag=flag)
            assert same(g, r), f"w={w} flag={flag:#x}: {diff(g, r)}"
            if w < 511:
                # a lone narrow pair takes the band kernel; the general
                # kernel's band exit too
                with pytest.MonkeyPatch.context() as mp:
                    mp.setenv("KSW2_HIP_BAND_MIN", "512")
                    g = ksw2.extd2(q, t, w=w, flag=flag)
                assert same(g, r), f"general kernel w={w} flag={flag:#x}: {diff(g, r)}"


@pytest.mark.parametrize("name", ["band_medium", "zdrop_band", "wrapping", "positive_match"])
def test_scenarios_state_in_hbm(name, monkeypatch):
    """The general kernel with its per-position state in HBM (the form for
    bands wider than its LDS rings) on the same scenarios (narrow bands kept
    on the general kernel)."""
    monkeypatch.setenv("KSW2_HIP_GEN_RING", "0")
    monkeypatch.setenv("KSW2_HIP_BAND_MIN", "512")
    test_scenarios_vs_oracle(name)


@pytest.mark.parametrize("name", ["band_narrow", "band_medium"])
def test_narrow_bands_on_either_kernel(name, monkeypatch):
    """Narrow bands in the band kernel's domain: a small batch takes the
    band kernel (the plan's prediction), a fixed width rule keeps them on
    the general kernel; both equal the oracle."""
    test_scenarios_vs_oracle(name)
    monkeypatch.setenv("KSW2_HIP_BAND_MIN", "512")
    test_scenarios_vs_oracle(name)


@pytest.mark.parametrize("ring", ["1", "0"])
def test_long_banded_pairs(ring, monkeypatch):
    """Long pairs under a band: the band crosses the LDS rings many times
    over (positions far beyond the ring size), exact with the state in LDS
    rings and in HBM.  Includes a w = 4096 pair (the banded long-pair
    workload's band) and band exits."""
    monkeypatch.setenv("KSW2_HIP_GEN_RING", ring)
    monkeypatch.setenv("KSW2_HIP_BAND_MIN", "512")  # narrow bands on the general kernel
    rng = np.random.default_rng(61)
    t1 = W.random_seq(rng, 20000)
    q1 = W.mutate(rng, t1, 0.08, 0.02)
    t2 = W.random_seq(rng, 40000)
    q2 = W.big_indel(rng, W.mutate(rng, t2, 0.1, 0.02), 1000, 3000)
    t3 = W.random_seq(rng, 12000)
    q3 = np.concatenate([t3[:5000], t3[9000:]])  # length difference beyond a w = 300 band
    cases = [(q1, t1, dict(w=300)), (q1, t1, dict(w=300, zdrop=200)), (q2, t2, dict(w=4096)),
             (q3, t3, dict(w=300)), (q1, t1, dict(w=300, mat=O.default_matrix(2, -4)))]
    for k, (q, t, kw) in enumerate(cases):
        g = ksw2.extd2(q, t, **kw)
        r = O.run_oracle(q, t, **kw)
        assert same(g, r), f"case {k} ring={ring}: {diff(g, r)}"


@pytest.mark.parametrize("name", ["band_narrow", "zdrop", "zdrop_band"])
def test_wave_walk_on_every_pair(name, monkeypatch):
    """The one-wave walk with staged rows (long pairs' backtrack) forced on
    every general pair, short ones included: the reference vectors and the
    oracle scenarios over every flag (left/right alignment, reversed CIGAR,
    EXTZ_ONLY, SEMIGLOBAL_END, z-drop stops)."""
    monkeypatch.setenv("KSW2_HIP_TILED_MIN", "0")
    test_scenarios_vs_oracle(name)
    if name == "band_narrow":
        test_golden_vectors_batched()
