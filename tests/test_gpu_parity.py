"""HIP ksw_extd2 vs the oracle (bit-exact: every ez field and the CIGAR)."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu

FLAGS = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_REV_CIGAR,
         ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_EXTZ_ONLY,
         ksw2.KSW_EZ_SEMIGLOBAL_END | ksw2.KSW_EZ_RIGHT]


def check_batch(pairs, flag=0, **kw):
    got = ksw2.extd2_batch(pairs, flag=flag, **kw)
    bad = []
    for k, ((q, t), g) in enumerate(zip(pairs, got)):
        ref = O.run_oracle(q, t, flag=flag, **kw)
        if not same(g, ref):
            bad.append(f"pair {k} qlen={len(q)} tlen={len(t)} flag={flag:#x}: {diff(g, ref)}")
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(pairs)} differ)"


@pytest.mark.parametrize("flag", FLAGS)
def test_random_small(flag):
    check_batch(random_pairs(100 + flag, 60, 1, 700), flag=flag)


def test_edge_lengths():
    rng = np.random.default_rng(5)
    pairs = []
    for ql, tl in [(1, 1), (1, 7), (7, 1), (1, 513), (513, 1), (63, 64), (64, 63), (511, 512),
                   (512, 513), (513, 512), (1024, 1), (2, 1025)]:
        pairs.append((W.random_seq(rng, ql), W.random_seq(rng, tl)))
    check_batch(pairs)


def test_identical_and_gap_only():
    rng = np.random.default_rng(6)
    t = W.random_seq(rng, 900)
    pairs = [(t.copy(), t), (t[:300].copy(), t), (t, t[:250].copy()),
             (np.concatenate([t[:400], t[700:]]), t)]
    check_batch(pairs)


def test_single_config_pair():
    check_batch(W.single(seed=1))


def test_fixed_config_sample():
    check_batch(W.fixed(seed=2, n=48))


def _check_against_cpu(pairs, **kw):
    """Large cases: compare with the compiled reference when it is present
    (oracle/_ref, built from the reference sources), else the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    fn = (lambda p: O.run_ref(p[0], p[1], **kw)) if O.ref_available() else (lambda p: O.run_oracle(p[0], p[1], **kw))
    got = ksw2.extd2_batch(pairs, **kw)
    with ThreadPoolExecutor(16) as ex:
        exp = list(ex.map(fn, pairs))
    bad = [f"pair {k} qlen={len(q)} tlen={len(t)}: {diff(g, r)}"
           for k, ((q, t), g, r) in enumerate(zip(pairs, got, exp)) if not same(g, r)]
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(pairs)} differ)"


def test_uneven_sample_multistripe():
    """Uneven inter-anchor intervals: many stripes per pair run concurrently
    on different waves (progress-counter hand-off)."""
    pairs = [p for p in W.uneven(seed=3, n=400) if len(p[0]) * len(p[1]) <= 1.5e8][:120]
    assert max(len(t) for _, t in pairs) > 8 * 512
    _check_against_cpu(pairs)


def test_uneven_sample_flags():
    pairs = [p for p in W.uneven(seed=31, n=200) if len(p[0]) * len(p[1]) <= 6e7][:60]
    for flag in (ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_SCORE_ONLY):
        _check_against_cpu(pairs, flag=flag)


def test_mixed_flags_one_batch():
    """Pairs with different flags in one batch land in different kernel
    groups of one plan; each result must match its own flag."""
    pairs = random_pairs(77, 90, 1, 1500)
    flags = [FLAGS[k % len(FLAGS)] for k in range(len(pairs))]
    got = ksw2.extd2_batch([(q, t, f) for (q, t), f in zip(pairs, flags)])
    for k, ((q, t), f, g) in enumerate(zip(pairs, flags, got)):
        assert same(g, O.run_oracle(q, t, flag=f)), (k, f)


def test_direction_bytes_every_cell():
    """Every direction byte the device writes (not only those on the
    traceback path) equals the reference's byte for that cell."""
    rng = np.random.default_rng(12)
    pairs = random_pairs(91, 6, 300, 1400) + [W.fixed(seed=2, n=1)[0]]
    for flag in (0, ksw2.KSW_EZ_RIGHT):
        plan = ksw2.Plan(pairs, flag=flag)
        plan.run()
        got = plan.fetch()
        for k, (q, t) in enumerate(pairs):
            ref, rd = O.run_oracle(q, t, flag=flag, want_dirs=True)
            gd = plan.dirs(k, len(q), len(t))
            nbad = int((gd != rd).sum())
            assert nbad == 0, f"pair {k} flag {flag}: {nbad} cells differ, first {np.argwhere(gd != rd)[:4].tolist()}"
            assert same(got[k], ref)
        plan.close()


def test_repeat_runs_identical():
    """A plan run repeatedly (stripes of a pair on concurrent waves) gives
    identical, exact results every time."""
    pairs = random_pairs(100, 60, 1, 700) + [p for p in W.uneven(seed=9, n=60) if len(p[0]) * len(p[1]) < 4e7][:20]
    refs = [O.run_oracle(q, t) for q, t in pairs]
    plan = ksw2.Plan(pairs)
    for _ in range(4):
        plan.run()
        got = plan.fetch()
        bad = [k for k in range(len(pairs)) if not same(got[k], refs[k])]
        assert not bad, bad
    plan.close()


@pytest.mark.parametrize("team", [1, 2, 3, 16, -1])
def test_forced_team_sizes(team, monkeypatch):
    """Every team size gives the same exact results: the stripes of a pair
    handed between the waves of one workgroup (team > 1), or run in turn on
    one wave (team 1); -1: sizes 1..16 in turn within one launch (workgroups
    of 16 waves whose extra waves end at once)."""
    monkeypatch.setenv("KSW2_HIP_TEAM", str(team))
    pairs = random_pairs(300 + team, 12, 600, 5000) + \
        [p for p in W.uneven(seed=13, n=80) if len(p[0]) * len(p[1]) < 6e7][:12]
    for flag in (0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY):
        _check_against_cpu(pairs, flag=flag)


def test_direction_bytes_team_of_three(monkeypatch):
    """Direction bytes of every cell with stripes spread over three waves."""
    monkeypatch.setenv("KSW2_HIP_TEAM", "3")
    pairs = random_pairs(92, 3, 1600, 2600)
    plan = ksw2.Plan(pairs)
    plan.run()
    got = plan.fetch()
    for k, (q, t) in enumerate(pairs):
        ref, rd = O.run_oracle(q, t, want_dirs=True)
        gd = plan.dirs(k, len(q), len(t))
        assert int((gd != rd).sum()) == 0, k
        assert same(got[k], ref)
    plan.close()


@pytest.mark.parametrize("flag", [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_SINGLETRACK])
def test_generic_score_path(flag, monkeypatch):
    """The stripe kernel's two score forms give the same bytes: the per-step
    byte table (codes below 4, the default here) and, with it turned off,
    the xor/min/multiply-add form."""
    pairs = random_pairs(300 + flag, 40, 1, 1500)
    ref = ksw2.extd2_batch(pairs, flag=flag)
    monkeypatch.setenv("KSW2_HIP_LUT", "0")
    got = ksw2.extd2_batch(pairs, flag=flag)
    for k, (g, r) in enumerate(zip(got, ref)):
        assert same(g, r), f"pair {k} flag {flag:#x}: {diff(g, r)}"
    if flag != ksw2.KSW_EZ_SINGLETRACK:
        check_batch(pairs, flag=flag)


def test_fast_path_outside_the_byte_table():
    """Fast-path calls the byte table cannot serve take the generic form:
    code 4 (N) as an ordinary symbol under KSW_EZ_GENERIC_SC with a
    diagonal-constant matrix, and a mismatch score below -31."""
    rng = np.random.default_rng(41)
    pairs = []
    for n in (300, 1100, 2600):
        t = W.random_seq(rng, n)
        t[rng.integers(0, n, n // 20)] = 4
        q = W.mutate(rng, t, 0.05, 0.01)
        pairs.append((q, t))
    mat = ksw2.default_matrix(0, -6)  # N vs N scores as a match
    for p in pairs:
        assert ksw2.route(*p, mat=mat, flag=ksw2.KSW_EZ_GENERIC_SC) == 0
    got = ksw2.extd2_batch(pairs, mat=mat, flag=ksw2.KSW_EZ_GENERIC_SC)
    for k, ((q, t), g) in enumerate(zip(pairs, got)):
        r = O.run_oracle(q, t, mat=mat, flag=ksw2.KSW_EZ_GENERIC_SC)
        assert same(g, r), f"N pair {k}: {diff(g, r)}"
    # (a gap open of 20 keeps -40 within the reference's 2(q+e) bound)
    mat40 = ksw2.default_matrix(0, -40)
    plain = [(W.mutate(rng, t, 0.05, 0.01), t) for t in (W.random_seq(rng, n) for n in (700, 2100))]
    for p in plain:
        assert ksw2.route(*p, mat=mat40, q=20) == 0
    got = ksw2.extd2_batch(plain, mat=mat40, q=20)
    for k, ((q, t), g) in enumerate(zip(plain, got)):
        r = O.run_oracle(q, t, mat=mat40, q=20)
        assert same(g, r), f"mismatch -40 pair {k}: {diff(g, r)}"


def test_single_calls_reuse_the_thread_context():
    """Single calls on one thread reuse its stream, staging and device block:
    sizes growing and shrinking, both kernels, a release of the cached block
    in between, and a deadline stop followed by a normal call all give the
    reference's results."""
    rng = np.random.default_rng(44)
    sizes = [300, 2500, 700, 6000, 120, 3000]
    for k, n in enumerate(sizes):
        t = W.random_seq(rng, n)
        q = W.mutate(rng, t, 0.06, 0.015)
        w = 150 if k % 3 == 2 else -1  # every third call on the general kernel
        g = ksw2.extd2(q, t, w=w)
        r = O.run_oracle(q, t, w=w)
        assert same(g, r), f"call {k} n={n} w={w}: {diff(g, r)}"
        if k == 3:
            ksw2.release_cached_memory()
    t = W.random_seq(rng, 20000)
    q = W.mutate(rng, t, 0.06, 0.015)
    with ksw2.progress_callback(lambda r, total: r == 0):
        g = ksw2.extd2(q, t, w=400)
    assert g.stopped == 1
    q2, t2 = q[:900], t[:900]
    g = ksw2.extd2(q2, t2, w=400)
    assert same(g, O.run_oracle(q2, t2, w=400)) and g.stopped == 0
