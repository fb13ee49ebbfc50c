"""Pins the oracle (oracle/ksw2_oracle.c), CPU only.

1. Golden vectors written by the reference ksw_extd2_avx512 itself
   (tests/golden/make_golden.py), including the reference's FASTA test data.
2. Live differential against the reference compiled from its sources
   (oracle/_ref), all three ISAs, when that build is present.
3. Reference-free mathematics: brute-force enumeration of every alignment on
   tiny inputs, an independent textbook Gotoh DP, CIGAR rescoring, closed
   forms (identical sequences, prefix = one gap), CIGAR length conservation.
4. The fast kernel's eligibility predicate (ksw2_hip_route) against the
   instrumented restatement: no in-band int8 wrap, no binding match cap.
5. Planted mistakes in the oracle's inputs/ISA are caught.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import as_result, diff, golden_cases, same



def test_golden_oracle_avx512():
    bad = []
    n = 0
    for c, q, t, kw in golden_cases():
        got = O.run_oracle(q, t, isa=O.ISA_AVX512, **kw)
        exp = as_result(c["expect"])
        n += 1
        if not same(got, exp):
            bad.append(f"{c['name']}: {diff(got, exp)}")
    assert n >= 380
    assert not bad, "\n".join(bad[:5])


def test_golden_oracle_sse_semantics():
    """Where the reference's SSE and AVX-512 kernels agree, the SSE emulation
    must reproduce the vector too; where they differ, it must not (those are
    the banded cases whose padding differs between ISAs)."""
    for c, q, t, kw in golden_cases():
        got = O.run_oracle(q, t, isa=O.ISA_SSE, **kw)
        if c["sse_same"]:
            assert same(got, as_result(c["expect"])), c["name"]


def test_golden_mt_pair_cigar_consistent():
    """MT-orang vs MT-human (reference FASTA data): the golden CIGAR rescoring
    equals the golden score and consumes both sequences."""
    for c, q, t, kw in golden_cases():
        if not c["name"].startswith("MT-orang") or (kw["flag"] & ~ksw2.KSW_EZ_GENERIC_SC) not in (0, 2) \
                or kw["w"] >= 0:
            continue
        exp = c["expect"]
        assert O.cigar_lengths(exp["cigar"]) == (len(q), len(t))
        if True:
            assert O.cigar_score(exp["cigar"], q, t, **{k: kw[k] for k in ("mat", "m", "q", "e", "q2", "e2")}) \
                == exp["score"]


@pytest.mark.skipif(not O.ref_available(), reason="reference checker not built (oracle/_ref)")
def test_oracle_vs_reference_random():
    rng = np.random.default_rng(99)
    for it in range(250):
        t = W.random_seq(rng, int(rng.integers(1, 600)))
        q = W.mutate(rng, t, 0.1, 0.03) if rng.random() < 0.7 else W.random_seq(rng, int(rng.integers(1, 600)))
        q = q if len(q) else W.random_seq(rng, 1)
        if rng.random() < 0.3:
            q = W.big_indel(rng, q, 10, max(11, len(q) // 2))
        pen = (8, 2, 75, 1) if rng.random() < 0.5 else (int(rng.integers(0, 25)), int(rng.integers(0, 6)),
                                                       int(rng.integers(0, 100)), int(rng.integers(0, 4)))
        kw = dict(mat=O.default_matrix(int(rng.integers(0, 3)), -int(rng.integers(1, 9))),
                  q=pen[0], e=pen[1], q2=pen[2], e2=pen[3],
                  w=-1 if rng.random() < 0.5 else int(rng.integers(0, 200)),
                  zdrop=-1 if rng.random() < 0.6 else int(rng.integers(0, 200)),
                  end_bonus=int(rng.integers(0, 20)),
                  flag=int(rng.choice([0, 1, 2, 4, 8, 0x18, 0x40, 0x80, 0x800, 0x802])))
        for isa, oi in (("avx512", O.ISA_AVX512), ("sse", O.ISA_SSE)):
            r = O.run_ref(q, t, isa=isa, **kw)
            o = O.run_oracle(q, t, isa=oi, **kw)
            assert same(o, r), f"it={it} {isa} {kw['w']} {kw['flag']}: {diff(o, r)}"


# ---------------------------------------------------------------- mathematics
def gap_cost(length, q, e, q2, e2):
    return min(q + length * e, q2 + length * e2)


def brute_force(qs, ts, mat, q, e, q2, e2):
    """Best global score over every alignment (ops M, I, D)."""
    best = None
    n, m = len(ts), len(qs)

    def rec(i, j, ops):
        nonlocal best
        if i == n and j == m:
            s, k = 0, 0
            ii = jj = 0
            while k < len(ops):
                op = ops[k]
                if op == "M":
                    s += int(mat[ts[ii] * 5 + qs[jj]])
                    ii, jj, k = ii + 1, jj + 1, k + 1
                else:
                    L = 0
                    while k < len(ops) and ops[k] == op:
                        L, k = L + 1, k + 1
                    s -= gap_cost(L, q, e, q2, e2)
                    if op == "D":
                        ii += L
                    else:
                        jj += L
            best = s if best is None else max(best, s)
            return
        if i < n and j < m:
            rec(i + 1, j + 1, ops + "M")
        if i < n:
            rec(i + 1, j, ops + "D")
        if j < m:
            rec(i, j + 1, ops + "I")

    rec(0, 0, "")
    return best


PARAMS_EXACT = [(8, 2, 75, 1, 0, -6), (4, 3, 20, 1, 0, -4), (2, 4, 24, 1, 1, -3), (6, 2, 30, 0, 0, -9)]


@pytest.mark.parametrize("pen", PARAMS_EXACT)
def test_bruteforce_tiny(pen):
    q_, e, q2, e2, ma, mm = pen
    mat = O.default_matrix(ma, mm)
    rng = np.random.default_rng(sum(pen) + 1000)
    for _ in range(60):
        n, m = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        ts, qs = W.random_seq(rng, n), W.random_seq(rng, m)
        if rng.random() < 0.5:
            qs = ts[: max(1, m)].copy() if rng.random() < 0.5 else qs
        r = O.run_oracle(qs, ts, mat=mat, q=q_, e=e, q2=q2, e2=e2, flag=ksw2.KSW_EZ_GENERIC_SC)
        assert r.score == brute_force(qs.tolist(), ts.tolist(), mat, q_, e, q2, e2), (qs, ts)
        # the path the traceback emits scores exactly the optimum
        assert O.cigar_score(r.cigar, qs, ts, mat=mat, q=q_, e=e, q2=q2, e2=e2) == r.score


@pytest.mark.parametrize("pen", PARAMS_EXACT)
def test_gotoh_medium(pen):
    q_, e, q2, e2, ma, mm = pen
    mat = O.default_matrix(ma, mm)
    rng = np.random.default_rng(7 + sum(pen))
    for _ in range(12):
        t = W.random_seq(rng, int(rng.integers(50, 700)))
        qq = W.mutate(rng, t, 0.1, 0.04)
        if rng.random() < 0.5:
            qq = W.big_indel(rng, qq, 20, 200)
        qq = qq if len(qq) else W.random_seq(rng, 3)
        for flag in (0, ksw2.KSW_EZ_RIGHT):
            r = O.run_oracle(qq, t, mat=mat, q=q_, e=e, q2=q2, e2=e2, flag=flag)
            g = O.gotoh2_global(qq, t, mat=mat, q=q_, e=e, q2=q2, e2=e2)
            assert r.score == g
            assert O.cigar_lengths(r.cigar) == (len(qq), len(t))
            assert O.cigar_score(r.cigar, qq, t, mat=mat, q=q_, e=e, q2=q2, e2=e2) == r.score


def test_closed_forms():
    rng = np.random.default_rng(3)
    t = W.random_seq(rng, 800)
    r = O.run_oracle(t.copy(), t)
    assert r.score == 0 and r.cigar == ((800 << 4) | 0,)
    for k in (1, 5, 66, 67, 68, 300, 799):
        r = O.run_oracle(t[:k].copy(), t)  # query = prefix: one deletion gap
        assert r.score == -gap_cost(800 - k, 8, 2, 75, 1), k
        r = O.run_oracle(t, t[:k].copy())  # target = prefix: one insertion gap
        assert r.score == -gap_cost(800 - k, 8, 2, 75, 1), k


# ------------------------------------------------------- fast-path predicate
def test_route_predicate_is_sound():
    """For every scoring the router sends to the fast kernel, the reference's
    int8 lanes never wrap in-band and z never exceeds mat[0] before the cap
    (so int32 lanes without the cap compute identical values)."""
    rng = np.random.default_rng(2024)
    checked = 0
    for _ in range(400):
        pen = (int(rng.integers(0, 40)), int(rng.integers(0, 12)), int(rng.integers(0, 128)), int(rng.integers(0, 8)))
        mat = O.default_matrix(0, -int(rng.integers(0, 60)))
        t = W.random_seq(rng, int(rng.integers(20, 300)))
        qq = W.mutate(rng, t, 0.15, 0.05) if rng.random() < 0.6 else W.random_seq(rng, int(rng.integers(20, 300)))
        if len(qq) == 0 or rng.random() < 0.3:
            qq = np.concatenate([qq, W.random_seq(rng, int(rng.integers(1, 150)))])
        if ksw2.route(qq, t, mat=mat, q=pen[0], e=pen[1], q2=pen[2], e2=pen[3]) != 0:
            continue
        with O.RangeProbe() as p:
            O.run_oracle(qq, t, mat=mat, q=pen[0], e=pen[1], q2=pen[2], e2=pen[3])
        assert p.wraps == 0 and p.cap_binds == 0, (pen, p.wraps, p.cap_binds)
        checked += 1
    assert checked > 60


def test_route_rejects_exactly_when_needed():
    t = np.zeros(100, np.uint8)
    qq = np.ones(90, np.uint8)
    assert ksw2.route(qq, t) == 0
    assert ksw2.route(qq, t, w=50) & 1
    assert ksw2.route(qq, t, w=100) == 0          # band as wide as the matrix
    assert ksw2.route(qq, t, zdrop=100) & 2
    assert ksw2.route(qq, t, flag=ksw2.KSW_EZ_APPROX_MAX) & 4
    assert ksw2.route(qq, t, mat=O.default_matrix(2, -4)) & 8   # positive match score
    assert ksw2.route(qq, t, q=8, e=1, q2=75, e2=1) & 8         # e == e2
    tn = t.copy()
    tn[3] = 4
    assert ksw2.route(qq, tn) & 16                              # wildcard in match/mismatch mode
    assert ksw2.route(qq, tn, flag=ksw2.KSW_EZ_GENERIC_SC) == 0  # generic matrix is diagonal-constant
    assert ksw2.route(qq, t, mat=O.default_matrix(0, -30)) == -1  # reference returns early
    assert ksw2.route(qq[:0], t) == -1


# ------------------------------------------------------------ planted mistakes
def test_planted_mistakes_are_caught():
    """The golden comparison is sensitive: a one-off change of a gap
    parameter, or the SSE padding on cases where the ISAs differ, fails it."""
    changed = 0
    for c, q, t, kw in golden_cases():
        if c["name"].startswith("rand") and not (kw["flag"] & 1) and len(q) > 30 and len(t) > 30:
            kw2 = dict(kw)
            kw2["e"] = kw["e"] + 1
            if not same(O.run_oracle(q, t, **kw2), as_result(c["expect"])):
                changed += 1
    assert changed > 50
    differing = [(c, q, t, kw) for c, q, t, kw in golden_cases() if not c["sse_same"]]
    for c, q, t, kw in differing:
        assert not same(O.run_oracle(q, t, isa=O.ISA_SSE, **kw), as_result(c["expect"]))


# ------------------------------------------------------- Singletrack traceback
def _gotoh_h(qs, ts, mat, m, q, e, q2, e2):
    """Full textbook Gotoh matrix with two gap pieces, H[i+1][j+1] = H(i, j),
    row/column 0 the boundary; independent of the difference recurrence."""
    neg = -(1 << 40)
    tl, ql = len(ts), len(qs)

    def gap(k):
        return -min(q + k * e, q2 + k * e2) if k > 0 else 0

    H = [[0] * (ql + 1) for _ in range(tl + 1)]
    E1 = [[neg] * (ql + 1) for _ in range(tl + 1)]
    E2 = [[neg] * (ql + 1) for _ in range(tl + 1)]
    F1 = [[neg] * (ql + 1) for _ in range(tl + 1)]
    F2 = [[neg] * (ql + 1) for _ in range(tl + 1)]
    for j in range(1, ql + 1):
        H[0][j] = gap(j)
    for i in range(1, tl + 1):
        H[i][0] = gap(i)
        for j in range(1, ql + 1):
            E1[i][j] = max(H[i][j - 1] - q - e, E1[i][j - 1] - e)
            E2[i][j] = max(H[i][j - 1] - q2 - e2, E2[i][j - 1] - e2)
            F1[i][j] = max(H[i - 1][j] - q - e, F1[i - 1][j] - e)
            F2[i][j] = max(H[i - 1][j] - q2 - e2, F2[i - 1][j] - e2)
            H[i][j] = max(H[i - 1][j - 1] + int(mat[int(ts[i - 1]) * m + int(qs[j - 1])]),
                          E1[i][j], E2[i][j], F1[i][j], F2[i][j])
    return np.array(H, dtype=np.int64)


def test_singletrack_tracks_are_gotoh_differences():
    """The two stored tracks (minimap2/ksw2_extd2_avx.c:432-436: z - u and
    z - v) are the horizontal and vertical differences of H.  Wherever no
    int8 intermediate wraps and the match cap never binds, they equal the
    differences of an independent textbook Gotoh matrix, cell by cell."""
    rng = np.random.default_rng(31)
    checked = 0
    for _ in range(120):
        t = W.random_seq(rng, int(rng.integers(1, 40)))
        q = W.mutate(rng, t, 0.15, 0.08) if rng.random() < 0.7 else W.random_seq(rng, int(rng.integers(1, 40)))
        q = q if len(q) else W.random_seq(rng, 1)
        pen = (8, 2, 75, 1) if rng.random() < 0.6 else (int(rng.integers(1, 12)), int(rng.integers(1, 4)),
                                                       int(rng.integers(12, 40)), 1)
        mat = O.default_matrix(0, -int(rng.integers(1, 8)))
        kw = dict(mat=mat, m=5, q=pen[0], e=pen[1], q2=pen[2], e2=pen[3])
        with O.RangeProbe() as rp:
            res, tr = O.run_oracle(q, t, flag=ksw2.KSW_EZ_SINGLETRACK, want_dirs=True, **kw)
        if rp.wraps or rp.cap_binds or res.score <= -(1 << 29):
            continue
        H = _gotoh_h(q, t, np.asarray(mat).reshape(-1), 5, *pen)
        assert np.array_equal(tr[0], H[1:, 1:] - H[1:, :-1])
        assert np.array_equal(tr[1], H[1:, 1:] - H[:-1, 1:])
        checked += 1
    assert checked > 60


def test_singletrack_cigar_properties():
    """A Singletrack CIGAR (when the walk succeeds) consumes both sequences,
    never scores above the optimum, and is a single match run for identical
    sequences; a failed walk leaves no CIGAR and sets zdropped
    (minimap2/ksw2_singletrack.cpp:160-165, ksw2_extd2_avx.c:756-757)."""
    rng = np.random.default_rng(32)
    valid = failed = 0
    for _ in range(300):
        t = W.random_seq(rng, int(rng.integers(1, 150)))
        q = W.mutate(rng, t, 0.1, 0.05)
        q = q if len(q) else W.random_seq(rng, 1)
        kw = dict(mat=O.default_matrix(0, -6), m=5, q=8, e=2, q2=75, e2=1)
        r = O.run_oracle(q, t, flag=ksw2.KSW_EZ_SINGLETRACK, **kw)
        if r.n_cigar == 0:
            assert r.zdropped == 1
            failed += 1
            continue
        valid += 1
        assert O.cigar_lengths(r.cigar) == (len(q), len(t))
        assert O.cigar_score(r.cigar, q, t, **kw) <= O.gotoh2_global(q, t, **kw)
    assert valid > 200
    t = W.random_seq(rng, 300)
    r = O.run_oracle(t, t, flag=ksw2.KSW_EZ_SINGLETRACK, mat=O.default_matrix(0, -6))
    assert list(r.cigar) == [300 << 4] and r.score == 0


@pytest.mark.skipif(not O.ref_available(), reason="reference checker not built (oracle/_ref)")
def test_singletrack_oracle_vs_reference_random():
    """Live differential of the Singletrack path against the reference
    compiled from its sources, over bands, z-drops, scorings and flags."""
    rng = np.random.default_rng(33)
    for it in range(400):
        t = W.random_seq(rng, int(rng.integers(1, 300)))
        q = W.mutate(rng, t, 0.1, 0.04) if rng.random() < 0.7 else W.random_seq(rng, int(rng.integers(1, 300)))
        q = q if len(q) else W.random_seq(rng, 1)
        if rng.random() < 0.3:
            q = W.big_indel(rng, q, 5, max(6, len(q) // 2))
        flag = ksw2.KSW_EZ_SINGLETRACK | int(rng.choice([0, 0, 0x02, 0x80, 0x04, 0x40, 0x800, 0x01]))
        pen = (8, 2, 75, 1) if rng.random() < 0.5 else (int(rng.integers(0, 20)), int(rng.integers(1, 6)),
                                                       int(rng.integers(0, 90)), int(rng.integers(0, 3)))
        kw = dict(mat=O.default_matrix(int(rng.integers(0, 3)), -int(rng.integers(1, 9))),
                  q=pen[0], e=pen[1], q2=pen[2], e2=pen[3],
                  w=-1 if rng.random() < 0.6 else int(rng.integers(0, 120)),
                  zdrop=-1 if rng.random() < 0.8 else int(rng.integers(0, 150)), flag=flag)
        r = O.run_ref(q, t, isa="avx512", **kw)
        o = O.run_oracle(q, t, **kw)
        assert same(o, r), f"it={it} {kw['w']} {flag:#x}: {diff(o, r)}"


def test_planted_singletrack_mistake_is_caught():
    """The Singletrack golden vectors are sensitive to the walk: the same
    calls without the flag (the direction-byte backtrack) differ on many."""
    changed = n = 0
    for c, q, t, kw in golden_cases():
        if not (kw["flag"] & ksw2.KSW_EZ_SINGLETRACK) or kw["flag"] & 0x841:
            continue
        n += 1
        kw2 = dict(kw, flag=kw["flag"] & ~ksw2.KSW_EZ_SINGLETRACK)
        if not same(O.run_oracle(q, t, **kw2), as_result(c["expect"])):
            changed += 1
    assert n > 40 and changed > n // 4
