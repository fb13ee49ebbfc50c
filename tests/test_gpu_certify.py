"""ksw2_hip_certify_batch (score-only plus score-certified banding) against
the oracle's restatement of alignScoreCertifiedKsw2 (oracle/certify_py.py):
every task's status, score, optimal score, final band, attempt count, the
certified CIGAR and the best-effort CIGAR, bit-exact, in one batch mixing
every case."""
import numpy as np
import pytest

import certify_py as CP
from anchorwave_amd import ksw2

pytestmark = pytest.mark.gpu

PEN = dict(mis=-6, o1=-8, e1=-2, o2=-75, e2=-1)
GPU_PEN = dict(mismatch=-6, open1=-8, extend1=-2, open2=-75, extend2=-1)


def _dna(rng, n, alphabet="ACGT"):
    return "".join(alphabet[i] for i in rng.integers(0, len(alphabet), n))


def _mutate(rng, s, sub, indel):
    out = []
    for c in s:
        r = rng.random()
        if r < sub:
            out.append("ACGT"[int(rng.integers(0, 4))])
        elif r < sub + indel / 2:
            continue
        elif r < sub + indel:
            out.append(c + "ACGT"[int(rng.integers(0, 4))])
        else:
            out.append(c)
    return "".join(out)


def _tasks(seed, n):
    rng = np.random.default_rng(seed)
    tasks = []
    for k in range(n):
        kind = k % 8
        t = _dna(rng, int(rng.integers(30, 1200)))
        q = _mutate(rng, t, 0.06, 0.03)
        if kind == 1:  # a long gap: the band must widen past it
            a = int(rng.integers(0, len(t) // 2))
            q = q[:a] + q[a + int(rng.integers(40, 200)):]
        elif kind == 2:  # wildcards
            q = "".join("N" if rng.random() < 0.03 else c for c in q)
        elif kind == 3:  # mixed case (rescoring compares characters)
            q = "".join(c.lower() if rng.random() < 0.01 else c for c in q)
        elif kind == 4:  # unrelated sequences
            q = _dna(rng, int(rng.integers(30, 600)))
        initial = int(rng.choice([0, 1, 5, 30, 100, 400]))
        diff = abs(len(q) - len(t))
        maximum = int(rng.choice([diff, diff + 7, 256, 1000, 1 << 20, max(diff - 1, 0)]))
        pre = None
        if kind == 5:
            pre = CP.score_only(q, t, **PEN)  # a correct precomputed optimum
        elif kind == 6:
            pre = 5  # unreachable: never certifies, every band is tried
        tasks.append((q, t, initial, maximum, pre))
    # edge cases
    tasks += [("", "ACGTAC", 3, 10, None), ("ACG", "", 3, 10, None), ("", "", 1, 1, None),
              ("ACGTACGTAC", "ACGTA", 5, 2, None), ("ACGT", "ACGT", -1, 5, None),
              ("ACGT", "ACGT", 1, 10, CP.KSW_NEG_INF)]
    return tasks


def _compare(tasks, got):
    bad = []
    for k, (task, g) in enumerate(zip(tasks, got)):
        q, t, a, b, pre = task[:5]
        anchors = task[5] if len(task) > 5 else None
        e = CP.certify(q, t, a, b, precomputed=pre, anchors=anchors, **PEN)
        gd = (g.status, g.score, g.optimal_score, g.final_band_width, g.traceback_attempts, list(g.cigar),
              list(g.best_effort_cigar))
        ed = (e.status, e.score, e.optimal_score, e.final_band_width, e.traceback_attempts, list(e.cigar),
              list(e.best_effort_cigar))
        if gd != ed:
            bad.append(f"task {k} (qlen {len(q)}, tlen {len(t)}, bands {a}/{b}, pre {pre}): "
                       f"got {gd[:5]} expected {ed[:5]}, cigars equal {gd[5:] == ed[5:]}")
    return bad


def test_certify_batch_matches_the_reference_logic():
    tasks = _tasks(81, 160)
    got = ksw2.certify_batch(tasks, **GPU_PEN)
    bad = _compare(tasks, got)
    statuses = {s: sum(1 for g in got if g.status == s) for s in range(4)}
    assert statuses[CP.COMPLETED] > 50 and statuses[CP.NOT_CERTIFIED] > 10, statuses
    assert not bad, "\n".join(bad[:6]) + f"\n({len(bad)} of {len(tasks)} differ)"


def test_certify_batch_wide_bands_take_the_band_kernel():
    """5 kbp intervals with a 600-900 bp indel: the certifying bands are
    512 wide and more, where banded calls run on the band kernel."""
    rng = np.random.default_rng(82)
    tasks = []
    for k in range(4):
        t = _dna(rng, 5000)
        q = _mutate(rng, t, 0.05, 0.01)
        a = int(rng.integers(500, 3000))
        q = q[:a] + q[a + int(rng.integers(600, 900)):]
        tasks.append((q, t, 1, 1 << 20, None))
    qc, _ = CP.encode(tasks[0][0])
    tc, _ = CP.encode(tasks[0][1])
    assert ksw2.route(qc, tc, w=2048) == 1  # band only: the band kernel at w >= 512
    got = ksw2.certify_batch(tasks, **GPU_PEN)
    bad = _compare(tasks, got)
    assert all(g.status == CP.COMPLETED and g.final_band_width >= 512 for g in got), [
        (g.status, g.final_band_width) for g in got]
    assert not bad, "\n".join(bad)


def test_certify_batch_deadline():
    """A deadline already passed: every task that needs device work ends
    with TimeLimit; the empty-side task completes without it."""
    rng = np.random.default_rng(83)
    t = _dna(rng, 3000)
    tasks = [(_mutate(rng, t, 0.05, 0.02), t, 10, 1 << 20), ("", "ACGT", 1, 4)]
    got = ksw2.certify_batch(tasks, max_seconds=1e-9, **GPU_PEN)
    assert got[0].status == CP.TIME_LIMIT
    assert got[1].status == CP.COMPLETED


def _exact_runs(cigar, q, t, min_len):
    out, qp, tp = [], 0, 0
    for c in cigar:
        ln, op = c >> 4, c & 0xF
        if op == 0:
            k = 0
            while k < ln:
                if q[qp + k] == t[tp + k]:
                    s = k
                    while k < ln and q[qp + k] == t[tp + k]:
                        k += 1
                    if k - s >= min_len:
                        out.append((qp + s, tp + s, min(k - s, 65535)))
                else:
                    k += 1
            qp, tp = qp + ln, tp + ln
        elif op == 1:
            qp += ln
        else:
            tp += ln
    return out


def test_certify_batch_with_selection_plans():
    """Tasks with a selection plan take the chained traceback first
    (tryChainedKsw2Traceback): anchors from an optimal alignment certify at
    band -2; perturbed plans (anchors shifted off the path, mismatching,
    overlapping, too few, gaps wider than the band) fall through to the
    banded loop; every field equals the restatement's."""
    rng = np.random.default_rng(84)
    tasks = []
    for k in range(36):
        t = _dna(rng, int(rng.integers(300, 4000)))
        q = _mutate(rng, t, 0.05, 0.02)
        if k % 4 == 1:
            a = int(rng.integers(0, len(t) // 2))
            q = q[:a] + q[a + int(rng.integers(50, 300)):]
        if k % 6 == 2:
            q = "".join("N" if rng.random() < 0.02 else c for c in q)
        base = CP.certify(q, t, 0, 1 << 20, **PEN)
        anchors = _exact_runs(base.cigar, q, t, 10)
        kind = k % 5
        if kind == 1 and len(anchors) > 3:  # one anchor moved off the path
            aq, ar, al = anchors[len(anchors) // 2]
            anchors[len(anchors) // 2] = (aq + 1, ar, al)
        elif kind == 2:
            anchors = anchors[:1]
        elif kind == 3:  # narrow maximum band: some gaps may exceed it
            tasks.append((q, t, 3, max(abs(len(q) - len(t)), 40), None, anchors))
            continue
        elif kind == 4:  # overlapping duplicates are skipped
            anchors = sorted(anchors + anchors[::3])
        tasks.append((q, t, int(rng.choice([0, 5, 40])), 1 << 20, None, anchors))
    tasks.append(("ACGTACGTACGT", "ACGTACGTACGT", 2, 64, None, []))  # a plan with no anchors
    got = ksw2.certify_batch(tasks, **GPU_PEN)
    bad = _compare(tasks, got)
    chained = sum(1 for g in got if g.final_band_width == -2)
    assert chained >= 10, chained
    assert not bad, "\n".join(bad[:6]) + f"\n({len(bad)} of {len(tasks)} differ)"
