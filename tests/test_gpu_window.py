"""ksw2_hip_window_batch (the sliding-window chain) against the oracle's
restatement of alignSlidingWindow (oracle/window_py.py): status, rescored
score, call count and the end-to-end alignment, bit-exact, for a batch that
mixes window sizes' worth of chains, wildcards, mixed case, long gaps and
empty sides."""
import numpy as np
import pytest

import window_py as WP
from anchorwave_amd import ksw2

pytestmark = pytest.mark.gpu

PEN = dict(mis=-6, o1=-8, e1=-2, o2=-75, e2=-1)
GPU_PEN = dict(mismatch=-6, open1=-8, extend1=-2, open2=-75, extend2=-1)


def _dna(rng, n):
    return "".join("ACGT"[i] for i in rng.integers(0, 4, n))


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
        t = _dna(rng, int(rng.integers(20, 4000)))
        q = _mutate(rng, t, 0.05, 0.02)
        if k % 6 == 1:
            a = int(rng.integers(0, max(1, len(q) - 1)))
            q = q[:a] + q[a + int(rng.integers(50, 400)):]
        elif k % 6 == 2:
            q = "".join("N" if rng.random() < 0.02 else c for c in q)
        elif k % 6 == 3:
            q = "".join(c.lower() if rng.random() < 0.02 else c for c in q)
        elif k % 6 == 4:
            q = q + _dna(rng, int(rng.integers(100, 1500)))  # a long query tail
        tasks.append((q, t))
    tasks += [("", "ACGTACGT"), ("ACGT", ""), ("", ""), ("A", "A"), ("ACGT" * 100, "T")]
    return tasks


@pytest.mark.parametrize("window", [0, 37, 250, 1000])
def test_window_batch_matches_the_reference_logic(window):
    tasks = _tasks(90 + window, 60)
    got = ksw2.window_batch(tasks, window, **GPU_PEN)
    bad = []
    for k, ((q, t), g) in enumerate(zip(tasks, got)):
        e = WP.align(q, t, window, **PEN)
        if (g.status, g.score, g.windows, g.cigar) != (e.status, e.score, e.windows, e.cigar):
            bad.append(f"task {k} (qlen {len(q)}, tlen {len(t)}): got {(g.status, g.score, g.windows)} "
                       f"expected {(e.status, e.score, e.windows)}, cigars equal {g.cigar == e.cigar}")
    assert not bad, f"window {window}\n" + "\n".join(bad[:6]) + f"\n({len(bad)} of {len(tasks)} differ)"
    if window == 250:
        assert sum(g.windows for g in got) > 2 * len(tasks)  # real chains
