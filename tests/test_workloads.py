"""Synthetic workload generators: determinism and shape."""
import numpy as np

from anchorwave_amd import workloads as W


def test_deterministic():
    a, b = W.fixed(seed=2, n=8), W.fixed(seed=2, n=8)
    assert all((x[0] == y[0]).all() and (x[1] == y[1]).all() for x, y in zip(a, b))


def test_shapes_and_rates():
    (q, t), = W.single(seed=1)
    assert len(t) == 5000 and 4500 < len(q) < 5500
    assert set(np.unique(t)) <= {0, 1, 2, 3}
    fx = W.fixed(seed=2, n=200)
    assert all(len(t) == 2000 for _, t in fx)
    big = sum(1 for q, t in fx if abs(len(q) - len(t)) >= 45)
    assert 5 <= big <= 45  # ~10% carry a 50-300 bp indel
    un = W.uneven(seed=3, n=300)
    tl = np.array([len(t) for _, t in un])
    assert tl.min() >= 100 and tl.max() <= 50000 and 1500 < np.median(tl) < 6000
    ratios = np.array([len(q) / len(t) for q, t in un])
    assert ratios.min() > 0.3 and np.median(ratios) < 2.5  # 20% also carry a 1-20 kbp insertion


def test_mutate_rates():
    rng = np.random.default_rng(0)
    t = W.random_seq(rng, 200000)
    q = W.mutate(rng, t, 0.05, 0.0)
    assert abs((q != t).mean() - 0.05) < 0.005
    q = W.mutate(rng, t, 0.0, 0.02)
    assert abs(len(q) - len(t)) < 3000


def test_banded_cells_matches_reference_band():
    # w >= max(qlen, tlen): full matrix
    assert W.banded_cells(300, 200, -1) == 300 * 200
    assert W.banded_cells(300, 200, 400) == 300 * 200
    assert W.banded_cells(1000, 1000, 10) < 1000 * 25


def test_serving_mix_and_uneven_ratio():
    """Serving tasks: 50/40/10 single/fixed/uneven, Poisson arrivals, and the
    uneven tasks drawn like the uneven batch, query/target length ratio
    0.5-2.0 before mutation (so some queries are much shorter than their
    targets)."""
    ts = W.serving(seed=5, n=3000, rate_per_s=1000.0)
    kinds = [t.kind for t in ts]
    frac = {k: kinds.count(k) / len(ts) for k in ("single", "fixed", "uneven")}
    assert abs(frac["single"] - 0.5) < 0.04 and abs(frac["fixed"] - 0.4) < 0.04 and abs(frac["uneven"] - 0.1) < 0.03
    gaps = np.diff([t.arrival_s for t in ts])
    assert (gaps >= 0).all() and abs(gaps.mean() - 1e-3) < 1e-4
    ratios = np.array([len(t.query) / len(t.target) for t in ts if t.kind == "uneven"])
    assert ratios.min() < 0.7 and (ratios < 1.0).mean() > 0.2  # shortened queries occur
    assert np.median(ratios) > 0.8
