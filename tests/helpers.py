"""Shared helpers for the tests (inputs and field-by-field comparison)."""
from __future__ import annotations

import numpy as np

from anchorwave_amd import workloads as W

FIELDS = ("score", "max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q",
          "n_cigar", "reach_end", "stopped", "cigar")


def same(a, b) -> bool:
    return all(getattr(a, f) == getattr(b, f) for f in FIELDS)


def diff(a, b) -> str:
    return ", ".join(f"{f}: {getattr(a, f)!r} vs {getattr(b, f)!r}" for f in FIELDS
                     if getattr(a, f) != getattr(b, f))[:600]


def random_pairs(seed: int, n: int, lo: int, hi: int, mutated: float = 0.7):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        t = W.random_seq(rng, int(rng.integers(lo, hi + 1)))
        if rng.random() < mutated:
            q = W.mutate(rng, t, 0.08, 0.03)
            if rng.random() < 0.2:
                q = W.big_indel(rng, q, 5, max(6, len(q) // 3))
        else:
            q = W.random_seq(rng, int(rng.integers(lo, hi + 1)))
        if len(q) == 0:
            q = W.random_seq(rng, 1)
        out.append((q, t))
    return out
