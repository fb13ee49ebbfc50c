"""Seeded synthetic inputs for the five ksw_extd2 workloads (codes A,C,G,T = 0..3).

  single    one 5,000 bp target; query with 8% substitutions, 2% indels
  fixed     4,096 pairs of 2,000 bp; 5% subs, 1% indels, 10% with one 50-300 bp indel
  uneven    20,000 pairs, target length log-normal (median 3 kbp, sigma 1,
            clipped to 100 bp..50 kbp), query/target length ratio U(0.5, 2);
            6% subs, 1.5% indels, 20% with one 1-20 kbp random insertion
  banded    64 pairs of 100-300 kbp; 10% subs, 2% indels, 2-5 SVs of 5-50 kbp; w = 4096
  serving   interval tasks drawn 50/40/10 from single / fixed / uneven

Short indels are 50/50 insertion/deletion with geometric lengths of mean 3
(p = 1/3); inserted bases are uniform.  Every generator takes a seed and
returns lists of uint8 arrays (query, target) - the orientation of the
reference call ksw_extd2(..., qlen, query, tlen, target, ...).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

Pair = Tuple[np.ndarray, np.ndarray]

# AnchorWave genoAli defaults (controlLayer.cpp:297-303), negated as the
# adapters pass them (alignSlidingWindow.cpp:841-845).
SCORING = dict(m=5, q=8, e=2, q2=75, e2=1)
MATCH, MISMATCH = 0, -6
BANDED_W = 4096


def random_seq(rng: np.random.Generator, n: int) -> np.ndarray:
    return rng.integers(0, 4, n, dtype=np.uint8)


def mutate(rng: np.random.Generator, t: np.ndarray, sub: float, indel: float,
           mean_len: float = 3.0) -> np.ndarray:
    """Substitutions (uniform among the 3 other bases), then short indels at
    per-base rate ``indel`` (half insertions, half deletions, geometric
    lengths with the given mean)."""
    n = len(t)
    out = t.copy()
    m = rng.random(n) < sub
    out[m] = (out[m] + rng.integers(1, 4, int(m.sum()), dtype=np.uint8)) % 4
    ev = np.flatnonzero(rng.random(n) < indel)
    if len(ev) == 0:
        return out
    is_ins = rng.random(len(ev)) < 0.5
    lens = rng.geometric(1.0 / mean_len, len(ev))
    keep = np.ones(n, dtype=bool)
    dpos, dlen = ev[~is_ins], lens[~is_ins]
    if len(dpos):
        idx = np.repeat(dpos, dlen) + (np.arange(dlen.sum()) - np.repeat(np.cumsum(dlen) - dlen, dlen))
        keep[idx[idx < n]] = False
    ipos, ilen = ev[is_ins], lens[is_ins]
    newpos = np.cumsum(keep) - keep  # position of each original base after deletions
    base = out[keep]
    if len(ipos):
        at = np.repeat(newpos[ipos], ilen)
        base = np.insert(base, at, random_seq(rng, int(ilen.sum())))
    return base


def big_indel(rng: np.random.Generator, s: np.ndarray, lo: int, hi: int,
              ins_only: bool = False) -> np.ndarray:
    L = int(rng.integers(lo, hi + 1))
    pos = int(rng.integers(0, len(s) + 1))
    if ins_only or rng.random() < 0.5:
        return np.concatenate([s[:pos], random_seq(rng, L), s[pos:]])
    return np.concatenate([s[:pos], s[pos + L:]])


def _nonempty(rng, s: np.ndarray) -> np.ndarray:
    return s if len(s) else random_seq(rng, 1)


def single(seed: int = 1, tlen: int = 5000) -> List[Pair]:
    rng = np.random.default_rng(seed)
    t = random_seq(rng, tlen)
    q = mutate(rng, t, 0.08, 0.02)
    return [(_nonempty(rng, q), t)]


def fixed(seed: int = 2, n: int = 4096, tlen: int = 2000) -> List[Pair]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        t = random_seq(rng, tlen)
        q = mutate(rng, t, 0.05, 0.01)
        if rng.random() < 0.10:
            q = big_indel(rng, q, 50, 300)
        out.append((_nonempty(rng, q), t))
    return out


def uneven_lengths(rng: np.random.Generator, n: int) -> np.ndarray:
    return np.clip(rng.lognormal(np.log(3000.0), 1.0, n), 100, 50000).astype(np.int64)


def _uneven_pair(rng: np.random.Generator, tlen: int, ratio: float) -> Pair:
    """One inter-anchor interval: target of ``tlen`` bases, query realised at
    ``ratio`` times that length (one uniform segment deleted or inserted),
    then 6% substitutions, 1.5% short indels and, for 20% of pairs, one
    1-20 kbp random insertion."""
    t = random_seq(rng, int(tlen))
    want = max(1, int(round(tlen * ratio)))
    q = t.copy()
    if want < len(q):
        d = len(q) - want
        pos = int(rng.integers(0, len(q) - d + 1))
        q = np.concatenate([q[:pos], q[pos + d:]])
    elif want > len(q):
        pos = int(rng.integers(0, len(q) + 1))
        q = np.concatenate([q[:pos], random_seq(rng, want - len(q)), q[pos:]])
    q = mutate(rng, q, 0.06, 0.015)
    if rng.random() < 0.20:
        q = big_indel(rng, q, 1000, 20000, ins_only=True)
    return _nonempty(rng, q), t


def uneven(seed: int = 3, n: int = 20000) -> List[Pair]:
    rng = np.random.default_rng(seed)
    tl = uneven_lengths(rng, n)
    ratio = rng.uniform(0.5, 2.0, n)
    return [_uneven_pair(rng, int(tl[k]), float(ratio[k])) for k in range(n)]


def banded(seed: int = 4, n: int = 64, lo: int = 100_000, hi: int = 300_000) -> List[Pair]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        t = random_seq(rng, int(rng.integers(lo, hi + 1)))
        q = mutate(rng, t, 0.10, 0.02)
        for _ in range(int(rng.integers(2, 6))):
            q = big_indel(rng, q, 5000, 50000)
        out.append((_nonempty(rng, q), t))
    return out


@dataclass
class Task:
    arrival_s: float
    query: np.ndarray
    target: np.ndarray
    kind: str


def serving(seed: int = 5, n: int = 2000, rate_per_s: float = 100.0) -> List[Task]:
    """Poisson arrivals; sizes drawn 50/40/10 from single/fixed/uneven (the
    uneven tasks with the uneven batch's 0.5-2.0 query/target length ratio)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, n)
    tasks = []
    tnow = 0.0
    for k in range(n):
        tnow += gaps[k]
        u = rng.random()
        if u < 0.5:
            t = random_seq(rng, 5000)
            q = mutate(rng, t, 0.08, 0.02)
            kind = "single"
        elif u < 0.9:
            t = random_seq(rng, 2000)
            q = mutate(rng, t, 0.05, 0.01)
            if rng.random() < 0.10:
                q = big_indel(rng, q, 50, 300)
            kind = "fixed"
        else:
            tl = int(uneven_lengths(rng, 1)[0])
            q, t = _uneven_pair(rng, tl, float(rng.uniform(0.5, 2.0)))
            kind = "uneven"
        tasks.append(Task(tnow, _nonempty(rng, q), t, kind))
    return tasks


def cells(pairs: List[Pair]) -> int:
    """Algorithmic cells of unbanded runs: sum of qlen * tlen."""
    return int(sum(len(q) * len(t) for q, t in pairs))


def banded_cells(qlen: int, tlen: int, w: int) -> int:
    """In-band cells of ksw_extd2's anti-diagonal band, sum of en0-st0+1
    (minimap2/ksw2_extd2_sse.c:147-150)."""
    if w < 0:
        w = max(qlen, tlen)
    r = np.arange(qlen + tlen - 1, dtype=np.int64)
    st = np.maximum(np.maximum(0, r - qlen + 1), (r - w + 1) >> 1)
    en = np.minimum(np.minimum(tlen - 1, r), (r + w) >> 1)
    ok = st <= en
    # the reference stops at the first empty anti-diagonal
    if not ok.all():
        stop = int(np.argmin(ok))
        st, en = st[:stop], en[:stop]
    return int(np.sum(en - st + 1))
