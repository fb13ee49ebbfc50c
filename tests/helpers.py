"""Shared helpers for the tests (inputs and field-by-field comparison)."""
from __future__ import annotations

import gzip
import json
import os

import numpy as np

import oracle_py as O

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


HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "ksw2_golden.json.gz")


def read_fasta(path):
    recs, name, buf = [], None, []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith(">"):
                if name is not None:
                    recs.append("".join(buf))
                name, buf = line[1:], []
            elif line:
                buf.append(line)
    if name is not None:
        recs.append("".join(buf))
    return recs


def encode(s):
    t = {"A": 0, "C": 1, "G": 2, "T": 3, "a": 0, "c": 1, "g": 2, "t": 3}
    return np.array([t.get(ch, 4) for ch in s], dtype=np.uint8)


def golden_cases():
    with gzip.open(GOLDEN, "rt") as f:
        data = json.load(f)
    fa = {}
    for c in data["cases"]:
        if "query" in c:
            q = np.frombuffer(c["query"].encode(), np.uint8) - 48
            t = np.frombuffer(c["target"].encode(), np.uint8) - 48
        else:
            (qf, qi), (tf, ti) = c["source"]["query"], c["source"]["target"]
            for fn in (qf, tf):
                if fn not in fa:
                    fa[fn] = read_fasta(os.path.join(HERE, "golden", "fasta", fn))
            q, t = encode(fa[qf][qi]), encode(fa[tf][ti])
        kw = {k: c[k] for k in ("m", "q", "e", "q2", "e2", "w", "zdrop", "end_bonus", "flag")}
        kw["mat"] = np.array(c["mat"], dtype=np.int8)
        yield c, q.astype(np.uint8), t.astype(np.uint8), kw


def as_result(d):
    d = dict(d)
    d["cigar"] = tuple(d["cigar"])
    return O.Result(**d)
