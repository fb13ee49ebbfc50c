#!/usr/bin/env python3
"""Writes tests/golden/ksw2_golden.json.gz: input/output vectors of the
reference ksw_extd2_avx512 (with the SSE and AVX2 results recorded beside it),
produced by the library that oracle/Makefile compiles in place from the
reference sources (minimap2/ksw2_extd2_{sse,avx}.c).  Run in the build
container, where the reference tree exists:

    make -C oracle ref && python tests/golden/make_golden.py

Inputs: seeded random pairs over all flags / bands / z-drops / scorings, small
samples of the benchmark workloads, and the FASTA test data the reference
ships (minimap2/test/*.fa, copied to tests/golden/fasta/) encoded the way
AnchorWave's adapters encode (alignSlidingWindow.cpp:265-276: ACGT -> 0..3,
anything else -> 4 with KSW_EZ_GENERIC_SC).
"""
from __future__ import annotations

import gzip
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle_py as O  # noqa: E402
from anchorwave_amd import workloads as W  # noqa: E402

FLAG_CHOICES = [0, 0, 0, 0x02, 0x01, 0x80, 0x800, 0x802, 0x40, 0x04, 0x08, 0x18, 0x801, 0x40 | 0x80]


def read_fasta(path):
    recs, name, buf = [], None, []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith(">"):
                if name is not None:
                    recs.append((name, "".join(buf)))
                name, buf = line[1:].split()[0], []
            elif line:
                buf.append(line)
    if name is not None:
        recs.append((name, "".join(buf)))
    return recs


def encode(s: str):
    table = {"A": 0, "C": 1, "G": 2, "T": 3, "a": 0, "c": 1, "g": 2, "t": 3}
    codes = np.array([table.get(ch, 4) for ch in s], dtype=np.uint8)
    return codes, bool((codes == 4).any())


def case(name, q, t, mat, m, q_, e, q2, e2, w, zdrop, end_bonus, flag, inline=True, src=None):
    kw = dict(mat=mat, m=m, q=q_, e=e, q2=q2, e2=e2, w=w, zdrop=zdrop, end_bonus=end_bonus, flag=flag)
    r512 = O.run_ref(q, t, isa="avx512", **kw)
    rsse = O.run_ref(q, t, isa="sse", **kw)
    ravx2 = O.run_ref(q, t, isa="avx2", **kw)
    c = {"name": name, "m": m, "mat": [int(x) for x in np.asarray(mat).reshape(-1)], "q": q_, "e": e,
         "q2": q2, "e2": e2, "w": w, "zdrop": zdrop, "end_bonus": end_bonus, "flag": flag,
         "expect": r512.as_dict(), "sse_same": rsse == r512, "avx2_same": ravx2 == r512}
    if inline:
        c["query"] = "".join(map(str, q.tolist()))
        c["target"] = "".join(map(str, t.tolist()))
    else:
        c["source"] = src
    return c


def main():
    assert O.ref_available(), "build the reference checker first: make -C oracle ref"
    rng = np.random.default_rng(20261202)
    cases = []
    for k in range(360):
        tl = int(rng.integers(1, 400))
        t = W.random_seq(rng, tl)
        if rng.random() < 0.65:
            q = W.mutate(rng, t, 0.08, 0.03)
            if rng.random() < 0.2:
                q = W.big_indel(rng, q, 5, max(6, tl // 2))
        else:
            q = W.random_seq(rng, int(rng.integers(1, 400)))
        if len(q) == 0:
            q = W.random_seq(rng, 1)
        flag = int(rng.choice(FLAG_CHOICES))
        if rng.random() < 0.15:
            t = t.copy()
            t[rng.integers(0, len(t), max(1, len(t) // 50))] = 4
        w = -1 if rng.random() < 0.55 else int(rng.integers(0, 160))
        zdrop = -1 if rng.random() < 0.7 else int(rng.integers(0, 150))
        if rng.random() < 0.6:
            m, mat, pen = 5, O.default_matrix(), (8, 2, 75, 1)
        else:
            ma, mm = int(rng.integers(0, 4)), -int(rng.integers(1, 10))
            m, mat = 5, O.default_matrix(ma, mm)
            pen = (int(rng.integers(0, 20)), int(rng.integers(1, 6)), int(rng.integers(0, 90)),
                   int(rng.integers(0, 3)))
        if flag & 0x04 and rng.random() < 0.5:  # a non-uniform generic matrix
            mat = rng.integers(-8, 3, 25).astype(np.int8)
        cases.append(case(f"rand{k}", q, t, mat, m, *pen, w, zdrop, int(rng.integers(0, 30)), flag))
    for k, (q, t) in enumerate(W.fixed(seed=2, n=6)):
        cases.append(case(f"fixed{k}", q, t, O.default_matrix(), 5, 8, 2, 75, 1, -1, -1, 0, 0))
    for k, (q, t) in enumerate(W.single(seed=1)):
        cases.append(case(f"single{k}", q, t, O.default_matrix(), 5, 8, 2, 75, 1, -1, -1, 0, 0))
    # reference test data (FASTA), AnchorWave encoding
    fa = {os.path.basename(p): read_fasta(os.path.join(HERE, "fasta", p))
          for p in ("MT-human.fa", "MT-orang.fa", "q2.fa", "t2.fa", "q-inv.fa", "t-inv.fa")}
    pairs = [("MT-orang.fa", 0, "MT-human.fa", 0), ("q2.fa", 0, "t2.fa", 0),
             ("q-inv.fa", 0, "t-inv.fa", 0), ("q-inv.fa", 1, "t-inv.fa", 0)]
    for qf, qi, tf, ti in pairs:
        qs, qn = encode(fa[qf][qi][1])
        ts, tn = encode(fa[tf][ti][1])
        gflag = 0x04 if (qn or tn) else 0
        src = {"query": [qf, qi], "target": [tf, ti]}
        for flag in (0, 0x02, 0x01, 0x800):
            cases.append(case(f"{qf}:{qi}~{tf}:{ti}/{flag:#x}", qs, ts, O.default_matrix(), 5, 8, 2, 75, 1,
                              -1, -1, 0, flag | gflag, inline=False, src=src))
        cases.append(case(f"{qf}:{qi}~{tf}:{ti}/w500", qs, ts, O.default_matrix(), 5, 8, 2, 75, 1,
                          500, -1, 0, gflag, inline=False, src=src))
    # KSW_EZ_SINGLETRACK (the two-track traceback, minimap2/ksw2_singletrack.cpp),
    # drawn from a separate stream so the cases above stay as they were
    rng2 = np.random.default_rng(20261203)
    st_flags = [0x1000, 0x1000, 0x1000, 0x1002, 0x1080, 0x1004, 0x1040, 0x1800, 0x1001]
    for k in range(120):
        tl = int(rng2.integers(1, 400))
        t = W.random_seq(rng2, tl)
        if rng2.random() < 0.7:
            q = W.mutate(rng2, t, 0.08, 0.03)
            if rng2.random() < 0.3:
                q = W.big_indel(rng2, q, 5, max(6, tl // 2))
        else:
            q = W.random_seq(rng2, int(rng2.integers(1, 400)))
        if len(q) == 0:
            q = W.random_seq(rng2, 1)
        flag = int(rng2.choice(st_flags))
        if rng2.random() < 0.15:
            t = t.copy()
            t[rng2.integers(0, len(t), max(1, len(t) // 50))] = 4
            flag |= 0x04
        w = -1 if rng2.random() < 0.6 else int(rng2.integers(0, 160))
        zdrop = -1 if rng2.random() < 0.85 else int(rng2.integers(0, 150))
        if rng2.random() < 0.6:
            m, mat, pen = 5, O.default_matrix(), (8, 2, 75, 1)
        else:
            m, mat = 5, O.default_matrix(0, -int(rng2.integers(1, 10)))
            pen = (int(rng2.integers(0, 20)), int(rng2.integers(1, 6)), int(rng2.integers(0, 90)),
                   int(rng2.integers(0, 3)))
        cases.append(case(f"st{k}", q, t, mat, m, *pen, w, zdrop, 0, flag))
    for k, (q, t) in enumerate(W.fixed(seed=2, n=3)):
        cases.append(case(f"st_fixed{k}", q, t, O.default_matrix(), 5, 8, 2, 75, 1, -1, -1, 0, 0x1000))
    for qf, qi, tf, ti in pairs[:2]:
        qs, qn = encode(fa[qf][qi][1])
        ts, tn = encode(fa[tf][ti][1])
        gflag = 0x04 if (qn or tn) else 0
        src = {"query": [qf, qi], "target": [tf, ti]}
        cases.append(case(f"{qf}:{qi}~{tf}:{ti}/st", qs, ts, O.default_matrix(), 5, 8, 2, 75, 1,
                          -1, -1, 0, 0x1000 | gflag, inline=False, src=src))
        cases.append(case(f"{qf}:{qi}~{tf}:{ti}/st-w500", qs, ts, O.default_matrix(), 5, 8, 2, 75, 1,
                          500, -1, 0, 0x1000 | gflag, inline=False, src=src))
    out = os.path.join(HERE, "ksw2_golden.json.gz")
    with gzip.open(out, "wt") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "reference": "ksw_extd2_avx512",
                   "cases": cases}, f)
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main()
