#!/usr/bin/env python3
"""Autotuner for the stripe kernel's lone-pair layout on the GPU it runs on.

A plan that leaves most of the device idle (a single ksw_extd2_hip call, a
few pairs) is latency-bound: its time is the chain of passes, so rows per
lane (R, the height of a lane's tile: 8 or 4), the team size (waves sharing
the pair's passes) and the team's layout (one workgroup on one CU, or
workgroups of four waves spread over CUs) decide it.  For each bin of swept
and striped lengths (the kernel's orientation: the striped sequence is the
one with more passes) this script times every candidate layout on a
representative pair, beside the layout the planner's own rules pick, and
writes the fastest to a table when it beats the planner's by at least 5 %
(--gain), one line per bin:

    bin <swept_lo> <swept_hi> <striped_lo> <striped_hi> <rows> <team> <passes> <spread>

where passes is the timed pair's pass count at those rows: the library
scales the team to a pair's own passes (team * its passes / passes, at most
one wave per pass), so "one wave per pass" carries over to every pair of the
bin.

The library reads the table (anchorwave_amd/tuning/<arch>.txt, or the file
KSW2_HIP_TUNING names; KSW2_HIP_TUNING=0 ignores it) and uses it for plans
in that regime; the planner's own rules apply elsewhere and whenever a bin
is missing.  Candidates are forced with the library's tuning switches
(KSW2_HIP_ROWS, KSW2_HIP_TEAM, KSW2_HIP_WIDE), each result is checked
against the first candidate's, and the table records the measured times.

    python -m anchorwave_amd.autotune [--out PATH] [--reps N]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from anchorwave_amd import ksw2, workloads as W  # noqa: E402

# bin edges of the swept and striped lengths, and the length timed per bin
EDGES = [0, 1500, 3000, 6000, 12000, 25000, 1 << 30]
CENTRES = [1000, 2200, 4500, 9000, 18000, 40000]
SWITCHES = ("KSW2_HIP_ROWS", "KSW2_HIP_TEAM", "KSW2_HIP_WIDE", "KSW2_HIP_TUNING")


def passes(striped: int, rows: int) -> int:
    stripes = (striped + 64 * rows - 1) // (64 * rows)
    return (stripes + 1) // 2


def candidates(striped: int):
    """(rows, team, spread) layouts worth timing for a pair of this striped
    length: one-workgroup teams of 1..16 waves (powers of two and the pass
    count), and teams of 8..64 waves spread over CUs in workgroups of four."""
    out = []
    for rows in (8, 4):
        p = passes(striped, rows)
        teams = sorted({t for t in (1, 2, 4, 8, 16) if t <= p} | ({p} if p <= 16 else set()))
        out += [(rows, t, 0) for t in teams]
        wide = sorted({t for t in (8, 16, 32, 64) if t < p + 4} | {min(64, (p + 3) // 4 * 4)})
        out += [(rows, t, 1) for t in wide if t >= 8]
    return out


def force(rows: int, team: int, spread: int) -> None:
    """Forces one layout; rows 0 leaves the planner's own rules (no table)."""
    os.environ["KSW2_HIP_TUNING"] = "0"
    for k in ("KSW2_HIP_ROWS", "KSW2_HIP_TEAM", "KSW2_HIP_WIDE"):
        os.environ.pop(k, None)
    if rows == 0:
        return
    os.environ["KSW2_HIP_ROWS"] = str(rows)
    if spread:
        os.environ["KSW2_HIP_WIDE"] = str(team // 4)
    else:
        os.environ["KSW2_HIP_TEAM"] = str(team)


def time_layout(pair, reps: int) -> tuple[float, object]:
    plan = ksw2.Plan([pair], mat=ksw2.default_matrix(W.MATCH, W.MISMATCH))
    plan.run()  # first run: lane streams, code objects
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.run()
        best = min(best, time.perf_counter() - t0)
    res = plan.fetch()[0]
    plan.close()
    return best, res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "gfx950.txt"))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--max-len", type=int, default=1 << 30, help="skip bins whose timed length exceeds this")
    ap.add_argument("--gain", type=float, default=0.95,
                    help="a bin's entry is written only when its best layout takes at most this fraction "
                         "of the planner's own layout's time")
    args = ap.parse_args()
    saved = {k: os.environ.get(k) for k in SWITCHES}
    rng = np.random.default_rng(17)
    lines = []
    for i, kq in enumerate(CENTRES):
        for j, kt in enumerate(CENTRES):
            if j < i or max(kq, kt) > args.max_len:
                continue  # the striped sequence is the one with more passes
            t = W.random_seq(rng, kt)
            q = W.mutate(rng, W.random_seq(rng, kq), 0.06, 0.015) if kq != kt else W.mutate(rng, t, 0.06, 0.015)
            pair = (q, t) if len(q) <= len(t) else (q[:kt], t)
            timed = []
            ref = None
            for rows, team, spread in [(0, 0, 0)] + candidates(kt):
                force(rows, team, spread)
                sec, res = time_layout(pair, args.reps)
                if ref is None:
                    ref = res
                elif res.__dict__ != ref.__dict__:
                    raise SystemExit(f"layout {rows}/{team}/{spread} changed the result of bin {kq}x{kt}")
                timed.append((sec, rows, team, spread))
            planner = timed[0][0]
            timed = sorted(timed[1:])
            sec, rows, team, spread = timed[0]
            base = next(s for s, r, tm, sp in timed if (r, tm, sp) == (8, 1, 0))
            print(f"[autotune] swept {len(pair[0])} striped {len(pair[1])}: rows {rows} team {team} "
                  f"{'spread' if spread else 'one workgroup'} {sec * 1e3:.2f} ms (planner's own layout "
                  f"{planner * 1e3:.2f} ms, one wave of eight rows {base * 1e3:.2f} ms; {len(timed)} layouts)",
                  flush=True)
            where = f"bin {EDGES[i]} {EDGES[i + 1]} {EDGES[j]} {EDGES[j + 1]}"
            if sec <= args.gain * planner:
                lines.append(f"{where} {rows} {team} {passes(len(pair[1]), rows)} {spread}"
                             f"  # {sec * 1e3:.2f} ms; planner {planner * 1e3:.2f} ms; one wave of eight rows "
                             f"{base * 1e3:.2f} ms")
            else:
                lines.append(f"# {where}: the planner's layout kept ({planner * 1e3:.2f} ms; best forced "
                             f"rows {rows} team {team} spread {spread} {sec * 1e3:.2f} ms)")
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("# Stripe-kernel layout of a lone pair per bin of swept and striped lengths, timed on this\n"
                "# GPU by anchorwave_amd/autotune.py: bin swept_lo swept_hi striped_lo striped_hi rows team "
                "passes spread\n")
        f.write("\n".join(lines) + "\n")
    print(f"[autotune] wrote {args.out} ({len(lines)} bins)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
