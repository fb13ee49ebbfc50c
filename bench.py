#!/usr/bin/env python3
"""Benchmark: GCUPS of batched ksw_extd2 (score + CIGAR, bit-exact) on one MI355X.

Workload (one step): the fixed batch - 4,096 pairs, target 2,000 bp uniform
ACGT, query with 5% substitutions, 1% geometric indels and 10% of pairs with
one 50-300 bp indel; AnchorWave scoring (match 0, mismatch -6, q/e 8/2,
q2/e2 75/1), global alignment, full CIGAR, all pairs in one launch.  Inputs are
uploaded to HBM before the timed region; a step is the DP sweep plus the
device backtrack that leaves every score and CIGAR in HBM.

Prints one JSON line (rank 0).  Extra legs:
  cpu_baseline  the reference ksw_extd2_avx512 (oracle/_ref, compiled from the
                reference sources) over the same 4,096 pairs on all host cores
                (or the C restatement when that build is absent);
  exact         every GPU result compared field by field with those CPU results;
  one_at_a_time the same kernels launched on single-pair batches.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# Hardware queues of this process (read by the HIP runtime at its start, so
# set before the library loads; an inherited value, often HIP's default of
# 4, is replaced).  The library runs one stream per concurrent flight
# (serving) or calling thread (ksw_extd2_hip); with 4 queues the fifth such
# stream waits behind another's kernels.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np  # noqa: E402

from anchorwave_amd import ksw2, workloads as W  # noqa: E402

METRIC = "GCUPS, batched ksw_extd2 score+CIGAR (bit-exact); p50/p99 task latency"
# Algorithmic integer work per cell of the reference inner loop with
# traceback (minimap2/ksw2_extd2_sse.c:237-283): 51 ops.
OPS_PER_CELL = 51
N_CU, LANES, CLK_GHZ = 256, 64, 2.4
PEAK_PACKED16_TOPS = N_CU * LANES * CLK_GHZ * 2 / 1e3   # 78.6 T int16 lane-ops/s
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def read_profile_value(name: str, key: str):
    """A figure of a committed rocprofv3 summary (a line 'key: N'), or None:
    traffic_bytes_per_dispatch (HBM bytes per dispatch of the DP kernel,
    FETCH_SIZE and WRITE_SIZE scaled by a copy of known size) and
    valu_issue_frac (SQ_ACTIVE_INST_VALU over the busy SIMD cycles)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            m = re.search(key + r":\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def host_cpu() -> dict:
    """Model name and the widest vector ISA of the host (the reference's
    compile-time dispatch, alignSlidingWindow.cpp:840-846, takes AVX-512 when
    the host has AVX512BW)."""
    model, flags = "unknown", ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = line.split(":", 1)[1]
    except OSError:
        pass
    fl = set(flags.split())
    isa = "AVX-512BW" if "avx512bw" in fl else "AVX2" if "avx2" in fl else "SSE4.1" if "sse4_1" in fl else "SSE2"
    return {"cpu_model": model, "isa": isa}


def cpu_baseline(pairs, threads: int, flag: int = 0, w: int = -1):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O  # checker / baseline only

    if O.ref_available():
        kind = "reference"
        fn = lambda p: O.run_ref(p[0], p[1], mat=W_MAT, flag=flag, w=w, **SC)  # noqa: E731
    else:
        kind = "port"
        fn = lambda p: O.run_oracle(p[0], p[1], mat=W_MAT, flag=flag, w=w, **SC)  # noqa: E731
    # largest-cells-first, one pair per task
    order = sorted(range(len(pairs)), key=lambda k: -len(pairs[k][0]) * len(pairs[k][1]))
    res = [None] * len(pairs)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        for k, r in zip(order, ex.map(lambda k: fn(pairs[k]), order)):
            res[k] = r
    dt = time.perf_counter() - t0
    return kind, dt, res


def singletrack_leg(args, pairs, cells) -> dict:
    """The fixed batch with KSW_EZ_SINGLETRACK (two difference tracks per
    cell, the Singletrack walk): resident inputs, timed like the headline,
    checked against the CPU reference with the same flag."""
    plan = ksw2.Plan(pairs, mat=W_MAT, flag=ksw2.KSW_EZ_SINGLETRACK, **SC)
    plan.run()
    plan.run()  # set-up as for the headline: streams, then the graph capture
    dp = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run()
        dp.append(plan.timing()[0])
    wall = time.perf_counter() - t0
    res = plan.fetch()
    plan.close()
    gcups = cells / (wall / args.steps) / 1e9
    out = {"gcups": round(gcups, 2), "ms_per_step": round(wall * 1e3 / args.steps, 3),
           "device_ms": round(float(np.mean(dp)), 3),
           "track_bytes_per_cell": 2, "walk": "wave-parallel Singletrack walk in the stripe kernel's epilogue"}
    log(f"[bench] singletrack: {gcups:.1f} GCUPS ({out['device_ms']:.2f} ms device)")
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        kind, dt, cpu_res = cpu_baseline(pairs, threads, flag=ksw2.KSW_EZ_SINGLETRACK)
        exact = sum(1 for g, c in zip(res, cpu_res) if g.__dict__ == c.__dict__)
        out["exact"] = {"agree": exact, "of": len(pairs), "frac": exact / len(pairs)}
        out["cpu"] = {"kind": kind, "gcups": round(cells / dt / 1e9, 3), "cores": threads}
        log(f"[bench] singletrack CPU {cells / dt / 1e9:.2f} GCUPS; exact {exact}/{len(pairs)}")
    return out


def band_cells(qlen: int, tlen: int, w: int) -> int:
    """In-band cells the reference visits: the sum over anti-diagonals of
    en0 - st0 + 1 (minimap2/ksw2_extd2_sse.c:147-156), up to a band exit."""
    r = np.arange(qlen + tlen - 1, dtype=np.int64)
    st = np.maximum(np.maximum(0, r - qlen + 1), (r - w + 1) >> 1)
    en = np.minimum(np.minimum(tlen - 1, r), (r + w) >> 1)
    bad = np.nonzero(st > en)[0]
    stop = int(bad[0]) if len(bad) else len(r)
    return int((en[:stop] - st[:stop] + 1).sum())


def single_pair_leg(args) -> dict:
    """One 5 kbp pair (8% substitutions, 2% indels; global, full CIGAR): one
    GPU call (median of repeats) against the reference on one host core."""
    q, t = W.single(seed=1)[0]
    ksw2.extd2(q, t, mat=W_MAT, **SC)
    ts = []
    for _ in range(9):
        t1 = time.perf_counter()
        g = ksw2.extd2(q, t, mat=W_MAT, **SC)
        ts.append(time.perf_counter() - t1)
    cells = len(q) * len(t)
    out = {"qlen": len(q), "tlen": len(t), "gpu_ms": round(float(np.median(ts)) * 1e3, 3),
           "gpu_gcups": round(cells / float(np.median(ts)) / 1e9, 3),
           "timing": "wall clock of one ksw_extd2_hip call (upload, DP, backtrack, CIGAR), median of 9"}
    if not args.no_cpu:
        kind, dt, res = cpu_baseline([(q, t)], 1)
        out["cpu"] = {"kind": kind, "ms": round(dt * 1e3, 3), "gcups": round(cells / dt / 1e9, 3), "cores": 1}
        out["exact"] = g.__dict__ == res[0].__dict__
    log(f"[bench] single 5 kbp pair: {out}")
    return out


def threads_leg(args, pairs) -> dict:
    """AnchorWave's per-interval pattern: -t executor threads each calling the
    drop-in ksw_extd2_hip on their own pairs at once (fixed-batch pairs,
    2 kbp; 16 threads), against the reference ksw_extd2_avx512 on the same
    pairs and threads."""
    threads = 16
    sub = pairs[: args.threads_leg]
    cells = W.cells(sub)
    ksw2.extd2(*sub[0], mat=W_MAT, **SC)  # device initialised
    res = [None] * len(sub)

    def work(w):
        for k in range(w, len(sub), threads):
            res[k] = ksw2.extd2(*sub[k], mat=W_MAT, **SC)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(threads)))  # every thread's first calls set up its context
        t0 = time.perf_counter()
        list(ex.map(work, range(threads)))
        dt = time.perf_counter() - t0
    out = {"pairs": len(sub), "threads": threads, "cells": cells, "seconds": round(dt, 3),
           "gcups": round(cells / dt / 1e9, 2), "calls_per_s": round(len(sub) / dt, 1),
           "timing": "wall clock of all calls, 16 host threads each calling ksw_extd2_hip one pair at a time"}
    if not args.no_cpu:
        kind, cdt, cres = cpu_baseline(sub, threads)
        out["cpu"] = {"kind": kind, "gcups": round(cells / cdt / 1e9, 3), "cores": threads}
        exact = sum(1 for g, c in zip(res, cres) if g.__dict__ == c.__dict__)
        out["exact"] = {"agree": exact, "of": len(sub), "frac": exact / len(sub)}
    log(f"[bench] {threads} threads calling ksw_extd2_hip: {out}")
    return out


def banded_leg(args) -> dict:
    """64 long pairs (100-300 kbp; 10% substitutions, 2% indels, 2-5
    structural indels of 5-50 kbp), w = 4096, zdrop off, full traceback,
    through ksw_extd2_hip_batch; GCUPS over in-band cells.  Pairs whose
    length difference exceeds the band leave it (zdropped), as in the
    reference."""
    log(f"[bench] banded batch: generating {args.banded} pairs")
    pb = W.banded(seed=4, n=args.banded)
    w = 4096
    cells = sum(band_cells(len(q), len(t), w) for q, t in pb)
    # a first call, untimed: the device block it takes stays cached for the next
    ksw2.extd2_batch(pb, mat=W_MAT, w=w, **SC)
    tm = {}
    gres = ksw2.extd2_batch(pb, mat=W_MAT, w=w, **SC, timing=tm)
    dt = tm["seconds"]
    exits = sum(1 for g in gres if g.zdropped)
    log(f"[bench] banded batch: {cells:.3g} in-band cells in {dt:.2f} s = {cells / dt / 1e9:.1f} GCUPS "
        f"({exits} band exits)")
    res = {"pairs": len(pb), "w": w, "in_band_cells": cells, "seconds": round(dt, 3),
           "gcups": round(cells / dt / 1e9, 2), "band_exits": exits,
           "timing": "wall clock of a second ksw_extd2_hip_batch call: upload, DP, backtrack, CIGAR download"}
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        kind, cdt, cres = cpu_baseline(pb, threads, w=w)
        exact = sum(1 for g, c in zip(gres, cres) if g.__dict__ == c.__dict__)
        res["exact"] = {"agree": exact, "of": len(pb), "frac": exact / len(pb)}
        res["cpu"] = {"kind": kind, "gcups": round(cells / cdt / 1e9, 3), "cores": threads}
        log(f"[bench] banded CPU {cells / cdt / 1e9:.2f} GCUPS; exact {exact}/{len(pb)}")
    return res


def general_leg(args, pairs) -> dict:
    """The general kernel's workload in AnchorWave: narrow bands, as the
    score-certified path asks for (alignSlidingWindow.cpp:1297-1310 widens
    the band from twice the length difference).  The fixed batch's pairs at
    w = 256, below the band kernel's minimum width, through
    ksw_extd2_hip_batch; GCUPS over in-band cells up to each band exit."""
    w = args.general_w
    sub = pairs[: args.general]
    cells = sum(band_cells(len(q), len(t), w) for q, t in sub)
    # a first call, untimed: the device block it takes stays cached for the next
    ksw2.extd2_batch(sub, mat=W_MAT, w=w, **SC)
    tm = {}
    gres = ksw2.extd2_batch(sub, mat=W_MAT, w=w, **SC, timing=tm)
    dt = tm["seconds"]
    exits = sum(1 for g in gres if g.zdropped)
    res = {"pairs": len(sub), "w": w, "in_band_cells": cells, "seconds": round(dt, 4),
           "gcups": round(cells / dt / 1e9, 2), "band_exits": exits, "kernel": "ksw_general_kernel",
           "timing": "wall clock of a second ksw_extd2_hip_batch call: upload, DP, backtrack, CIGAR download"}
    log(f"[bench] general kernel (w={w}): {cells:.3g} in-band cells in {dt * 1e3:.1f} ms = "
        f"{cells / dt / 1e9:.1f} GCUPS ({exits} band exits)")
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        kind, cdt, cres = cpu_baseline(sub, threads, w=w)
        exact = sum(1 for g, c in zip(gres, cres) if g.__dict__ == c.__dict__)
        res["exact"] = {"agree": exact, "of": len(sub), "frac": exact / len(sub)}
        res["cpu"] = {"kind": kind, "gcups": round(cells / cdt / 1e9, 3), "cores": threads}
        log(f"[bench] general-kernel CPU {cells / cdt / 1e9:.2f} GCUPS; exact {exact}/{len(sub)}")
    return res


def uneven_leg(args) -> dict:
    """The uneven inter-anchor batch through ksw_extd2_hip_batch (host
    buffers in and out, chunked by the device memory budget), against the
    same kind of call made one pair at a time on a seeded sample, and checked
    bit-exact against the CPU reference on every pair."""
    log(f"[bench] uneven batch: generating {args.uneven} pairs")
    pu = W.uneven(seed=3, n=args.uneven)
    cu = W.cells(pu)
    ksw2.extd2(pu[0][0][:64], pu[0][1][:64], mat=W_MAT, **SC)  # device initialised
    tm = {}
    gres = ksw2.extd2_batch(pu, mat=W_MAT, **SC, timing=tm)
    dt = tm["seconds"]
    log(f"[bench] uneven batch: {cu:.3g} cells in {dt:.2f} s = {cu / dt / 1e9:.1f} GCUPS (host API)")
    rng = np.random.default_rng(7)
    idx = rng.choice(len(pu), size=min(args.single_sample or 24, len(pu)), replace=False)
    tot, cs = 0.0, 0
    for k in idx:
        q, t = pu[k]
        t1 = time.perf_counter()
        ksw2.extd2(q, t, mat=W_MAT, **SC)
        tot += time.perf_counter() - t1
        cs += len(q) * len(t)
    single = cs / tot / 1e9
    log(f"[bench] uneven one-at-a-time: {single:.2f} GCUPS over {len(idx)} sampled pairs")
    res = {"pairs": len(pu), "cells": cu, "seconds": round(dt, 3), "gcups": round(cu / dt / 1e9, 2),
           "one_at_a_time_gcups": round(single, 3), "one_at_a_time_sample": int(len(idx)),
           "batched_over_one_at_a_time": round(cu / dt / 1e9 / single, 1),
           "timing": "wall clock of ksw_extd2_hip_batch: upload, DP, backtrack, CIGAR download"}
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        log(f"[bench] uneven batch: CPU reference on {threads} threads (exactness check)")
        kind, cdt, cres = cpu_baseline(pu, threads)
        exact = sum(1 for g, c in zip(gres, cres) if g.__dict__ == c.__dict__)
        res["exact"] = {"agree": exact, "of": len(pu), "frac": exact / len(pu)}
        res["cpu"] = {"kind": kind, "gcups": round(cu / cdt / 1e9, 3), "cores": threads}
        log(f"[bench] uneven CPU {cu / cdt / 1e9:.2f} GCUPS; exact {exact}/{len(pu)}")
    return res


def serving_leg(args) -> dict:
    """Continuous batching under Poisson arrivals (ksw2_hip_server_*): task
    sizes drawn 50/40/10 from the single-pair, fixed-batch and uneven-batch
    distributions.  Capacity is measured with every task submitted at once;
    then each offered load (a fraction of capacity) replays the tasks with
    exponential inter-arrival gaps.  Latency is submission to completion,
    timed inside the library."""
    log(f"[bench] serving: generating {args.serve} tasks")
    tasks = W.serving(seed=5, n=args.serve)
    cells = np.array([len(t.query) * len(t.target) for t in tasks], dtype=np.float64)
    # admission budget for DP and traceback state: 200 GB of HBM (the block
    # the batch legs keep cached is released first)
    ksw2.release_cached_memory()
    srv = ksw2.Server(mat=W_MAT, device_bytes=SERVE_BUDGET, **SC)

    kinds = np.array([t.kind for t in tasks])

    def run(arrivals):
        n = len(tasks)
        lat = np.zeros(n)
        kind_of = {}
        t0 = time.perf_counter()
        i = done = 0
        while done < n:
            now = time.perf_counter() - t0
            while i < n and (arrivals is None or arrivals[i] <= now):
                kind_of[srv.submit(tasks[i].query, tasks[i].target)] = i
                i += 1
            if i >= n or arrivals is None or (i & 15) == 0:
                for tk, _, lt in srv.poll(max_n=4096, wait_ms=0 if i < n else 1, results=False):
                    lat[kind_of.pop(tk)] = lt
                    done += 1
            if i < n and arrivals is not None:
                gap = arrivals[i] - (time.perf_counter() - t0)
                if gap > 2e-4:
                    time.sleep(min(gap, 1e-3))
        return time.perf_counter() - t0, lat

    run(None)  # warm-up: device buffers, code objects
    wall, lat = run(None)
    cap = len(tasks) / wall
    log(f"[bench] serving batch capacity (every task submitted at once): {cap:.0f} tasks/s, "
        f"{cells.sum() / wall / 1e9:.1f} GCUPS")
    rng = np.random.default_rng(11)

    def level(rate, load):
        arr = np.cumsum(rng.exponential(1.0 / rate, len(tasks)))
        wall, lat = run(arr)
        lv = {"load": load, "offered_tasks_per_s": round(rate, 1),
              "achieved_tasks_per_s": round(len(tasks) / wall, 1),
              "gcups": round(float(cells.sum() / wall / 1e9), 2),
              "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
              "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
              "by_kind_p50_p99_ms": {k: [round(float(np.percentile(lat[kinds == k], 50)) * 1e3, 2),
                                         round(float(np.percentile(lat[kinds == k], 99)) * 1e3, 2)]
                                     for k in ("single", "fixed", "uneven")}}
        return lv

    # sustained capacity: the highest offered Poisson rate (fractions of the
    # batch capacity) that the server keeps up with, achieved >= 0.98 offered
    search = []
    for f in [float(x) for x in args.search.split(",") if x]:
        lv = level(f * cap, None)
        search.append({k: lv[k] for k in ("offered_tasks_per_s", "achieved_tasks_per_s", "p50_ms", "p99_ms")})
        log(f"[bench] serving capacity search {f:.2f} of batch capacity: {search[-1]}")
        if lv["achieved_tasks_per_s"] < 0.98 * lv["offered_tasks_per_s"]:
            break
    kept = [lv for lv in search if lv["achieved_tasks_per_s"] >= 0.98 * lv["offered_tasks_per_s"]]
    sustained = max(lv["offered_tasks_per_s"] for lv in kept) if kept else search[0]["achieved_tasks_per_s"]
    log(f"[bench] serving sustained capacity: {sustained:.0f} tasks/s ({sustained / cap:.2f} of batch capacity)")
    # the sweep: fractions of the sustained capacity
    levels = []
    for load in [float(x) for x in args.loads.split(",") if x]:
        levels.append(level(load * sustained, load))
        log(f"[bench] serving load {load:.2f}: {levels[-1]}")
    ld = srv.load()
    srv.close()
    # the same tasks one call at a time (seeded sample per kind)
    single = {}
    for kind in ("single", "fixed", "uneven"):
        idx = [k for k, t in enumerate(tasks) if t.kind == kind][:12]
        ts = []
        for k in idx:
            t1 = time.perf_counter()
            ksw2.extd2(tasks[k].query, tasks[k].target, mat=W_MAT, **SC)
            ts.append(time.perf_counter() - t1)
        single[kind] = {"p50_ms": round(float(np.median(ts)) * 1e3, 3),
                        "gcups": round(float(cells[idx].sum() / sum(ts) / 1e9), 3), "sample": len(idx)}
    log(f"[bench] serving one-at-a-time: {single}")
    return {"tasks_per_level": len(tasks), "mix": "50% 5 kbp pair, 40% 2 kbp pair, 10% uneven-batch pair",
            "batch_capacity_tasks_per_s": round(cap, 1),
            "batch_capacity_gcups": round(cells.sum() * cap / len(tasks) / 1e9, 2),
            "capacity_search": search,
            "sustained_tasks_per_s": round(sustained, 1),
            "sustained": "highest offered Poisson rate (fractions of the batch capacity, every task submitted "
                         "at once) with achieved >= 0.98 offered; the sweep's loads are fractions of it",
            "admission_budget_bytes": SERVE_BUDGET,
            "sequences": {"where": "device arena, uploaded at submission", "arena_bytes": ld["seq_arena_bytes"],
                          "resident_tasks": ld["resident_tasks"], "staged_flights": ld["staged_flights"]},
            "levels": levels, "one_at_a_time": single,
            "latency": "submission to completion inside the library (queueing + device + CIGAR delivery)"}


def copy_calibration(nbytes: int) -> None:
    """One hipMemcpy device-to-device of nbytes: a plain copy kernel whose
    FETCH_SIZE / WRITE_SIZE readings scale the counters to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy calibration: {nbytes} bytes device-to-device")


SC = dict(m=5, q=8, e=2, q2=75, e2=1)
SERVE_BUDGET = 200 * 10**9
W_MAT = ksw2.default_matrix(W.MATCH, W.MISMATCH)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=4096, help="pairs in the fixed batch")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / cross-check")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all host cores (max 16)")
    ap.add_argument("--single-sample", type=int, default=16, help="pairs timed one at a time")
    ap.add_argument("--uneven", type=int, default=20000,
                    help="pairs in the uneven-batch leg (0: skip it)")
    ap.add_argument("--serve", type=int, default=12000,
                    help="tasks per offered-load level in the serving leg (0: skip it)")
    ap.add_argument("--search", default="0.15,0.2,0.25,0.3,0.35,0.4,0.5,0.6,0.8",
                    help="offered rates tried for the serving leg's sustained capacity, as fractions of "
                         "its batch capacity (stops at the first one not kept up with)")
    ap.add_argument("--loads", default="0.2,0.4,0.6,0.8,1.0,1.2",
                    help="offered loads of the serving leg, as fractions of its sustained capacity")
    ap.add_argument("--single-pair", type=int, default=1,
                    help="also time one 5 kbp pair on the GPU and on one host core (0: skip)")
    ap.add_argument("--banded", type=int, default=64,
                    help="pairs in the banded long-pair leg, w = 4096 (0: skip it)")
    ap.add_argument("--singletrack", type=int, default=1,
                    help="also time the fixed batch with KSW_EZ_SINGLETRACK (0: skip)")
    ap.add_argument("--threads-leg", type=int, default=1024,
                    help="pairs of the 16-thread ksw_extd2_hip leg (0: skip it)")
    ap.add_argument("--general", type=int, default=4096,
                    help="fixed-batch pairs of the narrow-band (general kernel) leg (0: skip it)")
    ap.add_argument("--general-w", type=int, default=256, help="band of the general-kernel leg")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (PMC byte-counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: every pair is independent, so "
                          "more GPUs would each take a share of the pairs; only --gpus 1 is measured"}))
        return 2

    if args.copy_calib > 0:
        copy_calibration(args.copy_calib << 20)
    log(f"[bench] generating fixed batch ({args.pairs} pairs x 2000 bp)")
    pairs = W.fixed(seed=2, n=args.pairs)
    cells = W.cells(pairs)
    log("[bench] uploading plan")
    plan = ksw2.Plan(pairs, mat=W_MAT, **SC)
    # plan set-up: the first run creates the plan's lane streams and the
    # second captures its launches as a hipGraph (replayed from then on)
    plan.run()
    plan.run()
    for _ in range(args.warmup):
        plan.run()
    log(f"[bench] warmup done; timing {args.steps} steps")
    dp_ms, bt_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run()  # synchronous: DP + backtrack complete on return
        d, b, _c = plan.timing()
        dp_ms.append(d)
        bt_ms.append(b)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    gcups = cells / (ms_per_step * 1e-3) / 1e9
    dp_avg = float(np.mean(dp_ms))
    log(f"[bench] {ms_per_step:.2f} ms/step, {gcups:.1f} GCUPS (DP kernel {dp_avg:.2f} ms, "
        f"backtrack {np.mean(bt_ms):.2f} ms)")
    t1 = time.perf_counter()
    ftm = {}
    gpu_res = plan.fetch(timing=ftm)
    fetch_py_ms = (time.perf_counter() - t1) * 1e3
    fetch_ms = ftm["seconds"] * 1e3
    log(f"[bench] fetch of every score and CIGAR to host ez structs (ksw2_hip_plan_fetch): {fetch_ms:.2f} ms; "
        f"with the Python binding's result objects {fetch_py_ms:.2f} ms")
    dir_bytes = sum(((len(t) + 511) // 512 + 1) // 2 * (len(q) + 127) * 1024 for q, t in pairs)
    plan.close()

    # one pair per launch (inputs resident), the batching baseline
    single = None
    if args.single_sample > 0:
        sp = pairs[: args.single_sample]
        tot, cs = 0.0, 0
        for q, t in sp:
            p1 = ksw2.Plan([(q, t)], mat=W_MAT, **SC)
            p1.run()
            t1 = time.perf_counter()
            p1.run()
            tot += time.perf_counter() - t1
            cs += len(q) * len(t)
            p1.close()
        single = cs / tot / 1e9
        log(f"[bench] one-at-a-time: {single:.2f} GCUPS over {len(sp)} pairs")

    achieved = cells * OPS_PER_CELL / (dp_avg * 1e-3) / 1e12
    roofline = {
        "bound": "valu",
        "achieved": round(achieved, 3),
        "peak": round(PEAK_PACKED16_TOPS, 1),
        "unit": "Tops/s",
        "frac": round(achieved / PEAK_PACKED16_TOPS, 4),
        "traffic": read_profile_value("stripe_kernel_fixed_4096x2000_pmc.md", "traffic_bytes_per_dispatch"),
        "kernel": "ksw_stripe_kernel",
        "ops_per_cell": OPS_PER_CELL,
        "avg_launch_ms": round(dp_avg, 3),
        # measured VALU issue: busy fraction of the SIMDs' VALU issue slots (PMC)
        "valu_issue_frac": read_profile_value("stripe_kernel_fixed_4096x2000_pmc.md", "valu_issue_frac"),
        "traceback_write_GBs": round(dir_bytes / (dp_avg * 1e-3) / 1e9, 1),
        "traceback_write_frac_hbm": round(dir_bytes / (dp_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    out = {
        "metric": METRIC,
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16 (packed pairs of cells; exact to the reference's int8)",
        "data": "synthetic",
        "config": {"workload": f"fixed batch: {args.pairs} pairs x 2000 bp, 5% subs, 1% indels, "
                               "10% with a 50-300 bp indel; global, full CIGAR",
                   "pairs": args.pairs, "cells": cells, "scoring": "0/-6, 8/2, 75/1",
                   "device_ms_dp": round(dp_avg, 3), "device_ms_backtrack": round(float(np.mean(bt_ms)), 3),
                   "timed": "plan runs on resident inputs; scores and CIGARs left in HBM",
                   "fetch_ms": round(fetch_ms, 3),
                   "fetch": "ksw2_hip_plan_fetch of every score and CIGAR into ksw_extz_t structs, after the "
                            "timed steps; with the Python binding's result objects "
                            f"{fetch_py_ms:.2f} ms",
                   "gcups_with_fetch": round(cells / ((ms_per_step + fetch_ms) * 1e-3) / 1e9, 2),
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
        "roofline": roofline,
        "one_at_a_time_gcups": None if single is None else round(single, 3),
        "batched_over_one_at_a_time": None if not single else round(gcups / single, 1),
    }
    if args.single_pair:
        out["single_pair"] = single_pair_leg(args)
    if args.singletrack:
        out["singletrack"] = singletrack_leg(args, pairs, cells)
    if args.threads_leg > 0:
        out["threads"] = threads_leg(args, pairs)
    if args.general > 0:
        out["general"] = general_leg(args, pairs)
    if args.banded > 0:
        out["banded"] = banded_leg(args)
    if args.uneven > 0:
        out["uneven"] = uneven_leg(args)
    if args.serve > 0:
        out["serving"] = serving_leg(args)
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        log(f"[bench] CPU baseline on {threads} threads")
        kind, dt, cpu_res = cpu_baseline(pairs, threads)
        exact = sum(1 for g, c in zip(gpu_res, cpu_res) if g.__dict__ == c.__dict__)
        out["cpu_baseline"] = {"value": round(cells / dt / 1e9, 3), "unit": "GCUPS", "cores": threads,
                               "kind": kind, **host_cpu(),
                               "sample": f"the full fixed batch ({len(pairs)} pairs, {cells:.3g} cells), "
                                         "ksw_extd2_avx512 one pair per task, largest first"}
        out["exact"] = {"agree": exact, "of": len(pairs), "frac": exact / len(pairs)}
        log(f"[bench] CPU {cells / dt / 1e9:.2f} GCUPS; exact {exact}/{len(pairs)}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
