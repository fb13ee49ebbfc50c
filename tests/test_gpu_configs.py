"""The workloads at their real shapes, bit-exact against the reference.

- the banded long pairs (100-300 kbp, 2-5 structural indels of 5-50 kbp,
  w = 4096): pairs that stay in the band and pairs that leave it;
- the tail of the uneven inter-anchor batch (pairs above 1.5e8 cells), on
  the team sizes the planner picks for them (wide teams of several
  workgroups, no override);
- a Poisson replay of the serving mix whose uneven tasks are drawn like
  the uneven batch (length ratio 0.5-2.0);
- the admission footprint model against the device memory a plan takes.

Every ksw_extz_t field and the CIGAR are compared with the compiled
reference ksw_extd2_avx512 (oracle/_ref) when it is present, else with the
oracle restatement."""
import ctypes as C
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, same

pytestmark = pytest.mark.gpu


def _cpu(p, threads=8, **kw):
    fn = O.run_ref if O.ref_available() else O.run_oracle

    def one(x):
        q, t = x[0], x[1]
        return fn(q, t, **kw)

    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, p))


def _compare(pairs, got, exp, what):
    bad = [f"{what} pair {k} qlen={len(q)} tlen={len(t)}: {diff(g, r)}"
           for k, ((q, t), g, r) in enumerate(zip(pairs, got, exp)) if not same(g, r)]
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(pairs)} differ)"


def test_banded_long_pairs_at_scale():
    """Five pairs of the banded long-pair workload (seed 4) under w = 4096:
    pairs 5 and 29 stay in the band (|qlen - tlen| <= w), pairs 0, 3 and 4
    leave it (zdropped, traceback from the maximum), as the reference."""
    allp = W.banded(seed=4)
    idx = [0, 3, 4, 5, 29]
    pairs = [allp[k] for k in idx]
    assert min(len(t) for _, t in pairs) >= 100_000
    print(f"banded pairs {idx}: " + ", ".join(f"{len(q)}x{len(t)}" for q, t in pairs), flush=True)
    got = ksw2.extd2_batch(pairs, w=W.BANDED_W)
    exp = _cpu(pairs, threads=5, w=W.BANDED_W)
    _compare(pairs, got, exp, "banded")
    exits = [k for k, r in zip(idx, exp) if r.zdropped]
    assert exits == [0, 3, 4], exits
    # the pairs that stay in the band have a full path
    assert all(r.n_cigar > 0 for k, r in zip(idx, exp) if k not in exits)


def test_uneven_tail_on_planner_teams():
    """The largest pairs of the uneven batch's first 4,000 (1.5e8 to 3e9
    cells, up to ~50 kbp targets), in one batch with the planner's own team
    sizes: at this batch size the 50 kbp target's team exceeds 16 waves (a
    wide team over several workgroups) and the others, of 5-16 waves, are
    spread over workgroups on different CUs."""
    un = W.uneven(seed=3, n=4000)
    cells = np.array([len(q) * len(t) for q, t in un], dtype=np.int64)
    cand = [k for k in np.argsort(-cells) if 1.5e8 < cells[k] <= 3e9]
    pick = cand[::max(1, len(cand) // 6)][:6]
    pairs = [un[k] for k in pick]
    print("uneven tail: " + ", ".join(f"{len(q)}x{len(t)}" for q, t in pairs), flush=True)
    assert len(pairs) == 6 and max(len(t) for _, t in pairs) > 16 * 1024
    got = ksw2.extd2_batch(pairs)
    exp = _cpu(pairs, threads=3)
    _compare(pairs, got, exp, "uneven tail")


def test_serving_poisson_replay_with_uneven_tasks():
    """The serving mix (50% 5 kbp, 40% 2 kbp, 10% uneven-batch pairs with the
    0.5-2.0 length ratio) submitted at Poisson arrival times while earlier
    tasks run: every result equals the reference's and the pair's lone
    ksw_extd2_hip result."""
    tasks = [t for t in W.serving(seed=21, n=200, rate_per_s=2000.0)
             if len(t.query) * len(t.target) <= 4e8]
    assert sum(t.kind == "uneven" for t in tasks) >= 10
    srv = ksw2.Server()
    tickets, got = {}, {}
    t0 = time.perf_counter()
    for k, task in enumerate(tasks):
        while time.perf_counter() - t0 < task.arrival_s:
            for tk, res, _ in srv.poll(wait_ms=0):
                got[tickets[tk]] = res
            time.sleep(0.0002)
        tickets[srv.submit(task.query, task.target)] = k
    while len(got) < len(tasks):
        for tk, res, _ in srv.poll(wait_ms=2000):
            got[tickets[tk]] = res
    srv.close()
    pairs = [(t.query, t.target) for t in tasks]
    exp = _cpu(pairs, threads=8)
    _compare(pairs, [got[k] for k in range(len(tasks))], exp, "served")
    alone_idx = [k for k, t in enumerate(tasks) if t.kind == "uneven"][:6] + [0, 1]
    for k in alone_idx:
        alone = ksw2.extd2(*pairs[k])
        assert same(got[k], alone), f"task {k}: mix vs alone: {diff(got[k], alone)}"


def _hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def _free_bytes(hip):
    fr, tot = C.c_size_t(), C.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(C.byref(fr), C.byref(tot)) == 0
    return fr.value


@pytest.mark.parametrize("kind", ["stripe", "band", "general"])
def test_footprint_bounds_device_memory(kind):
    """ksw2_hip_footprint is what the batch and serving paths admit by: the
    device memory a plan actually takes (hipMemGetInfo before creation and
    after a run) stays within the summed footprints plus a small constant,
    and the footprints do not overstate it by more than a quarter, for each
    of the three kernels."""
    rng = np.random.default_rng(33)
    if kind == "stripe":
        pairs = W.fixed(seed=7, n=48)
        w = -1
    elif kind == "band":
        ts = [W.random_seq(rng, 40000) for _ in range(4)]
        pairs = [(W.mutate(rng, t, 0.05, 0.01), t) for t in ts]
        w = 2048
    else:
        ts = [W.random_seq(rng, 12000) for _ in range(8)]
        pairs = [(W.mutate(rng, t, 0.05, 0.01), t) for t in ts]
        w = 300
    routes = {ksw2.route(q, t, w=w) for q, t in pairs}
    assert routes == ({0} if kind == "stripe" else {1}), routes
    fp = sum(ksw2.footprint(q, t, w=w) for q, t in pairs)
    hip = _hip()
    warm = ksw2.Plan(pairs, w=w)  # one-time runtime allocations (code objects, queues)
    warm.run()
    warm.close()
    f0 = _free_bytes(hip)
    plan = ksw2.Plan(pairs, w=w)
    plan.run()
    f1 = _free_bytes(hip)
    res = plan.fetch()
    plan.close()
    used = f0 - f1
    print(f"{kind}: footprint {fp / 2**20:.1f} MiB, device memory taken {used / 2**20:.1f} MiB", flush=True)
    assert all(r.n_cigar > 0 for r in res)
    assert used <= fp + (16 << 20), (used, fp)
    assert used >= 0.75 * fp - (4 << 20), (used, fp)
