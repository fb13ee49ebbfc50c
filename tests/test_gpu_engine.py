"""Continuous batching: the serving API and batches larger than the device
memory budget.  Every result must be exactly what the same pair gives alone
(ksw_extd2_hip) and what the oracle gives."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu


def _mix(seed, n):
    """Pairs of mixed size, flags, bands and wildcard content."""
    rng = np.random.default_rng(seed)
    out = []
    for k, (q, t) in enumerate(random_pairs(seed, n, 1, 3000)):
        flag = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_EXTZ_ONLY][k % 5]
        w = -1 if k % 3 else int(rng.integers(20, 400))
        zdrop = -1 if k % 4 else int(rng.integers(30, 200))
        if k % 7 == 0:
            t = t.copy()
            t[rng.integers(0, len(t), max(1, len(t) // 40))] = 4
        out.append((q, t, flag, w, zdrop, int(rng.integers(0, 20))))
    return out


def _oracle(p):
    q, t, flag, w, zdrop, eb = p
    return O.run_oracle(q, t, flag=flag, w=w, zdrop=zdrop, end_bonus=eb)


def test_server_results_alone_and_in_the_mix():
    pairs = _mix(81, 60) + [(q, t, 0, -1, -1, 0) for q, t in W.fixed(seed=2, n=8)] + \
        [(q, t, 0, -1, -1, 0) for q, t in W.single(seed=1)]
    srv = ksw2.Server()
    tickets = {}
    for k, p in enumerate(pairs):
        tickets[srv.submit(p[0], p[1], w=p[3], zdrop=p[4], end_bonus=p[5], flag=p[2])] = k
    got = {}
    while len(got) < len(pairs):
        for tk, res, lat in srv.poll(wait_ms=2000):
            got[tickets[tk]] = res
            assert lat >= 0
    ld = srv.load()
    srv.close()
    # every pair's sequences went to the device arena at submission; a flight
    # uploads its own only when an upload it needs has not landed yet
    assert ld["seq_arena_bytes"] > 0 and ld["resident_tasks"] == len(pairs), ld
    assert ld["staged_flights"] <= len(pairs) // 4, ld
    for k, p in enumerate(pairs):
        alone = ksw2.extd2(p[0], p[1], flag=p[2], w=p[3], zdrop=p[4], end_bonus=p[5])
        exp = _oracle(p)
        assert same(got[k], alone), f"pair {k}: mix vs alone: {diff(got[k], alone)}"
        assert same(got[k], exp), f"pair {k}: vs oracle: {diff(got[k], exp)}"


def test_server_submissions_while_running():
    """Pairs submitted while earlier ones are on the device are admitted as
    they arrive and complete independently."""
    srv = ksw2.Server()
    batches = [random_pairs(90 + b, 25, 200, 2500) for b in range(4)]
    expect, got = {}, {}
    for b, pairs in enumerate(batches):
        for q, t in pairs:
            expect[srv.submit(q, t)] = (q, t)
        for tk, res, _ in srv.poll(wait_ms=0):
            got[tk] = res
    while len(got) < len(expect):
        for tk, res, _ in srv.poll(wait_ms=2000):
            got[tk] = res
    srv.close()
    for tk, (q, t) in expect.items():
        assert same(got[tk], O.run_oracle(q, t))


def test_server_budget_too_small_for_a_pair():
    """A pair whose footprint exceeds the server's whole budget completes at
    once with zdropped = 1 (the AVX kernels' allocation-failure convention);
    the others are unaffected."""
    srv = ksw2.Server(device_bytes=8 << 20)
    rng = np.random.default_rng(3)
    big_t = W.random_seq(rng, 6000)
    small = random_pairs(93, 5, 100, 400)
    tk_big = srv.submit(W.mutate(rng, big_t, 0.05, 0.01), big_t)
    tks = {srv.submit(q, t): (q, t) for q, t in small}
    got = {}
    while len(got) < 6:
        for tk, res, _ in srv.poll(wait_ms=2000):
            got[tk] = res
    assert srv.status[tk_big] == -3 and all(srv.status[tk] == 0 for tk in tks)  # ENOMEM for that ticket only
    srv.close()
    assert got[tk_big].zdropped == 1 and got[tk_big].n_cigar == 0
    for tk, (q, t) in tks.items():
        assert same(got[tk], O.run_oracle(q, t))


def test_batch_larger_than_budget_streams_through_the_engine(monkeypatch):
    """ksw_extd2_hip_batch over a batch several times the memory budget:
    admitted in flights by footprint, longest first, with backfill."""
    monkeypatch.setenv("KSW2_HIP_MEM_BUDGET", str(48 << 20))
    pairs = _mix(95, 80) + [(q, t, 0, -1, -1, 0) for q, t in W.fixed(seed=2, n=12)]
    got = ksw2.extd2_batch(pairs)
    for k, (p, g) in enumerate(zip(pairs, got)):
        exp = _oracle(p)
        assert same(g, exp), f"pair {k}: {diff(g, exp)}"


def test_engine_block_kept_between_batch_calls(monkeypatch):
    """Batches larger than the budget, one after another: the second takes
    the device block the first left cached, a server then takes it, and
    after release the next batch allocates again; every result exact."""
    monkeypatch.setenv("KSW2_HIP_MEM_BUDGET", str(40 << 20))
    pairs = _mix(97, 60) + [(q, t, 0, -1, -1, 0) for q, t in W.fixed(seed=5, n=10)]
    exp = [_oracle(p) for p in pairs]

    def check(got):
        for k, (g, e) in enumerate(zip(got, exp)):
            assert same(g, e), f"pair {k}: {diff(g, e)}"

    check(ksw2.extd2_batch(pairs))
    check(ksw2.extd2_batch(pairs))
    srv = ksw2.Server()
    tks = {srv.submit(q, t, w=w, zdrop=z, end_bonus=eb, flag=f): k for k, (q, t, f, w, z, eb) in enumerate(pairs[:20])}
    got = {}
    while len(got) < len(tks):
        for tk, res, _ in srv.poll(wait_ms=2000):
            got[tks[tk]] = res
    srv.close()
    for k, g in got.items():
        assert same(g, exp[k]), f"server pair {k}: {diff(g, exp[k])}"
    ksw2.release_cached_memory()
    check(ksw2.extd2_batch(pairs))
    ksw2.release_cached_memory()


def test_try_submit_defers_when_memory_is_committed():
    """A server with room for a few 2 kbp pairs: try_submit defers the rest
    (retry 250 ms to 5 s) instead of queueing them; held + queued footprint
    never exceeds the budget; every deferred pair, resubmitted, completes
    exactly as it does alone."""
    pairs = W.fixed(seed=12, n=40)
    fp = max(ksw2.footprint(q, t) for q, t in pairs)
    budget = 6 * fp + (1 << 20)
    srv = ksw2.Server(device_bytes=budget)
    waiting = list(range(len(pairs)))
    tickets, deferred, retries, done = {}, 0, [], {}
    for _round in range(10_000):
        still = []
        for k in waiting:
            tk, retry = srv.try_submit(*pairs[k])
            if tk is None:
                deferred += 1
                retries.append(retry)
                still.append(k)
            else:
                tickets[tk] = k
        waiting = still
        ld = srv.load()
        assert ld["held_bytes"] + ld["queued_bytes"] <= ld["budget_bytes"], ld
        for tk, res, _lat in srv.poll(wait_ms=20):
            done[tickets[tk]] = res
        if not waiting and len(done) == len(pairs):
            break
    srv.close()
    assert len(done) == len(pairs)
    assert deferred > 0 and all(250 <= r <= 5000 for r in retries), (deferred, retries[:5])
    for k, (q, t) in enumerate(pairs):
        alone = ksw2.extd2(q, t)
        assert same(done[k], alone), f"pair {k}: {diff(done[k], alone)}"


def test_try_submit_pair_above_the_budget_fails_like_submit():
    q, t = W.single(seed=1)[0]
    srv = ksw2.Server(device_bytes=ksw2.footprint(q, t) // 2)
    tk, retry = srv.try_submit(q, t)
    assert tk is not None and retry == 0
    (got,) = [r for _t, r, _l in srv.poll(wait_ms=-1)]
    srv.close()
    assert got.zdropped == 1


def test_failed_flight_is_reported_per_ticket(monkeypatch):
    """A flight that fails before launch (forced from the host for the first
    flight) completes its pairs with zdropped = 1 and no CIGAR, and poll
    reports the failure: per ticket (status EDEVICE, the others OK) and as
    the server's first failure; later flights are exact."""
    monkeypatch.setenv("KSW2_HIP_FAIL_FLIGHTS", "1")
    pairs = random_pairs(94, 30, 200, 1500)
    srv = ksw2.Server()
    tks = {srv.submit(q, t): k for k, (q, t) in enumerate(pairs)}
    got = {}
    while len(got) < len(pairs):
        for tk, res, _ in srv.poll(wait_ms=2000):
            got[tks[tk]] = (res, srv.status[tk])
    assert srv.last_error == -4
    srv.close()
    failed = [k for k, (_r, st) in got.items() if st != 0]
    assert failed and all(got[k][1] == -4 for k in failed)
    for k in failed:
        assert got[k][0].zdropped == 1 and got[k][0].n_cigar == 0
    for k, (res, st) in got.items():
        if st == 0:
            assert same(res, O.run_oracle(*pairs[k])), k


def test_server_without_sequence_arena_stages_per_flight():
    """A server too small for a sequence arena (under 16 MiB) uploads each
    flight's sequences itself, with the same results."""
    pairs = random_pairs(95, 12, 50, 400)
    srv = ksw2.Server(device_bytes=12 << 20)
    tickets = {srv.submit(q, t): k for k, (q, t) in enumerate(pairs)}
    got = {}
    while len(got) < len(pairs):
        for tk, res, _ in srv.poll(wait_ms=2000):
            got[tickets[tk]] = res
    ld = srv.load()
    srv.close()
    assert ld["seq_arena_bytes"] == 0 and ld["resident_tasks"] == 0 and ld["staged_flights"] > 0, ld
    for k, (q, t) in enumerate(pairs):
        assert same(got[k], O.run_oracle(q, t)), k
