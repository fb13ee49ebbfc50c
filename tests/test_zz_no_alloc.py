"""ksw_extd2_hip_noalloc against ksw_extd2_hip: the same pairs through both
entry points give identical ksw_extz_t fields and CIGARs, on two calls in a
row (the second smaller, so that the scratch still holds the first call's
bytes), on each kernel the library routes to; and the calls take no device
memory."""
import ctypes as C

import numpy as np
import pytest

from anchorwave_amd import ksw2, workloads as W
from helpers import diff, same

pytestmark = pytest.mark.gpu


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    fr, tot = C.c_size_t(), C.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(C.byref(fr), C.byref(tot)) == 0
    return fr.value


def _both(q, t, **kw):
    st, got = ksw2.extd2_noalloc(q, t, **kw)
    assert st == 0, ksw2.STATUS.get(st, st)
    ref = ksw2.extd2(q, t, **kw)
    assert same(got, ref), f"qlen={len(q)} tlen={len(t)} {kw}: {diff(got, ref)}"
    return got


def _pair(rng, tlen, qlen=None):
    t = W.random_seq(rng, tlen)
    q = W.mutate(rng, t, 0.08, 0.02)
    if qlen is not None:
        q = q[:qlen] if len(q) >= qlen else np.concatenate([q, W.random_seq(rng, qlen - len(q))])
    return q, t


def test_set_up_then_two_calls_second_smaller():
    """The single-pair workload (5 kbp target, 8% substitutions, 2% indels)
    then a 700 bp pair: both equal ksw_extd2_hip's results."""
    assert ksw2.lib().ksw2_hip_init(0) == 0
    rng = np.random.default_rng(5)
    big = W.single(seed=1)[0]
    small = _pair(rng, 700)
    print(f"noalloc: {len(big[0])}x{len(big[1])} then {len(small[0])}x{len(small[1])}", flush=True)
    r1 = _both(*big)
    assert r1.n_cigar > 0
    r2 = _both(*small)
    assert r2.n_cigar > 0


@pytest.mark.parametrize("kind", ["stripe", "stripe_right", "score_only", "singletrack", "general", "band",
                                  "zdrop_semiglobal"])
def test_each_kernel_two_calls(kind):
    rng = np.random.default_rng(len(kind))
    kw = {}
    if kind == "stripe_right":
        kw = dict(flag=ksw2.KSW_EZ_RIGHT)
    elif kind == "score_only":
        kw = dict(flag=ksw2.KSW_EZ_SCORE_ONLY)
    elif kind == "singletrack":
        kw = dict(flag=ksw2.KSW_EZ_SINGLETRACK)
    elif kind == "general":
        kw = dict(w=100)
    elif kind == "band":
        kw = dict(w=600)
    elif kind == "zdrop_semiglobal":
        kw = dict(zdrop=100, flag=ksw2.KSW_EZ_SEMIGLOBAL_END)
    sizes = [(3000, 3100), (1200, 900)] if kind != "band" else [(6000, 6500), (2500, 2400)]
    for tlen, qlen in sizes:
        q, t = _pair(rng, tlen, qlen)
        print(f"noalloc {kind}: {len(q)}x{len(t)} route {ksw2.route(q, t, **{k: v for k, v in kw.items() if k in ('w', 'zdrop', 'flag')})}",
              flush=True)
        _both(q, t, **kw)


def test_takes_no_device_memory():
    """After the set-up, calls leave the device's free memory unchanged."""
    rng = np.random.default_rng(9)
    q, t = _pair(rng, 4000)
    _both(q, t)  # set-up done, plan code paths warm
    f0 = _free_bytes()
    for _ in range(3):
        st, r = ksw2.extd2_noalloc(q, t)
        assert st == 0 and r.n_cigar > 0
    f1 = _free_bytes()
    assert f1 == f0, (f0, f1)


def test_rejects_more_than_the_scratch():
    """A pair whose direction bytes exceed the scratch is rejected before any
    device work; the next call still runs."""
    rng = np.random.default_rng(11)
    q, t = _pair(rng, 12000)
    st, r = ksw2.extd2_noalloc(q, t)
    assert st == ksw2.ETOOBIG and r.zdropped == 1 and r.n_cigar == 0
    q2, t2 = _pair(rng, 1500)
    _both(q2, t2)
