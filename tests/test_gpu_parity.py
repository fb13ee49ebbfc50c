"""HIP ksw_extd2 vs the oracle (bit-exact: every ez field and the CIGAR)."""
import numpy as np
import pytest

import oracle_py as O
from anchorwave_amd import ksw2, workloads as W
from helpers import diff, random_pairs, same

pytestmark = pytest.mark.gpu

FLAGS = [0, ksw2.KSW_EZ_RIGHT, ksw2.KSW_EZ_SCORE_ONLY, ksw2.KSW_EZ_REV_CIGAR,
         ksw2.KSW_EZ_SEMIGLOBAL_END, ksw2.KSW_EZ_EXTZ_ONLY,
         ksw2.KSW_EZ_SEMIGLOBAL_END | ksw2.KSW_EZ_RIGHT]


def check_batch(pairs, flag=0, **kw):
    got = ksw2.extd2_batch(pairs, flag=flag, **kw)
    bad = []
    for k, ((q, t), g) in enumerate(zip(pairs, got)):
        ref = O.run_oracle(q, t, flag=flag, **kw)
        if not same(g, ref):
            bad.append(f"pair {k} qlen={len(q)} tlen={len(t)} flag={flag:#x}: {diff(g, ref)}")
    assert not bad, "\n".join(bad[:5]) + f"\n({len(bad)} of {len(pairs)} differ)"


@pytest.mark.parametrize("flag", FLAGS)
def test_random_small(flag):
    check_batch(random_pairs(100 + flag, 60, 1, 700), flag=flag)


def test_edge_lengths():
    rng = np.random.default_rng(5)
    pairs = []
    for ql, tl in [(1, 1), (1, 7), (7, 1), (1, 513), (513, 1), (63, 64), (64, 63), (511, 512),
                   (512, 513), (513, 512), (1024, 1), (2, 1025)]:
        pairs.append((W.random_seq(rng, ql), W.random_seq(rng, tl)))
    check_batch(pairs)


def test_identical_and_gap_only():
    rng = np.random.default_rng(6)
    t = W.random_seq(rng, 900)
    pairs = [(t.copy(), t), (t[:300].copy(), t), (t, t[:250].copy()),
             (np.concatenate([t[:400], t[700:]]), t)]
    check_batch(pairs)


def test_single_config_pair():
    check_batch(W.single(seed=1))


def test_fixed_config_sample():
    check_batch(W.fixed(seed=2, n=48))
