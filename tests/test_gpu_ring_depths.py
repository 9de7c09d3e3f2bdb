"""Penalty sets whose gap components keep rings of different depths on the
device (M: the score scope; I1/D1: e1/g + 1 slots; I2/D2: e2/g + 1), on short
pairs (LDS rings) and long pairs (HBM rings), bit-exact with the oracle:
status, score, every CIGAR operation and the diagonal-step count.
"""
import sys

import pytest

import oracle as O
import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu

SETS = [
    ("affine", (0, 5, 7, 3)),           # g = 1: I1/D1 rings of 4
    ("affine", (0, 6, 8, 4)),           # g = 2: rings of 3
    ("affine2p", (0, 6, 4, 2, 12, 3)),  # g = 1: I1/D1 of 3, I2/D2 of 4
    ("affine2p", (0, 8, 4, 4, 16, 4)),  # g = 4: every gap ring of 2
    ("affine2p", (0, 3, 2, 2, 9, 1)),   # I2/D2 read one step back
]
IDS = ["%s%s" % (m, p) for m, p in SETS]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check(metric, pen, ps, memory="high", score_only=False):
    al = wfa_amd.Aligner(metric=metric, penalties=pen, memory=memory, score_only=score_only)
    db = al.upload(ps)
    assert db.run() in (0, 1)
    got = db.download()
    ref = O.oracle_align(O.Conf(metric, pen, memory, score_only), O.Pairs(ps.patterns(), ps.texts()))
    bad = [i for i in range(ps.n)
           if got.status[i] != ref.status[i] or got.score[i] != ref.score[i] or
           (not score_only and got.ops(i) != ref.ops(i))]
    assert not bad, "pairs differ: %s (first: status %d/%d score %d/%d)" % (
        bad[:10], got.status[bad[0]], ref.status[bad[0]], got.score[bad[0]], ref.score[bad[0]])
    if not score_only:
        assert db.stats().diagonal_steps == int(ref.dsteps.sum())


@pytest.mark.parametrize("metric,pen", SETS, ids=IDS)
def test_short_pairs_vs_oracle(metric, pen):
    say("ring depths, short pairs: %s %s" % (metric, pen))
    a, b = datagen.generate(1, 5101, 800), datagen.generate(0, 5102, 20)
    ps = datagen.PairSet.from_lists(a.patterns() + b.patterns(), a.texts() + b.texts())
    for mem, so in (("high", False), ("high", True), ("med", False)):
        check(metric, pen, ps, mem, so)


@pytest.mark.parametrize("metric,pen", SETS, ids=IDS)
def test_long_pairs_vs_oracle(metric, pen):
    say("ring depths, long pairs: %s %s" % (metric, pen))
    ps = datagen.generate(2, 5103, 10)
    ps = ps.subset([i for i in range(ps.n) if max(ps.plen[i], ps.tlen[i]) <= 8000][:4])
    check(metric, pen, ps)
