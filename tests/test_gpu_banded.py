"""Banded heuristics on the GPU (wavefront_heuristic.c:451-500): banded
static and banded adaptive, alone and after wf-adaptive, in every metric and
in memory modes high and ultralow, bit-exact with the oracle (status, score,
every CIGAR operation).  Bands that keep the end out of reach end as the
reference does: WF_STATUS_ALG_PARTIAL with an empty CIGAR for the gap
metrics (END_UNREACHABLE), WF_STATUS_UNATTAINABLE in BiWFA, the step limit
for indel / edit.  A pair whose reference backtrace would take operations
back through a wavefront the cut-off emptied (the oracle's stale flag) is
declined with WF_STATUS_UNSUPPORTED, and only such a pair."""
import sys

import pytest

import oracle as O
import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu

BANDED = [
    (1, 0, 0, 0, -10, 10),
    (1, 0, 0, 0, -3, 40),
    (1, 0, 0, 0, 5, 10),        # diagonal 0 outside the band: nothing is reachable
    (2, 0, 0, 1, -10, 10),
    (2, 0, 0, 3, -5, 5),
    (5, 10, 50, 1, -20, 20),
    (6, 10, 50, 2, -8, 8),
]
METRICS = [("affine", (0, 4, 6, 2)), ("affine", (-3, 5, 3, 2)), ("affine2p", (0, 4, 6, 2, 24, 1)),
           ("linear", (0, 4, 2)), ("edit", (0, 1, 1)), ("indel", (0, 1, 1))]
MAX_STEPS = 30000


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pairs(seed):
    a, b = datagen.generate(1, seed, 800), datagen.generate(0, seed + 1, 40)
    s = datagen.semiglobal(1, seed + 2, 200)
    return datagen.PairSet.from_lists(a.patterns() + b.patterns() + s.patterns(),
                                      a.texts() + b.texts() + s.texts())


def check(metric, pen, heur, ps, memory, score_only):
    al = wfa_amd.Aligner(metric=metric, penalties=pen, memory=memory, score_only=score_only, heuristic=heur,
                         max_steps=MAX_STEPS)
    db = al.upload(ps)
    db.run()
    got = db.download()
    ref = O.oracle_align(O.Conf(metric, pen, memory, score_only, heuristic=heur, max_steps=MAX_STEPS),
                         O.Pairs(ps.patterns(), ps.texts()))
    declined = [i for i in range(ps.n) if got.status[i] == wfa_amd.WF_STATUS_UNSUPPORTED]
    assert declined == [i for i in range(ps.n) if ref.stale[i]], (heur, memory, score_only)
    bad = [i for i in range(ps.n) if not ref.stale[i] and
           (got.status[i] != ref.status[i] or got.score[i] != ref.score[i] or
            (not score_only and got.ops(i) != ref.ops(i)))]
    assert not bad, "%s %s so=%s: pairs differ %s (first: status %d/%d score %d/%d)" % (
        heur, memory, score_only, bad[:10], got.status[bad[0]], ref.status[bad[0]], got.score[bad[0]],
        ref.score[bad[0]])
    return got, ref


@pytest.mark.parametrize("metric,pen", METRICS, ids=["%s%s" % m for m in METRICS])
def test_banded_vs_oracle(metric, pen):
    ps = pairs(4301)
    for heur in BANDED:
        for mem in ("high", "ultralow"):
            for so in (False, True):
                say("banded %s %s %s %s so=%s" % (metric, pen, heur, mem, so))
                check(metric, pen, heur, ps, mem, so)


def test_banded_long_pairs_vs_oracle():
    """1-10 kbp pairs on the HBM-ring and BiWFA kernels."""
    ps = datagen.generate(2, 4302, 10)
    ps = ps.subset([i for i in range(ps.n) if ps.plen[i] <= 10000][:6])
    for metric, pen in (("affine", (0, 4, 6, 2)), ("edit", (0, 1, 1))):
        for heur in ((1, 0, 0, 0, -300, 300), (2, 0, 0, 1, -100, 100), (5, 10, 50, 1, -200, 200)):
            for mem in ("high", "ultralow"):
                say("banded long %s %s %s" % (metric, heur, mem))
                check(metric, pen, heur, ps, mem, False)


def test_banded_single_pair_entry_point():
    """wavefront_align with a band: an unreachable end comes back as
    WF_STATUS_ALG_PARTIAL, score INT32_MIN, empty CIGAR, like the batch."""
    say("banded single pair")
    ps = datagen.semiglobal(1, 4303, 60)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), heuristic=(1, 0, 0, 0, -5, 5))
    ref = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), heuristic=(1, 0, 0, 0, -5, 5)),
                         O.Pairs(ps.patterns(), ps.texts()))
    assert O.WF_STATUS_ALG_PARTIAL in set(int(x) for x in ref.status)
    for i in range(ps.n):
        if ref.stale[i]:
            continue
        assert al.align(ps.pattern(i), ps.text(i)) == (int(ref.status[i]), int(ref.score[i]), ref.ops(i)), i
