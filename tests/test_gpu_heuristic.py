"""The wf-adaptive heuristic on the GPU (the reference library's default
strategy), bit-exact with the oracle: status, score, every CIGAR operation
and the diagonal-step count.  The reference's own wf-adaptive golden files
(test.*.affine.wfapt{0,1}.alg) run in test_gpu_parity.py."""
import sys

import pytest

import oracle as O
import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu

SETS = [
    ("affine", (0, 4, 6, 2), (4, 10, 50, 1)),
    ("affine", (-3, 5, 3, 2), (4, 10, 50, 10)),
    ("affine2p", (0, 4, 6, 2, 24, 1), (4, 10, 50, 1)),
    ("linear", (0, 4, 2), (4, 50, 10, 3)),
    ("edit", (0, 1, 1), (4, 10, 50, 1)),
    ("indel", (0, 1, 1), (4, 10, 50, 10)),
    ("affine", (0, 4, 6, 2), (4, 1, 0, 1)),   # the tightest cut at every step
]
IDS = ["%s%s-h%s" % (m, p, h[1:]) for m, p, h in SETS]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(metric, pen, heur, ps, memory="high", score_only=False, **form):
    al = wfa_amd.Aligner(metric=metric, penalties=pen, memory=memory, score_only=score_only, heuristic=heur, **form)
    db = al.upload(ps)
    assert db.run() in (0, 1)
    got = db.download()
    ref = O.oracle_align(O.Conf(metric, pen, memory, score_only, heuristic=heur, **form),
                         O.Pairs(ps.patterns(), ps.texts()))
    return got, ref, db.stats()


def assert_same(got, ref, ps, score_only=False):
    bad = [i for i in range(ps.n)
           if got.status[i] != ref.status[i] or got.score[i] != ref.score[i] or
           (not score_only and got.ops(i) != ref.ops(i))]
    assert not bad, "pairs differ: %s (first: status %d/%d score %d/%d)" % (
        bad[:10], got.status[bad[0]], ref.status[bad[0]], got.score[bad[0]], ref.score[bad[0]])


def mixed(seed):
    a, b = datagen.generate(1, seed, 1200), datagen.generate(0, seed + 1, 60)
    return datagen.PairSet.from_lists(a.patterns() + b.patterns(), a.texts() + b.texts())


@pytest.mark.parametrize("metric,pen,heur", SETS, ids=IDS)
def test_wfadaptive_vs_oracle(metric, pen, heur):
    say("wf-adaptive %s %s %s" % (metric, pen, heur))
    ps = mixed(4101)
    for mem in ("high", "ultralow"):
        for so in (False, True):
            got, ref, st = run_both(metric, pen, heur, ps, mem, so)
            assert_same(got, ref, ps, so)
            if not so:
                assert st.diagonal_steps == int(ref.dsteps.sum()), mem


@pytest.mark.parametrize("metric,pen", [("affine", (0, 4, 6, 2)), ("affine2p", (0, 4, 6, 2, 24, 1)),
                                        ("edit", (0, 1, 1))])
def test_wfadaptive_long_pairs_vs_oracle(metric, pen):
    """1-10 kbp pairs (HBM-ring and BiWFA kernels), where the cut-off pays."""
    say("wf-adaptive long pairs %s %s" % (metric, pen))
    ps = datagen.generate(2, 4102, 10)
    ps = ps.subset([i for i in range(ps.n) if ps.plen[i] <= 10000][:6])
    for mem in ("high", "ultralow"):
        got, ref, st = run_both(metric, pen, (4, 10, 50, 1), ps, mem)
        assert_same(got, ref, ps)
        assert st.diagonal_steps == int(ref.dsteps.sum()), mem


def test_wfadaptive_endsfree_vs_oracle():
    say("wf-adaptive ends-free")
    ps = datagen.semiglobal(1, 4103, 800)
    for metric, pen, form in (("affine", (0, 4, 6, 2), dict(ends_free=(0, 0, 30, 30))),
                              ("affine", (0, 4, 6, 2), dict(extension=True)),
                              ("edit", (0, 1, 1), dict(ends_free=(0, 30, 0, 30)))):
        got, ref, st = run_both(metric, pen, (4, 10, 50, 1), ps, **form)
        assert_same(got, ref, ps)
        assert st.diagonal_steps == int(ref.dsteps.sum())


def test_cutoffs_on_endsfree_wavefronts_vs_oracle():
    """Cut-offs on ends-free wavefronts, including the score-0 wavefront over
    wide free beginnings: the tightest wf-adaptive cut, wfmash, a static band
    that narrows the score-0 range and an adaptive band.  A pair whose
    reference trace would take operations back through a wavefront a cut-off
    emptied (the oracle's stale flag) is declined, and only such a pair."""
    say("cut-offs on ends-free wavefronts")
    ps = datagen.semiglobal(1, 4105, 600)
    for heur, form in (((4, 1, 0, 1), dict(ends_free=(20, 20, 20, 20))),
                       ((4, 10, 50, 1), dict(ends_free=(30, 30, 0, 0))),
                       ((0x40, 10, 50, 1), dict(extension=True)),
                       ((1, 0, 0, 0, -5, 5), dict(ends_free=(10, 10, 10, 10))),
                       ((2, 0, 0, 1, -8, 8), dict(ends_free=(0, 0, 30, 30)))):
        for so in (False, True):
            got, ref, st = run_both("affine", (0, 4, 6, 2), heur, ps, "high", so, **form)
            declined = [i for i in range(ps.n) if got.status[i] == wfa_amd.WF_STATUS_UNSUPPORTED]
            assert declined == [i for i in range(ps.n) if (not so) and ref.stale[i]], (heur, form, so)
            bad = [i for i in range(ps.n) if i not in declined and
                   (got.status[i] != ref.status[i] or got.score[i] != ref.score[i] or
                    (not so and got.ops(i) != ref.ops(i)))]
            assert not bad, (heur, form, so, bad[:10])


def test_default_attributes_are_wfadaptive():
    """An aligner made from wavefront_aligner_attr_default (wf-adaptive
    10, 50, 1: wavefront_attributes.c:38-100) aligns like the oracle with
    that heuristic, through the single-pair entry point too."""
    say("default attributes")
    ps = datagen.generate(1, 4104, 200)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), heuristic_none=False)
    got = al.align_batch(ps)
    ref = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), heuristic=(4, 10, 50, 1)), O.Pairs(ps.patterns(), ps.texts()))
    assert_same(got, ref, ps)
    for i in range(0, ps.n, 20):
        assert al.align(ps.pattern(i), ps.text(i)) == (int(ref.status[i]), int(ref.score[i]), ref.ops(i))
