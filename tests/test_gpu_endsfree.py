"""Ends-free and extension forms on the GPU, through the C ABI, bit-exact
with the oracle (status, score, every CIGAR operation, diagonal-steps).

The oracle follows the reference's scalar ends-free extension, which ends at
the first diagonal in ascending order that reaches a free end
(wavefront_extend_kernels.c:154-196); extension trims the CIGAR to its best
prefix (alignment/cigar_utils.c:111-278) and reports a partial status when
it was cut.
"""
import sys

import pytest

import oracle as O
import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu

# every ends-free form the device path accepts
SETS = [
    ("affine", (0, 4, 6, 2), dict(ends_free=(0, 0, 30, 30))),
    ("affine", (0, 4, 6, 2), dict(ends_free=(30, 30, 0, 0))),
    ("affine", (0, 4, 6, 2), dict(ends_free=(10, 10, 10, 10))),
    ("affine", (-2, 4, 6, 2), dict(ends_free=(0, 30, 0, 30))),
    ("affine", (0, 4, 6, 2), dict(extension=True)),
    ("affine", (-1, 3, 5, 2), dict(extension=True)),
    ("affine2p", (0, 4, 6, 2, 24, 1), dict(ends_free=(20, 20, 20, 20))),
    ("affine2p", (0, 4, 6, 2, 24, 1), dict(extension=True)),
    ("linear", (0, 4, 2), dict(ends_free=(30, 30, 30, 30))),
    ("linear", (-2, 3, 4), dict(extension=True)),
    ("edit", (0, 1, 1), dict(ends_free=(0, 30, 0, 30))),
    ("edit", (0, 1, 1), dict(extension=True)),
    ("indel", (0, 1, 1), dict(ends_free=(0, 0, 0, 30))),
]
IDS = ["%s%s-%s" % (m, p, "ext" if f.get("extension") else "ef%s" % (f["ends_free"],)) for m, p, f in SETS]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(metric, pen, form, ps, memory="high", score_only=False):
    al = wfa_amd.Aligner(metric=metric, penalties=pen, memory=memory, score_only=score_only, **form)
    db = al.upload(ps)
    rc = db.run()
    assert rc in (0, 1), rc
    got = db.download()
    ref = O.oracle_align(O.Conf(metric, pen, memory, score_only, **form), O.Pairs(ps.patterns(), ps.texts()))
    return got, ref, db.stats()


def assert_same(got, ref, ps, score_only=False):
    bad = [i for i in range(ps.n)
           if got.status[i] != ref.status[i] or got.score[i] != ref.score[i] or
           (not score_only and got.ops(i) != ref.ops(i))]
    assert not bad, "pairs differ: %s (first: status %d/%d score %d/%d)" % (
        bad[:10], got.status[bad[0]], ref.status[bad[0]], got.score[bad[0]], ref.score[bad[0]])


@pytest.mark.parametrize("metric,pen,form", SETS, ids=IDS)
def test_endsfree_vs_oracle(metric, pen, form):
    say("ends-free %s %s %s" % (metric, pen, form))
    ps = datagen.semiglobal(1, 3101, 1500)
    mems = ("high",) if metric in ("edit", "indel") and form.get("extension") else ("high", "med")
    for mem in mems:
        for so in (False, True):
            got, ref, st = run_both(metric, pen, form, ps, mem, so)
            assert_same(got, ref, ps, so)
            if not so:
                assert st.diagonal_steps == int(ref.dsteps.sum())


@pytest.mark.parametrize("metric,pen,form", [
    ("affine", (0, 4, 6, 2), dict(ends_free=(0, 0, 300, 300))),
    ("affine2p", (0, 4, 6, 2, 24, 1), dict(extension=True)),
    ("edit", (0, 1, 1), dict(ends_free=(0, 300, 0, 300))),
])
def test_endsfree_long_pairs_vs_oracle(metric, pen, form):
    """1-8 kbp pairs with flanks of up to 300 bases (HBM-ring kernel)."""
    say("ends-free long pairs %s %s %s" % (metric, pen, form))
    ps = datagen.semiglobal(2, 3102, 8, flank=300)
    ps = ps.subset([i for i in range(ps.n) if max(ps.plen[i], ps.tlen[i]) <= 8000][:5])
    got, ref, st = run_both(metric, pen, form, ps)
    assert_same(got, ref, ps)
    assert st.diagonal_steps == int(ref.dsteps.sum())


def test_wide_free_beginning_vs_oracle():
    """A free beginning wider than the one-wave ring: the score-0 wavefront
    overflows it and the pair is re-run on a wider ring."""
    say("wide free beginning")
    ps = datagen.semiglobal(1, 3103, 64, flank=30)
    L = int(min(ps.plen.min(), ps.tlen.min()))
    form = dict(ends_free=(L, L, L, L))
    got, ref, st = run_both("affine", (0, 4, 6, 2), form, ps)
    assert_same(got, ref, ps)
    assert st.diagonal_steps == int(ref.dsteps.sum())


def test_endsfree_edge_cases_vs_oracle():
    say("ends-free edge cases")
    pats = [b"", b"ACGT", b"", b"A", b"ACGTACGT", b"NNACGTNN", b"acgtACGT", b"T" * 50]
    txts = [b"", b"", b"ACG", b"C", b"TTACGTACGTTT", b"ACGT", b"ACGT", b"G" * 60]
    ps = datagen.PairSet.from_lists(pats, txts)
    for metric, pen, form in (("affine", (0, 4, 6, 2), dict(ends_free=(0, 0, 0, 0))),
                              ("affine", (0, 4, 6, 2), dict(extension=True)),
                              ("affine2p", (0, 4, 6, 2, 24, 1), dict(extension=True)),
                              ("linear", (0, 4, 2), dict(ends_free=(0, 0, 0, 0))),
                              ("edit", (0, 1, 1), dict(extension=True))):
        got, ref, _ = run_both(metric, pen, form, ps)
        assert_same(got, ref, ps)


def test_single_pair_entry_and_partial_status():
    """wavefront_align returns the batch's result, including the partial
    status of a trimmed extension."""
    say("ends-free single-pair entry")
    ps = datagen.semiglobal(1, 3104, 40)
    for metric, pen, form in (("affine", (0, 4, 6, 2), dict(extension=True)),
                              ("linear", (0, 4, 2), dict(ends_free=(5, 5, 5, 5)))):
        al = wfa_amd.Aligner(metric=metric, penalties=pen, **form)
        res = al.align_batch(ps)
        for i in range(ps.n):
            assert al.align(ps.pattern(i), ps.text(i)) == (int(res.status[i]), int(res.score[i]), res.ops(i))
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), extension=True)
    st, sc, ops = al.align(b"ACGTACGTACGTACGTACGT" + b"TTTTTTTTTT", b"ACGTACGTACGTACGTACGT" + b"GGGGGGGGGG")
    assert st == wfa_amd.WF_STATUS_ALG_PARTIAL and ops == b"M" * 20 and sc == 20
