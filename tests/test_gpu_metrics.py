"""Indel, edit and gap-linear distances on the GPU (single-component
wavefronts: M only), through the C ABI, bit-exact with the oracle.

Status, score, every CIGAR operation and the diagonal-step count (which pins
each score step's [lo, hi], the trimming and, for edit, the exact pruning of
wavefront_compute_edit.c:219-278) in memory modes high and ultralow (BiWFA),
with CIGARs and score only.  The reference's own golden files for edit and
indel run in test_gpu_parity.py.
"""
import sys

import pytest

import oracle as O
import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu

SETS = [("edit", (0, 1, 1)), ("indel", (0, 1, 1)), ("linear", (0, 4, 2)), ("linear", (-1, 3, 2)),
        ("linear", (0, 3, 5))]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(metric, pen, ps, memory, score_only=False):
    al = wfa_amd.Aligner(metric=metric, penalties=pen, memory=memory, score_only=score_only)
    db = al.upload(ps)
    assert db.run() == 0
    got = db.download()
    conf = O.Conf(metric, pen, memory, score_only)
    ref = O.oracle_align(conf, O.Pairs(ps.patterns(), ps.texts()))
    return got, ref, db.stats()


def assert_same(got, ref, ps, score_only=False):
    bad = [i for i in range(ps.n)
           if got.status[i] != ref.status[i] or got.score[i] != ref.score[i] or
           (not score_only and got.ops(i) != ref.ops(i))]
    assert not bad, "pairs differ: %s (first: status %d/%d score %d/%d)" % (
        bad[:10], got.status[bad[0]], ref.status[bad[0]], got.score[bad[0]], ref.score[bad[0]])


def mixed(seed):
    a, b = datagen.generate(1, seed, 800), datagen.generate(0, seed + 1, 60)
    return datagen.PairSet.from_lists(a.patterns() + b.patterns(), a.texts() + b.texts())


@pytest.mark.parametrize("memory", ["high", "ultralow"])
@pytest.mark.parametrize("metric,pen", SETS, ids=["%s%s" % (m, p) for m, p in SETS])
def test_random_vs_oracle(metric, pen, memory):
    say("%s %s %s: short reads + 1 kbp" % (metric, pen, memory))
    ps = mixed(2101)
    got, ref, st = run_both(metric, pen, ps, memory)
    assert_same(got, ref, ps)
    assert st.diagonal_steps == int(ref.dsteps.sum())


@pytest.mark.parametrize("memory", ["high", "ultralow"])
@pytest.mark.parametrize("metric,pen", [("edit", (0, 1, 1)), ("indel", (0, 1, 1)), ("linear", (-1, 3, 2))])
def test_skewed_pairs_vs_oracle(metric, pen, memory):
    """Pairs with a 1.2-2.4 kbp insertion: wavefronts wider than 1000
    diagonals, where edit prunes (the step count pins every pruned range)."""
    say("%s %s %s: skewed pairs" % (metric, pen, memory))
    ps = datagen.skewed(2102, 12)
    got, ref, st = run_both(metric, pen, ps, memory)
    assert_same(got, ref, ps)
    assert st.diagonal_steps == int(ref.dsteps.sum())


@pytest.mark.parametrize("memory", ["high", "ultralow"])
@pytest.mark.parametrize("metric,pen", [("edit", (0, 1, 1)), ("linear", (0, 4, 2))])
def test_score_only_vs_oracle(metric, pen, memory):
    say("%s %s %s score only" % (metric, pen, memory))
    a, b = datagen.generate(1, 2103, 300), datagen.skewed(2104, 4)
    ps = datagen.PairSet.from_lists(a.patterns() + b.patterns(), a.texts() + b.texts())
    got, ref, _ = run_both(metric, pen, ps, memory, score_only=True)
    assert_same(got, ref, ps, score_only=True)


@pytest.mark.parametrize("metric,pen", [("edit", (0, 1, 1)), ("linear", (0, 4, 2))])
def test_long_pairs_vs_oracle(metric, pen):
    """1-8 kbp pairs on the HBM-ring kernel, both memory modes."""
    say("%s %s: 1-8 kbp pairs" % (metric, pen))
    ps = datagen.generate(2, 2105, 8)
    ps = ps.subset([i for i in range(ps.n) if ps.plen[i] <= 8000][:5])
    for memory in ("high", "ultralow"):
        got, ref, st = run_both(metric, pen, ps, memory)
        assert_same(got, ref, ps)
        assert st.diagonal_steps == int(ref.dsteps.sum())


def test_edge_cases_vs_oracle():
    say("single-component edge cases")
    pats = [b"", b"ACGT", b"", b"A", b"A", b"ACGTACGT", b"AAAAAAAAAA", b"T" * 70]
    txts = [b"", b"", b"ACG", b"A", b"C", b"TGCATGCA", b"A", b"G" * 70]
    ps = datagen.PairSet.from_lists(pats, txts)
    for metric, pen in SETS:
        for memory in ("high", "ultralow"):
            got, ref, _ = run_both(metric, pen, ps, memory)
            assert_same(got, ref, ps)


def test_single_pair_entry_matches_batch():
    say("wavefront_align vs batch, edit and linear")
    ps = datagen.generate(1, 2106, 40)
    for metric, pen in (("edit", (0, 1, 1)), ("linear", (-1, 3, 2))):
        al = wfa_amd.Aligner(metric=metric, penalties=pen)
        res = al.align_batch(ps)
        for i in range(ps.n):
            assert al.align(ps.pattern(i), ps.text(i)) == (int(res.status[i]), int(res.score[i]), res.ops(i))


def test_classic_scores():
    """cigar->score: the distance for edit / indel, the negated penalty for
    gap-linear, and matches minus penalties with a match bonus
    (wavefront_compute.c:108-120)."""
    say("classic scores")
    t = b"ACGTTGCAACGTACGGATCCATGACCAGTACGATCG"  # 36 bases
    sub = t[:10] + (b"A" if t[10:11] != b"A" else b"C") + t[11:]
    ps = datagen.PairSet.from_lists([t, sub, t[:20] + t[23:]], [t, t, t])
    for metric, pen, want in (("edit", (0, 1, 1), [0, 1, 3]), ("indel", (0, 1, 1), [0, 2, 3]),
                              ("linear", (0, 4, 2), [0, -4, -6]), ("linear", (-1, 3, 2), [36, 32, 27])):
        res = wfa_amd.Aligner(metric=metric, penalties=pen).align_batch(ps)
        assert list(res.status) == [0, 0, 0]
        assert [int(x) for x in res.score] == want, (metric, pen, list(res.score))
