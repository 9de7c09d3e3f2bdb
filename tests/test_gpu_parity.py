"""Parity of the HIP path with the reference, through the C ABI, on the GPU.

Bit-exact: status, cigar->score and every CIGAR operation, plus the number
of diagonal-steps (which pins the exact [lo,hi] evolution of every score
step, including the reference's trimming and null-wavefront limits).
"""
import sys

import numpy as np
import pytest

import golden_suite as G
import oracle as O
import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_run(metric, pen, ps, memory="high", score_only=False, max_steps=0):
    conf = O.Conf(metric, pen, memory, score_only, max_steps)
    return O.oracle_align(conf, O.Pairs(ps.patterns(), ps.texts()))


def assert_same(got, ref, ps, check_dsteps=None):
    bad = []
    for i in range(ps.n):
        if got.status[i] != ref.status[i] or got.score[i] != ref.score[i] or got.ops(i) != ref.ops(i):
            bad.append(i)
    assert not bad, "pairs differ: %s (first: status %d/%d score %d/%d)" % (
        bad[:10], got.status[bad[0]], ref.status[bad[0]], got.score[bad[0]], ref.score[bad[0]])
    if check_dsteps is not None:
        assert np.array_equal(check_dsteps, ref.dsteps)


GOLDEN_GPU = G.cases()  # every golden file, the wf-adaptive ones included


@pytest.fixture(scope="module")
def suite():
    p, t = G.read_seq()
    return datagen.PairSet.from_lists(p, t)


@pytest.mark.parametrize("case", GOLDEN_GPU, ids=[c[0] for c in GOLDEN_GPU])
def test_golden_suite(suite, case):
    name, metric, pen, mem, so, heur = case
    say("golden %s on GPU" % name)
    al = wfa_amd.Aligner(metric=metric, penalties=pen, memory=mem, score_only=so, heuristic=heur)
    res = al.align_batch(suite)
    lines = []
    conf = O.Conf(metric, pen, mem, so, heuristic=heur)
    for i in range(suite.n):
        assert res.status[i] == 0, (i, res.status[i])
        if so:
            lines.append("%d\t-" % res.score[i])
        else:
            lines.append("%d\t%s" % (O.cigar_score(conf, res.ops(i)), wfa_amd.rle(res.ops(i))))
    gold = G.golden_lines(name)
    diff = [i for i in range(len(gold)) if lines[i] != gold[i]]
    assert not diff, "%s: %d lines differ, first %d: %r vs %r" % (name, len(diff), diff[0], lines[diff[0]], gold[diff[0]])


@pytest.mark.parametrize("shape,n,seed", [(1, 4000, 101), (0, 300, 102)])
def test_random_affine_vs_oracle(shape, n, seed):
    say("random affine shape %d n=%d seed=%d" % (shape, n, seed))
    ps = datagen.generate(shape, seed, n)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    db = al.upload(ps)
    assert db.run() == 0
    got = db.download()
    ref = oracle_run("affine", (0, 4, 6, 2), ps)
    assert_same(got, ref, ps)
    assert db.stats().diagonal_steps == int(ref.dsteps.sum())


@pytest.mark.parametrize("metric,pen", [("affine", (-3, 5, 3, 2)), ("affine2p", (0, 4, 6, 2, 24, 1)),
                                        ("affine", (0, 1, 2, 1)), ("affine2p", (-2, 3, 4, 3, 12, 1))])
def test_random_penalties_vs_oracle(metric, pen):
    say("random %s %s" % (metric, pen))
    ps = datagen.generate(1, 202, 1500)
    al = wfa_amd.Aligner(metric=metric, penalties=pen)
    db = al.upload(ps)
    assert db.run() == 0
    got = db.download()
    ref = oracle_run(metric, pen, ps)
    assert_same(got, ref, ps)
    assert db.stats().diagonal_steps == int(ref.dsteps.sum())


def test_affine2p_long_pairs_vs_oracle():
    say("affine2p 1-20 kbp pairs")
    ps = datagen.generate(2, 303, 6)
    keep = [i for i in range(ps.n) if ps.plen[i] <= 8000]
    ps = ps.subset(keep[:4])
    al = wfa_amd.Aligner(metric="affine2p", penalties=(0, 4, 6, 2, 24, 1))
    db = al.upload(ps)
    assert db.run() == 0
    got = db.download()
    ref = oracle_run("affine2p", (0, 4, 6, 2, 24, 1), ps)
    assert_same(got, ref, ps)
    assert db.stats().diagonal_steps == int(ref.dsteps.sum())


def test_diagonal_steps_match_oracle():
    say("diagonal-step totals")
    ps = datagen.generate(1, 404, 2000)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    db = al.upload(ps)
    assert db.run() == 0
    ref = oracle_run("affine", (0, 4, 6, 2), ps)
    assert db.stats().diagonal_steps == int(ref.dsteps.sum())


def test_edge_cases_vs_oracle():
    say("edge cases")
    t = b"ACGTTGCAACGTACGGATCCATGACCAGTACGATCG"
    pats = [b"", b"ACGT", b"", b"A", b"A", t, t[:17] + t[18:], b"acgtNNacgt", b"ACGTNNACGT", b"G" * 70,
            t + t + t, b"T"]
    texts = [b"", b"", b"ACG", b"A", b"C", t, t, b"acgtNNacgt", b"acgtNNacgt", b"G" * 3, t, b"GGGGGGGGGGGGT"]
    ps = datagen.PairSet.from_lists(pats, texts)
    for metric, pen in (("affine", (0, 4, 6, 2)), ("affine2p", (0, 4, 6, 2, 24, 1)), ("affine", (-5, 1, 2, 1))):
        al = wfa_amd.Aligner(metric=metric, penalties=pen)
        got = al.align_batch(ps)
        ref = oracle_run(metric, pen, ps)
        assert_same(got, ref, ps)


def test_single_pair_entry_matches_batch():
    say("wavefront_align vs wavefront_align_batch")
    ps = datagen.generate(1, 505, 64)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    batch = al.align_batch(ps)
    for i in range(ps.n):
        st, score, ops = al.align(ps.pattern(i), ps.text(i))
        assert (st, score, ops) == (batch.status[i], batch.score[i], batch.ops(i))


def test_score_only_and_max_steps():
    say("score-only and max-steps")
    ps = datagen.generate(1, 606, 500)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), score_only=True)
    got = al.align_batch(ps)
    ref = oracle_run("affine", (0, 4, 6, 2), ps, score_only=True)
    assert np.array_equal(got.status, ref.status) and np.array_equal(got.score, ref.score)
    al2 = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), max_steps=40)
    got2 = al2.align_batch(ps)
    ref2 = oracle_run("affine", (0, 4, 6, 2), ps, max_steps=40)
    assert np.array_equal(got2.status, ref2.status) and np.array_equal(got2.score, ref2.score)
    assert (got2.status == wfa_amd.WF_STATUS_MAX_STEPS_REACHED).any()
    for i in range(ps.n):
        if ref2.status[i] == 0:
            assert got2.ops(i) == ref2.ops(i)


def test_device_batch_rerun_is_identical():
    say("device batch run twice")
    ps = datagen.generate(1, 707, 3000)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    db = al.upload(ps)
    assert db.run() == 0
    a = db.download()
    assert db.run() == 0
    b = db.download()
    assert np.array_equal(a.score, b.score)
    assert all(a.ops(i) == b.ops(i) for i in range(ps.n))


@pytest.mark.parametrize("metric,pen,shape,n,seed", [
    ("affine", (0, 4, 6, 2), 2, 6, 801),
    ("affine2p", (0, 4, 6, 2, 24, 1), 2, 4, 802),
    ("affine", (0, 4, 6, 2), 1, 300, 803),
    ("affine", (-3, 5, 3, 2), 2, 4, 804),
])
def test_biwfa_vs_oracle(metric, pen, shape, n, seed):
    say("BiWFA %s %s shape %d n=%d" % (metric, pen, shape, n))
    ps = datagen.generate(shape, seed, n)
    if shape == 2:
        keep = [i for i in range(ps.n) if ps.plen[i] <= 12000]
        ps = ps.subset(keep)
    al = wfa_amd.Aligner(metric=metric, penalties=pen, memory="ultralow")
    db = al.upload(ps)
    assert db.run() == 0
    got = db.download()
    ref = oracle_run(metric, pen, ps, memory="ultralow")
    assert_same(got, ref, ps)
    assert db.stats().diagonal_steps == int(ref.dsteps.sum())


def test_biwfa_score_only_vs_oracle():
    say("BiWFA score-only")
    ps = datagen.generate(2, 805, 4)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), memory="ultralow", score_only=True)
    got = al.align_batch(ps)
    ref = oracle_run("affine", (0, 4, 6, 2), ps, memory="ultralow", score_only=True)
    assert np.array_equal(got.status, ref.status) and np.array_equal(got.score, ref.score)


def test_biwfa_long_noisy_pair_vs_oracle():
    say("BiWFA 30 kbp 15% pair")
    # a 100 kbp / 15% pair cut to 30 kbp keeps the run short on the oracle
    full = datagen.generate(3, 806, 1)
    p, t = full.pattern(0)[:30000], full.text(0)[:30000]
    ps = datagen.PairSet.from_lists([p], [t])
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), memory="ultralow")
    got = al.align_batch(ps)
    ref = oracle_run("affine", (0, 4, 6, 2), ps, memory="ultralow")
    assert_same(got, ref, ps)


LEAN_CHILD = r"""
import sys
sys.path.insert(0, %r); sys.path.insert(0, %r); sys.path.insert(0, %r)
import numpy as np, oracle as O, wfa_amd
from wfa_amd import datagen
a, b = datagen.generate(1, 1001, 1500), datagen.generate(0, 1002, 150)
ps = datagen.PairSet.from_lists(a.patterns() + b.patterns(), a.texts() + b.texts())
al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
db = al.upload(ps)
assert db.run() == 0
got = db.download()
ref = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), "high"), O.Pairs(ps.patterns(), ps.texts()))
bad = [i for i in range(ps.n) if got.status[i] != ref.status[i] or got.score[i] != ref.score[i] or got.ops(i) != ref.ops(i)]
assert not bad, bad[:10]
assert db.stats().diagonal_steps == int(ref.dsteps.sum())
print("launches", db.stats().forward_launches)
"""


@pytest.mark.parametrize("env", [{"WFA_LEAN_WCAP": "224,224,4000"}, {"WFA_LEAN": "0"}],
                         ids=["lean-ring-overflow-retries", "generic-lds-family"])
def test_kernel_families_vs_oracle(env):
    """The lean family with its narrowest ring for every pair (the 1 kbp
    pairs overflow it and are re-run with wider rings), and the generic LDS
    family with the lean one turned off, both match the oracle."""
    import os
    import subprocess
    from conftest import ROOT
    say("kernel family %s" % env)
    code = LEAN_CHILD % (os.path.join(ROOT, "wfa2-lib_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"))
    r = subprocess.run([sys.executable, "-c", code], env=dict(os.environ, **env), capture_output=True, timeout=240)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    if "WFA_LEAN_WCAP" in env:
        assert int(r.stdout.split()[-1]) > 1  # the retries ran
