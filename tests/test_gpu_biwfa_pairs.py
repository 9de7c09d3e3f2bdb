"""BiWFA breakpoint searches on paired workgroups (the forward and the
reverse side of a large sub-problem on two workgroups, exchanging every
score) against the same searches on one workgroup, and against the oracle.

Paired searches run when a launch has fewer large sub-problems than half the
workgroups the device holds; WFA_PAIR_SEARCH=0 turns them off, which gives
the single-workgroup search (itself pinned against the oracle by the other
BiWFA tests) on the same pairs.  Status, score, every CIGAR operation and the
diagonal-step totals must agree."""
import os
import sys

import numpy as np
import pytest

import oracle as O
import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu

SETS = [("affine", (0, 4, 6, 2)), ("affine2p", (0, 4, 6, 2, 24, 1)), ("edit", (0, 1, 1)), ("indel", (0, 1, 1)),
        ("linear", (0, 4, 2))]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def long_pairs(seed, n, lo=6000, hi=16000):
    full = datagen.generate(2, seed, 40 * n)
    keep = [i for i in range(full.n) if lo <= max(full.plen[i], full.tlen[i]) <= hi][:n]
    return full.subset(keep)


def run(metric, pen, ps, paired):
    old = os.environ.get("WFA_PAIR_SEARCH")
    os.environ["WFA_PAIR_SEARCH"] = "1" if paired else "0"
    try:
        al = wfa_amd.Aligner(metric=metric, penalties=pen, memory="ultralow")
        db = al.upload(ps)
        assert db.run() == 0
        return db.download(), db.stats()
    finally:
        if old is None:
            del os.environ["WFA_PAIR_SEARCH"]
        else:
            os.environ["WFA_PAIR_SEARCH"] = old


def same(a, b, n):
    bad = [i for i in range(n) if a.status[i] != b.status[i] or a.score[i] != b.score[i] or a.ops(i) != b.ops(i)]
    assert not bad, bad[:10]


@pytest.mark.parametrize("metric,pen", SETS, ids=[m for m, _ in SETS])
def test_paired_search_equals_single_workgroup(metric, pen):
    ps = long_pairs(831, 6)
    say("paired vs single-workgroup BiWFA, %s %s: %d pairs of %s bp" % (metric, pen, ps.n, list(ps.plen)))
    got, st = run(metric, pen, ps, True)
    ref, st0 = run(metric, pen, ps, False)
    same(got, ref, ps.n)
    assert st.diagonal_steps == st0.diagonal_steps


@pytest.mark.parametrize("metric,pen", [("edit", (0, 1, 1)), ("affine", (0, 4, 6, 2))])
def test_paired_search_vs_oracle(metric, pen):
    ps = long_pairs(832, 2, 8500, 11000)
    say("paired BiWFA vs oracle, %s: %d pairs" % (metric, ps.n))
    got, st = run(metric, pen, ps, True)
    ref = O.oracle_align(O.Conf(metric, pen, "ultralow"), O.Pairs(ps.patterns(), ps.texts()))
    assert np.array_equal(got.status, ref.status) and np.array_equal(got.score, ref.score)
    assert all(got.ops(i) == ref.ops(i) for i in range(ps.n))
    assert st.diagonal_steps == int(ref.dsteps.sum())


def test_timed_out_paired_searches_fall_back_to_one_workgroup():
    """A wait bound of a few device-clock ticks makes paired searches give
    up (kBiPairAbort) instead of waiting for their partner; the host then
    searches those sub-problems again on one workgroup, with the same
    results.  Every wait stays bounded, so no workgroup can hang."""
    ps = long_pairs(834, 4, 8000, 14000)
    say("paired BiWFA with a 2-tick wait bound: %d pairs" % ps.n)
    old = os.environ.get("WFA_PAIR_WAIT_TICKS")
    os.environ["WFA_PAIR_WAIT_TICKS"] = "2"
    try:
        got, st = run("affine", (0, 4, 6, 2), ps, True)
    finally:
        if old is None:
            del os.environ["WFA_PAIR_WAIT_TICKS"]
        else:
            os.environ["WFA_PAIR_WAIT_TICKS"] = old
    ref, st0 = run("affine", (0, 4, 6, 2), ps, False)
    same(got, ref, ps.n)
