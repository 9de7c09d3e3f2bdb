"""Single-pair calls (wavefront_align) replay a captured graph per launch
configuration (upload, counter reset, forward, backtrace, downloads).  Every
replay must give the result of the uncaptured path: here hundreds of
replays, with the program's own uncaptured work (batches on the same
aligner, which move its buffers and force re-captures, and on another
aligner) and synchronisations between them, are each compared with the
uncaptured batch result and the oracle; pairs that overflow the first ring
width (a retry) and score-only / wf-adaptive aligners are included."""
import os
import sys

import numpy as np
import pytest

import oracle as O
import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def with_wide_pairs(ps, rng, k):
    """k pairs whose pattern carries a 150-base insertion (wavefronts wider
    than the first ring width: those calls retry on the general path)."""
    pats, txts = ps.patterns(), ps.texts()
    for j in rng.choice(len(pats), size=k, replace=False):
        ins = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), 150))
        p = pats[j]
        pats[j] = p[: len(p) // 2] + ins + p[len(p) // 2:]
    return pats, txts


@pytest.mark.parametrize("kind", ["exact", "score_only", "wfadaptive"])
def test_single_pair_replays_equal_uncaptured(kind):
    os.environ.pop("WFA_PLAIN_GRAPH", None)
    kw = {"score_only": kind == "score_only"}
    if kind == "wfadaptive":
        kw["heuristic"] = (4, 10, 50, 1)
    ps = datagen.generate(1, 4711, 400)
    pats, txts = with_wide_pairs(ps, np.random.default_rng(3), 6)
    say("single-pair graph replays (%s): %d calls" % (kind, len(pats)))
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), **kw)
    other = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), **kw)
    want = other.align_batch(pats, txts)
    conf = O.Conf("affine", (0, 4, 6, 2), "high", score_only=kw["score_only"],
                  heuristic=kw.get("heuristic"))
    ref = O.oracle_align(conf, O.Pairs(pats, txts))
    extra = datagen.generate(0, 4712, 30)
    for i in range(len(pats)):
        st, sc, ops = al.align(pats[i], txts[i])
        assert (st, sc) == (int(want.status[i]), int(want.score[i])), i
        assert (st, sc) == (int(ref.status[i]), int(ref.score[i])), i
        if not kw["score_only"]:
            assert ops == want.ops(i) == ref.ops(i), i
        if i % 97 == 50:
            # uncaptured work between replays, on the same aligner and another one
            got = al.align_batch(extra)
            got2 = other.align_batch(extra)
            assert np.array_equal(got.score, got2.score)


def test_single_pairs_of_growing_lengths():
    """Lengths from 10 bp to 3 kbp: bucket, ring-width and family changes
    re-capture; each call equals the uncaptured batch result."""
    os.environ.pop("WFA_PLAIN_GRAPH", None)
    rng = np.random.default_rng(9)
    pats, txts = [], []
    for L in [10, 40, 120, 300, 310, 600, 1000, 1030, 2000, 3000]:
        for _ in range(3):
            t = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), L))
            p = bytearray(t)
            for _ in range(max(1, L // 25)):
                p[int(rng.integers(0, len(p)))] = ord("ACGT"[int(rng.integers(0, 4))])
            pats.append(bytes(p))
            txts.append(t)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    want = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2)).align_batch(pats, txts)
    say("single pairs of growing lengths: %d calls" % len(pats))
    for rep in range(2):
        for i in range(len(pats)):
            st, sc, ops = al.align(pats[i], txts[i])
            assert (st, sc, ops) == (int(want.status[i]), int(want.score[i]), want.ops(i)), (rep, i)
