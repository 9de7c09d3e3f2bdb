"""Long pairs through the HIP path against the reference's own results
(tests/golden/long_pairs.json, written by tests/golden/make_long_fixtures.py
from WFA2-lib built from source):

* 8-20 kbp dual-cost gap-affine (0,4,6,2,24,1) pairs of the 1-20 kbp
  workload, memory mode high (the HBM-ring forward kernel with 256/1024
  threads per pair and the record arena at those lengths);
* full 100 kbp / 15% gap-affine pairs in memory mode ultralow (BiWFA: the
  breakpoint kernel at 200k-diagonal rings, the breadth-first recursion, the
  base jobs and the CIGAR gather).

Status, score and every CIGAR operation must be identical."""
import hashlib
import json
import os
import sys

import pytest

import wfa_amd
from conftest import GOLDEN
from wfa_amd import datagen

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(GOLDEN, "long_pairs.json")))["sets"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fixture_pairs(name):
    s = FIX[name]
    full = datagen.generate(s["shape"], s["seed"], s["first_n"])
    idx = [p["index"] for p in s["pairs"]]
    return full.subset(idx), s["pairs"]


def check(got, i, want):
    ops = got.ops(i)
    assert int(got.status[i]) == want["status"], (want["index"], got.status[i], want["status"])
    assert int(got.score[i]) == want["score"], (want["index"], got.score[i], want["score"])
    assert hashlib.sha256(ops).hexdigest() == want["sha256"], want["index"]
    assert wfa_amd.rle(ops) == want["cigar"], want["index"]


def test_affine2p_8_to_20_kbp_memory_high_equal_reference():
    ps, want = fixture_pairs("c2_tail")
    say("affine2p 8-20 kbp tail: %d pairs, lengths %s" % (ps.n, [int(x) for x in ps.plen]))
    al = wfa_amd.Aligner(metric="affine2p", penalties=(0, 4, 6, 2, 24, 1))
    got = al.align_batch(ps)
    for i in range(ps.n):
        check(got, i, want[i])


def test_affine2p_8_to_20_kbp_score_only_equal_reference():
    ps, want = fixture_pairs("c2_tail")
    say("affine2p 8-20 kbp tail, score only: %d pairs" % ps.n)
    al = wfa_amd.Aligner(metric="affine2p", penalties=(0, 4, 6, 2, 24, 1), score_only=True)
    got = al.align_batch(ps)
    for i in range(ps.n):
        assert int(got.status[i]) == want[i]["status"]
        assert int(got.score[i]) == want[i]["score"], want[i]["index"]


def test_biwfa_full_100kbp_equal_reference():
    ps, want = fixture_pairs("c3")
    say("BiWFA 100 kbp / 15%%: %d full pairs" % ps.n)
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), memory="ultralow")
    got = al.align_batch(ps)
    for i in range(ps.n):
        check(got, i, want[i])
