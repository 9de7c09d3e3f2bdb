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




WAVES_CHILD = r"""
import hashlib, json, sys
import numpy as np
sys.path.insert(0, %r)
import wfa_amd
from wfa_amd import datagen
# twelve pairs with texts of 17,000 bases and 7%% errors (substitutions,
# 1-base insertions and deletions), dual-cost gap-affine, memory mode high
rng = np.random.default_rng(17)
pats, txts = [], []
for _ in range(12):
    t = rng.choice(list(b"ACGT"), 17000).astype(np.uint8)
    p = []
    for c in t:
        r = rng.random()
        if r < 0.07 / 3:
            p.append(int(rng.choice(list(b"ACGT"))))
        elif r < 0.14 / 3:
            p.append(int(c)); p.append(int(rng.choice(list(b"ACGT"))))
        elif r < 0.07:
            continue
        else:
            p.append(int(c))
    pats.append(bytes(p)); txts.append(bytes(t))
ps = datagen.PairSet.from_lists(pats, txts)
al = wfa_amd.Aligner(metric="affine2p", penalties=(0, 4, 6, 2, 24, 1))
db = al.upload(ps)
rc = db.run()
got = db.download()
print(json.dumps({"rc": rc, "launches": db.stats().forward_launches, "dsteps": db.stats().diagonal_steps,
                  "res": [[int(got.status[i]), int(got.score[i]), hashlib.sha256(got.ops(i)).hexdigest()]
                          for i in range(ps.n)]}))
"""


def test_long_pairs_in_record_arena_waves_agree():
    """Twelve long pairs whose records (about 560 MB together, under 64 MB
    each) do not fit a 64 MB record arena at once: run together they fill it
    before any finishes, so they are re-run one workgroup at a time, in
    waves; the results equal those of one launch with a large arena."""
    import subprocess
    from conftest import ROOT
    code = WAVES_CHILD % os.path.join(ROOT, "wfa2-lib_amd")
    out = {}
    for name, env in (("one launch", {}), ("arena waves", {"WFA_RECORD_ARENA_MB": "64"})):
        say("long pairs: %s" % name)
        r = subprocess.run([sys.executable, "-c", code], env=dict(os.environ, **env), capture_output=True,
                           timeout=240)
        assert r.returncode == 0, (name, r.stderr.decode()[-2000:])
        out[name] = json.loads(r.stdout.decode().strip().splitlines()[-1])
        say("  %s: %d launches, %d diagonal-steps" % (name, out[name]["launches"], out[name]["dsteps"]))
        assert out[name]["rc"] == 0, (name, out[name])
    assert out["arena waves"]["launches"] > out["one launch"]["launches"]
    assert out["arena waves"]["res"] == out["one launch"]["res"]
    assert out["arena waves"]["dsteps"] == out["one launch"]["dsteps"]
    assert all(r[0] == 0 for r in out["one launch"]["res"])
