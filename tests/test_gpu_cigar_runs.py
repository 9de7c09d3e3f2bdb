"""Run-length CIGARs made on the device (wfa_device_batch_download_runs):
for every pair the runs decode to exactly the operations that
wfa_device_batch_download returns, in the form cigar_get_CIGAR gives with
show_mismatches, and status / score are the same; BiWFA, affine2p and
score-only batches included; the extension form is refused."""
import sys

import numpy as np
import pytest

import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check_same(al, ps, label):
    db = al.upload(ps)
    assert db.run() == 0
    want = db.download()
    got = db.download_runs()
    db.free()
    assert got.rc == 0
    assert np.array_equal(got.status, want.status)
    assert np.array_equal(got.score, want.score)
    for i in range(ps.n):
        assert got.ops(i) == want.ops(i), (label, i)
        assert got.rle(i) == wfa_amd.rle(want.ops(i)), (label, i)
    say("%s: %d pairs, %d runs for %d operation bytes" % (label, ps.n, len(got.runs),
                                                         int(sum(len(want.ops(i)) for i in range(ps.n)))))


def mixed(seed):
    parts = [datagen.generate(1, seed, 3000), datagen.generate(0, seed + 1, 100), datagen.generate(2, seed + 2, 6)]
    return datagen.PairSet.from_lists(sum((p.patterns() for p in parts), []), sum((p.texts() for p in parts), []))


def test_runs_equal_operations_affine():
    check_same(wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2)), mixed(5150), "affine")


def test_runs_equal_operations_affine2p_and_biwfa():
    check_same(wfa_amd.Aligner(metric="affine2p", penalties=(0, 4, 6, 2, 24, 1)), mixed(5160), "affine2p")
    check_same(wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), memory="ultralow"),
               datagen.generate(0, 5170, 200), "biwfa")


def test_runs_score_only_and_extension_refused():
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), score_only=True)
    ps = datagen.generate(1, 5180, 500)
    db = al.upload(ps)
    assert db.run() == 0
    want = db.download(cigars=False)
    got = db.download_runs()
    db.free()
    assert len(got.runs) == 0 and np.array_equal(got.score, want.score)
    ext = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), extension=True)
    db = ext.upload(ps)
    assert db.run() in (0, 1)
    assert wfa_amd.lib().wfa_device_batch_cigar_runs(ext.wf, db.h) == wfa_amd.WF_STATUS_UNSUPPORTED
    db.free()
