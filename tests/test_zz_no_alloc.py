"""wfa_device_batch_run_noalloc through the C ABI against wfa_device_batch_run
on the same resident batches, two calls in a row (the second on a smaller
batch): status, score, every CIGAR operation and the diagonal-step totals
must be identical.  The first batch's records exceed the static array's
record arena (so it runs in waves) and holds pairs of up to 6 kbp (the
HBM-ring family, whose grid the array's ring share limits).  The device's
free memory is the same before and after a call (no allocation), and memory
mode ultralow is refused before any device work."""
import ctypes
import sys

import numpy as np
import pytest

import wfa_amd
from wfa_amd import datagen

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def mixed(seed, n_short, n_kbp, n_long):
    parts = [datagen.generate(1, seed, n_short), datagen.generate(0, seed + 1, n_kbp)]
    longs = datagen.generate(2, seed + 2, 40)
    keep = [i for i in range(longs.n) if max(longs.plen[i], longs.tlen[i]) <= 6000][:n_long]
    parts.append(longs.subset(keep))
    pats = sum((p.patterns() for p in parts), [])
    txts = sum((p.texts() for p in parts), [])
    return datagen.PairSet.from_lists(pats, txts)


def compare(al, ps, label):
    db = al.upload(ps)
    rc = db.run()
    want = db.download()
    want_steps = db.stats().diagonal_steps
    rc2 = db.run_noalloc()
    got = db.download()
    got_steps = db.stats().diagonal_steps
    db.free()
    assert rc == rc2 == 0, (label, rc, rc2)
    assert got_steps == want_steps, label
    assert np.array_equal(got.status, want.status), label
    assert np.array_equal(got.score, want.score), label
    for i in range(ps.n):
        assert got.ops(i) == want.ops(i), (label, i)
    say("%s: %d pairs identical (%d diagonal-steps)" % (label, ps.n, got_steps))


def test_noalloc_equals_batch_run_two_calls():
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    big = mixed(2024, 150000, 300, 12)      # ~330 MB of records: several arena waves
    say("no-alloc run, call 1: %d pairs" % big.n)
    compare(al, big, "call 1")
    small = mixed(2030, 2000, 40, 4)
    say("no-alloc run, call 2: %d pairs" % small.n)
    compare(al, small, "call 2")


def test_noalloc_affine2p_score_only_and_cigar():
    for score_only in (False, True):
        al = wfa_amd.Aligner(metric="affine2p", penalties=(0, 4, 6, 2, 24, 1), score_only=score_only)
        ps = mixed(2040, 3000, 50, 6)
        say("no-alloc run, affine2p, score_only=%s: %d pairs" % (score_only, ps.n))
        db = al.upload(ps)
        assert db.run() == 0
        want = db.download(cigars=not score_only)
        assert db.run_noalloc() == 0
        got = db.download(cigars=not score_only)
        db.free()
        assert np.array_equal(got.status, want.status)
        assert np.array_equal(got.score, want.score)
        if not score_only:
            for i in range(ps.n):
                assert got.ops(i) == want.ops(i), i


def test_noalloc_allocates_nothing_and_refuses_ultralow():
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    ps = mixed(2050, 5000, 20, 2)
    db = al.upload(ps)
    assert db.run_noalloc() == 0       # first call on the device: looks the array up
    before = free_device_bytes()
    assert db.run_noalloc() == 0
    after = free_device_bytes()
    db.free()
    say("free device bytes before / after a no-alloc call: %d / %d" % (before, after))
    assert after == before
    bi = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), memory="ultralow")
    db = bi.upload(datagen.generate(1, 2060, 10))
    assert db.run_noalloc() == wfa_amd.WF_STATUS_UNSUPPORTED
    db.free()
