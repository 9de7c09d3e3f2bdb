"""The request loop (wfa_server_*) on the GPU: requests of three length
classes arrive over time (Poisson), are batched by free lanes while other
batches run, and every result equals the oracle's and the result of the same
request aligned alone with wavefront_align."""
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


def mixed_requests():
    short = datagen.generate(1, 1201, 240)                    # 100-300 bp
    medium = datagen.generate(0, 1202, 40)                     # 1 kbp
    longs = datagen.generate(2, 1203, 40)
    keep = [i for i in range(longs.n) if 2500 <= max(longs.plen[i], longs.tlen[i]) <= 6000][:8]
    longs = longs.subset(keep)
    pats = short.patterns() + medium.patterns() + longs.patterns()
    txts = short.texts() + medium.texts() + longs.texts()
    rng = np.random.default_rng(12)
    order = rng.permutation(len(pats))
    return [pats[i] for i in order], [txts[i] for i in order]


def collect_all(srv, n, timeout_s=120):
    got = {}
    t0 = srv.clock()
    while len(got) < n:
        for r in srv.collect(max_results=1024, timeout_ms=200):
            got[r.id] = r
        assert srv.clock() - t0 < timeout_s, "requests did not finish"
    return got


def test_served_results_equal_oracle_and_alone():
    say("serving: Poisson mix of short / medium / long requests")
    pats, txts = mixed_requests()
    n = len(pats)
    srv = wfa_amd.Server(short_max_length=400, long_min_length=2500, short_lanes=2, medium_lanes=1,
                         long_lanes=1, max_batch_pairs=64, lane_memory_bytes=2 << 30)
    arrivals = np.cumsum(np.random.default_rng(3).exponential(1.0 / 400.0, n))  # ~400 requests/s
    srv.replay(datagen.PairSet.from_lists(pats, txts), arrivals)
    got = collect_all(srv, n)
    srv.close()
    assert sorted(got) == list(range(n))  # ids are submission order
    # oracle, per class (long requests use BiWFA)
    is_long = np.array([max(len(p), len(t)) >= 2500 for p, t in zip(pats, txts)])
    ref = {}
    for mem, sel in (("high", ~is_long), ("ultralow", is_long)):
        idx = np.nonzero(sel)[0]
        r = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), mem), O.Pairs([pats[i] for i in idx], [txts[i] for i in idx]))
        for j, i in enumerate(idx):
            ref[i] = (int(r.status[j]), int(r.score[j]), r.ops(j))
    for i in range(n):
        assert (got[i].status, got[i].score, got[i].ops) == ref[i], i
    # alone: the same request through wavefront_align
    alone = {"high": wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2)),
             "ultralow": wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), memory="ultralow")}
    for i in range(n):
        st, sc, ops = alone["ultralow" if is_long[i] else "high"].align(pats[i], txts[i])
        assert (st, sc, ops) == (got[i].status, got[i].score, got[i].ops), i
    # batching happened, and batches of different lanes overlapped in time
    starts = {}
    for r in got.values():
        starts.setdefault((r.start, r.finish), []).append(r.id)
    assert max(len(v) for v in starts.values()) > 1
    spans = sorted(starts)
    overlap = any(spans[j][0] < spans[j - 1][1] for j in range(1, len(spans)))
    assert overlap
    lat = np.array([got[i].latency for i in range(n)])
    say("served %d requests, p50 %.2f ms, p99 %.2f ms, %d batches" % (
        n, np.percentile(lat, 50) * 1e3, np.percentile(lat, 99) * 1e3, len(spans)))


def test_submit_and_collect_by_id():
    say("serving: submit / collect")
    srv = wfa_amd.Server(short_lanes=1, medium_lanes=1, long_lanes=1)
    ps = datagen.generate(1, 1301, 20)
    ids = [srv.submit(ps.pattern(i), ps.text(i)) for i in range(ps.n)]
    got = collect_all(srv, ps.n)
    assert srv.pending() == 0
    srv.close()
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    for i, rid in enumerate(ids):
        st, sc, ops = al.align(ps.pattern(i), ps.text(i))
        assert (got[rid].status, got[rid].score, got[rid].ops) == (st, sc, ops)
        assert got[rid].submit <= got[rid].start <= got[rid].finish


def test_serving_mix_of_the_configuration_equal_reference():
    """The serving configuration's own mixture (60% 1 kbp / 5%, 35% 1-20 kbp /
    1-15% with geometric indels, 5% 100 kbp / 15% through BiWFA), default
    length classes, Poisson arrivals: every served result equals the
    reference's (tests/golden/long_pairs.json) and the same request aligned
    alone."""
    import hashlib
    import json
    import os
    from conftest import GOLDEN
    s = json.load(open(os.path.join(GOLDEN, "long_pairs.json")))["sets"]["c4_mix"]
    ps = datagen.generate(s["shape"], s["seed"], s["first_n"])
    want = s["pairs"]
    n = ps.n
    is_long = np.array([max(int(ps.plen[i]), int(ps.tlen[i])) >= s["long_min_length"] for i in range(n)])
    say("serving: configuration mix, %d requests (%d of 100 kbp)" % (n, int(is_long.sum())))
    srv = wfa_amd.Server(long_min_length=s["long_min_length"], lane_memory_bytes=8 << 30)
    arrivals = np.cumsum(np.random.default_rng(5).exponential(1.0 / 50.0, n))  # ~50 requests/s
    srv.replay(ps, arrivals)
    got = collect_all(srv, n, timeout_s=110)
    srv.close()
    assert sorted(got) == list(range(n))
    for i in range(n):
        w = want[i]
        assert (got[i].status, got[i].score) == (w["status"], w["score"]), i
        assert hashlib.sha256(got[i].ops).hexdigest() == w["sha256"], i
    alone = {"high": wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2)),
             "ultralow": wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2), memory="ultralow")}
    for i in range(n):
        st, sc, ops = alone["ultralow" if is_long[i] else "high"].align(ps.pattern(i), ps.text(i))
        assert (st, sc, ops) == (got[i].status, got[i].score, got[i].ops), i
    lat = np.array([got[i].latency for i in range(n)]) * 1e3
    say("served %d requests: p50 %.1f ms, p99 %.1f ms (short p50 %.1f ms, long p50 %.1f ms)" % (
        n, np.percentile(lat, 50), np.percentile(lat, 99), np.percentile(lat[~is_long], 50),
        np.percentile(lat[is_long], 50)))
