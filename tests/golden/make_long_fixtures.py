"""Writes tests/golden/long_pairs.json: the reference's own results (WFA2-lib
built from its sources by oracle/Makefile into oracle/_ref/) for long pairs
of the seeded workloads, which the CPU restatement is too slow to check on
every test run.  Run in the build container, where the reference sources are:

    python tests/golden/make_long_fixtures.py

Sets (pairs regenerate from (shape, seed, index) with wfa_amd.datagen):
  c2_tail  dual-cost gap-affine (0,4,6,2,24,1), memory mode med (CIGARs
           identical to memory high, tests/wfa.utest.sh:67-78 of the
           reference), pairs of 8-20 kbp from the 1-20 kbp workload;
  c3       gap-affine (0,4,6,2), memory mode ultralow (BiWFA), full 100 kbp /
           15% pairs;
  c4_mix   gap-affine (0,4,6,2) serving mixture (60% 1 kbp / 5%, 35% 1-20
           kbp, 5% 100 kbp); memory mode high (as med) for pairs under 32 kbp
           and ultralow for the 100 kbp ones, as the request loop's classes.
Each pair keeps the status, the score, the CIGAR in run-length form and the
SHA-256 of its operations."""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "wfa2-lib_amd"))

import oracle as O  # noqa: E402
from wfa_amd import datagen  # noqa: E402

OUT = os.path.join(HERE, "long_pairs.json")
LONG_MIN = 32768


def record(idx, ps, res, j, memory):
    ops = res.ops(j)
    return {"index": int(idx), "plen": int(ps.plen[j]), "tlen": int(ps.tlen[j]), "memory": memory,
            "status": int(res.status[j]), "score": int(res.score[j]), "cigar": O.rle(ops),
            "sha256": hashlib.sha256(ops).hexdigest()}


def run(conf, ps, threads):
    pairs = O.Pairs(ps.patterns(), ps.texts())
    t = time.time()
    r = O.ref_align(conf, pairs, threads)
    print("  %d pairs, %s %s: %.1f s" % (ps.n, conf.metric, conf.memory, time.time() - t), flush=True)
    return r


def main():
    threads = min(8, os.cpu_count() or 1)
    sets = {}

    # c2_tail: the first pairs of 8-20 kbp of the 1-20 kbp workload (seed 7)
    full = datagen.generate(2, 7, 200)
    L = [max(int(full.plen[i]), int(full.tlen[i])) for i in range(full.n)]
    idx = [i for i in range(full.n) if L[i] >= 8000][:8]
    ps = full.subset(idx)
    conf = O.Conf("affine2p", (0, 4, 6, 2, 24, 1), "med")
    r = run(conf, ps, threads)
    sets["c2_tail"] = {"shape": 2, "seed": 7, "first_n": 200, "metric": "affine2p",
                       "penalties": [0, 4, 6, 2, 24, 1],
                       "pairs": [record(i, ps, r, j, "med") for j, i in enumerate(idx)]}

    # c3: full 100 kbp / 15% pairs through BiWFA (seed 7, the first three)
    ps = datagen.generate(3, 7, 3)
    conf = O.Conf("affine", (0, 4, 6, 2), "ultralow")
    r = run(conf, ps, threads)
    sets["c3"] = {"shape": 3, "seed": 7, "first_n": 3, "metric": "affine", "penalties": [0, 4, 6, 2],
                  "pairs": [record(i, ps, r, i, "ultralow") for i in range(ps.n)]}

    # c4_mix: the serving mixture, memory mode by length class
    n = 60
    ps = datagen.generate(4, 4242, n)
    longs = [i for i in range(n) if max(int(ps.plen[i]), int(ps.tlen[i])) >= LONG_MIN]
    rest = [i for i in range(n) if i not in longs]
    recs = {}
    for mem, sel in (("med", rest), ("ultralow", longs)):
        sub = ps.subset(sel)
        r = run(O.Conf("affine", (0, 4, 6, 2), mem), sub, threads)
        for j, i in enumerate(sel):
            recs[i] = record(i, sub, r, j, mem)
    sets["c4_mix"] = {"shape": 4, "seed": 4242, "first_n": n, "metric": "affine", "penalties": [0, 4, 6, 2],
                      "long_min_length": LONG_MIN, "pairs": [recs[i] for i in range(n)]}

    meta = {"reference": "WFA2-lib v2.3 built from source by oracle/Makefile (%s)"
                         % os.path.basename(O.ref_lib_path()),
            "generator": "wfa_amd.datagen.generate(shape, seed, first_n); pair i of a set is pair 'index'",
            "sets": sets}
    with open(OUT, "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote %s (%d bytes)" % (OUT, os.path.getsize(OUT)))


if __name__ == "__main__":
    main()
