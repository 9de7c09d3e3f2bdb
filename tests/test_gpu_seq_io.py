"""The streaming .seq -> .alg entry point (wfa_align_seq_file) on the GPU:
the reference's regression input (tests/golden/wfa.utest.seq) aligned in each
configuration gives, byte for byte, the .alg file the reference's own
align_benchmark wrote for it (tests/wfa.utest.sh:35-64), and the full output
form carries the lengths and sequences."""
import os
import sys

import pytest

import golden_suite as G
import wfa_amd
from conftest import GOLDEN
from wfa_amd import datagen

pytestmark = pytest.mark.gpu

CASES = [c for c in G.cases() if c[0] in (
    "test.affine", "test.affine.p3", "test.score.affine", "test.pb.affine2p", "test.edit", "test.indel",
    "test.biwfa.affine", "test.biwfa.score.affine", "test.affine.wfapt0", "test.score.affine.wfapt1")]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_seq_file_equals_reference_alg(tmp_path, case):
    name, metric, pen, mem, so, heur = case
    say("seq file %s" % name)
    al = wfa_amd.Aligner(metric=metric, penalties=pen, memory=mem, score_only=so, heuristic=heur)
    out = os.path.join(str(tmp_path), name + ".alg")
    rc, st = al.align_seq_file(os.path.join(GOLDEN, "wfa.utest.seq"), out, batch_pairs=100)
    assert rc == 0
    assert st.pairs == 305
    with open(out, "rb") as f, open(os.path.join(GOLDEN, name + ".alg"), "rb") as g:
        assert f.read() == g.read()


def test_seq_file_full_form_and_edges(tmp_path):
    """Full output lines, sequences without a final newline, empty
    sequences, and batches smaller than the file."""
    say("seq file full form")
    ps = datagen.generate(1, 5101, 50)
    pats, txts = ps.patterns() + [b"", b"ACGT"], ps.texts() + [b"ACG", b""]
    src = os.path.join(str(tmp_path), "in.seq")
    with open(src, "wb") as f:
        body = b"".join(b">" + p + b"\n<" + t + b"\n" for p, t in zip(pats, txts))
        f.write(body[:-1])  # no newline after the last text
    al = wfa_amd.Aligner(metric="affine", penalties=(0, 4, 6, 2))
    out = os.path.join(str(tmp_path), "out.alg")
    rc, st = al.align_seq_file(src, out, batch_pairs=7, full=True)
    assert rc == 0 and st.pairs == len(pats)
    res = al.align_batch(datagen.PairSet.from_lists(pats, txts))
    lines = open(out, "rb").read().decode().splitlines()
    assert len(lines) == len(pats)
    for i, line in enumerate(lines):
        plen, tlen, score, p, t, cigar = line.split("\t")
        assert (int(plen), int(tlen), p.encode(), t.encode()) == (len(pats[i]), len(txts[i]), pats[i], txts[i])
        assert cigar == (wfa_amd.rle(res.ops(i)) if res.ops(i) else "-")
