"""A reference-style C caller (tests/c/dropin_example.c) compiled unchanged
against include/wfa_mi355x.h and linked with libwfa2_amd.so: it must build
without a GPU, and on the GPU print the same .alg lines as the oracle."""
import os
import subprocess

import pytest

import oracle as O
import wfa_amd
from conftest import ROOT
from wfa_amd import datagen

SRC = os.path.join(ROOT, "tests", "c", "dropin_example.c")
LIBDIR = os.path.dirname(wfa_amd.LIB_PATH)
README_PAIR = (b"TCTTTACTCGCGCGTTGGAGAAATACAATAGT", b"TCTATACTGCGCGTTTGGAGAAATAAAATAGT")


def build(tmp):
    exe = os.path.join(str(tmp), "dropin_example")
    cmd = ["gcc", "-std=gnu11", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIBDIR, "-lwfa2_amd", "-Wl,-rpath," + LIBDIR, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    return exe


def test_reference_style_caller_compiles_and_links(tmp_path):
    assert os.access(build(tmp_path), os.X_OK)


@pytest.mark.gpu
def test_reference_style_caller_matches_oracle(tmp_path):
    print("C drop-in caller on GPU", flush=True)
    exe = build(tmp_path)
    ps = datagen.generate(1, 909, 200)
    stdin = b"".join(ps.pattern(i) + b"\t" + ps.text(i) + b"\n" for i in range(ps.n))
    out = subprocess.run([exe], input=stdin, capture_output=True, timeout=120, check=True).stdout
    lines = out.decode().splitlines()
    conf = O.Conf("affine", (0, 4, 6, 2), "high")
    pats = [README_PAIR[0]] + ps.patterns()
    txts = [README_PAIR[1]] + ps.texts()
    ref = O.oracle_align(conf, O.Pairs(pats, txts))
    want = ["%d\t%s" % (O.cigar_score(conf, ref.ops(i)), wfa_amd.rle(ref.ops(i))) for i in range(len(pats))]
    assert lines == want
