"""The reference's own example callers, examples/wfa_basic.c and
examples/wfa_repeated.c, compiled unchanged against the forwarding header
include/wavefront/wavefront_align.h (plus the reference's utils/commons.h,
which they include for libc headers only) and linked with libwfa2_amd.so.

On the CPU they must compile and link; __graft_entry__.build() also builds
them into tests/c/_bin/ where the reference sources are present, so that a
GPU run can execute them: wfa_basic must report the oracle's score for the
README pair and its CIGAR must re-score to it; wfa_repeated must finish its
100,000 wavefront_align calls."""
import os
import re
import subprocess
import sys

import pytest

import oracle as O
import wfa_amd
from conftest import ROOT
sys.path.insert(0, os.path.join(ROOT, "tests", "c"))
from reference_examples import BIN, EXAMPLES, REFERENCE, compile_example  # noqa: E402

README_PAIR = (b"TCTTTACTCGCGCGTTGGAGAAATACAATAGT", b"TCTATACTGCGCGTTTGGAGAAATAAAATAGT")


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "examples")), reason="reference sources absent")
@pytest.mark.parametrize("name", EXAMPLES)
def test_reference_example_compiles_and_links(tmp_path, name):
    exe = compile_example(name, tmp_path)
    assert os.access(exe, os.X_OK)
    # the forwarding header, not the reference's own wavefront/ directory, was used
    syms = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True).stdout
    assert "wavefront_aligner_new" in syms and "wavefront_align" in syms


@pytest.mark.gpu
def test_reference_wfa_basic_on_gpu():
    exe = os.path.join(BIN, "wfa_basic")
    if not os.access(exe, os.X_OK):
        pytest.skip("tests/c/_bin/wfa_basic not built (needs the reference sources at build time)")
    print("reference examples/wfa_basic.c on GPU", flush=True)
    run = subprocess.run([exe], capture_output=True, timeout=120)
    assert run.returncode == 0, run.stderr.decode()
    err = run.stderr.decode()
    conf = O.Conf("affine", (0, 4, 6, 2), "high")
    ref = O.oracle_align(conf, O.Pairs([README_PAIR[0]], [README_PAIR[1]]))
    want = O.cigar_score(conf, ref.ops(0))
    got = int(re.search(r"WFA-Alignment returns score (-?\d+)", err).group(1))
    rescored = int(re.search(r"SCORE \(RE\)COMPUTED (-?\d+)", err).group(1))
    assert got == want and rescored == want, err
    assert "ALIGNMENT " + wfa_amd.rle(ref.ops(0)) in err, err


@pytest.mark.gpu
def test_reference_wfa_repeated_on_gpu():
    exe = os.path.join(BIN, "wfa_repeated")
    if not os.access(exe, os.X_OK):
        pytest.skip("tests/c/_bin/wfa_repeated not built (needs the reference sources at build time)")
    print("reference examples/wfa_repeated.c on GPU (100,000 wavefront_align calls)", flush=True)
    run = subprocess.run([exe], capture_output=True, timeout=110)
    assert run.returncode == 0, run.stderr.decode()[-2000:]
    err = run.stderr.decode()
    assert "[WFA]" not in err, err[-2000:]
    assert err.strip().splitlines()[-1] == "... done 100000 alignments"
