"""A C++ caller of the reference's binding interface (tests/cpp/
wfaligner_example.cpp) compiled against include/WFAligner.hpp: it must build
without a GPU; on the GPU each aligner / form matches the oracle, the batch
overloads match the single-pair calls, and the SAM CIGARs follow
alignment/cigar.c:180-239 (run lengths, 'X' folded into 'M' unless
mismatches are shown, then matches as '=')."""
import os
import subprocess

import pytest

import oracle as O
import wfa_amd
from conftest import ROOT
from wfa_amd import datagen

SRC = os.path.join(ROOT, "tests", "cpp", "wfaligner_example.cpp")
LIBDIR = os.path.dirname(wfa_amd.LIB_PATH)


def build(tmp):
    exe = os.path.join(str(tmp), "wfaligner_example")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIBDIR, "-lwfa2_amd", "-Wl,-rpath," + LIBDIR, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=180)
    return exe


def sam_cigar(ops: bytes, show_mismatches: bool) -> str:
    """Run-length SAM CIGAR from its definition; the first operation is taken
    as it is, as the reference does."""
    if not ops:
        return ""
    runs = []
    for i, c in enumerate(ops.decode()):
        op = c if (i == 0 or show_mismatches or c != "X") else "M"
        if runs and runs[-1][0] == op:
            runs[-1][1] += 1
        else:
            runs.append([op, 1])
    return "".join("%d%s" % (n, "=" if (show_mismatches and op == "M") else op) for op, n in runs)


def test_cpp_caller_compiles_and_links(tmp_path):
    assert os.access(build(tmp_path), os.X_OK)


MODES = {
    "affine": O.Conf("affine", (0, 4, 6, 2), heuristic=(4, 10, 50, 1)),   # the default attributes
    "affine-exact": O.Conf("affine", (0, 4, 6, 2)),
    "endsfree": O.Conf("affine", (0, 4, 6, 2), ends_free=(0, 0, 10, 10)),
    "extension": O.Conf("affine", (0, 4, 6, 2), extension=True),
    "edit": O.Conf("edit", (0, 1, 1)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("mode", list(MODES))
def test_cpp_caller_matches_oracle(tmp_path, mode):
    print("C++ WFAligner %s on GPU" % mode, flush=True)
    exe = build(tmp_path)
    ps = datagen.semiglobal(1, 919, 150, flank=10)
    stdin = b"".join(ps.pattern(i) + b"\t" + ps.text(i) + b"\n" for i in range(ps.n))
    single = subprocess.run([exe, mode], input=stdin, capture_output=True, timeout=120, check=True).stdout
    batch = subprocess.run([exe, mode, "batch"], input=stdin, capture_output=True, timeout=120, check=True).stdout
    ref = O.oracle_align(MODES[mode], O.Pairs(ps.patterns(), ps.texts()))
    s_lines, b_lines = single.decode().splitlines(), batch.decode().splitlines()
    assert len(s_lines) == len(b_lines) == ps.n
    for i in range(ps.n):
        st, sc, sam, sam_x, ops = s_lines[i].split("\t")
        want_ops = ref.ops(i)
        assert (int(st), int(sc), ops) == (int(ref.status[i]), int(ref.score[i]), want_ops.decode()), i
        assert sam == sam_cigar(want_ops, False) and sam_x == sam_cigar(want_ops, True), i
        bst, bsc, _, _, bops = b_lines[i].split("\t")
        assert (bst, bsc, bops) == (st, sc, ops), i
