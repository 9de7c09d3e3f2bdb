"""Builds the reference's example callers (examples/wfa_basic.c,
examples/wfa_repeated.c), unchanged, against this repository's forwarding
header include/wavefront/wavefront_align.h and links them with
libwfa2_amd.so.  Used by __graft_entry__.build() (output in tests/c/_bin/,
which travels with the tree to a GPU run) and by
tests/test_dropin_reference_examples.py."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REFERENCE = "/root/reference"
EXAMPLES = ("wfa_basic", "wfa_repeated")
BIN = os.path.join(ROOT, "tests", "c", "_bin")
LIBDIR = os.path.join(ROOT, "wfa2-lib_amd", "wfa_amd")


def compile_example(name, out_dir):
    """gcc on the unchanged example source; the include path holds include/
    (wavefront/wavefront_align.h) and a temporary directory outside the tree
    whose utils/ links to the reference's utils/ (the examples include
    utils/commons.h for libc headers).  Only the executable lands in out_dir."""
    exe = os.path.join(str(out_dir), name)
    with tempfile.TemporaryDirectory(prefix="wfa_ref_inc_") as inc:
        os.symlink(os.path.join(REFERENCE, "utils"), os.path.join(inc, "utils"))
        cmd = ["gcc", "-std=gnu11", "-O2", "-I", os.path.join(ROOT, "include"), "-I", inc,
               os.path.join(REFERENCE, "examples", name + ".c"),
               "-L", LIBDIR, "-lwfa2_amd", "-Wl,-rpath," + LIBDIR, "-o", exe]
        subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    return exe
