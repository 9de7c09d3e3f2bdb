"""The device code must write memory only through vector instructions: scan
the gfx950 assembly of every HIP source for scalar-cache writes, and check
that the kernels each source is meant to hold were really compiled into it.
The backtrace kernel's CIGAR writers must keep their 16-byte chunks in
registers (no LDS, no scratch)."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

BAD = re.compile(r"^\s+s_(store|atomic|dcache_wb|buffer_store|buffer_atomic|scratch_store)", re.M)
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CSRC = os.path.join(ROOT, "wfa2-lib_amd", "csrc")

# source -> kernel names that must appear in its assembly
SOURCES = {
    "wfa_kernels.hip": ["wfa_forward_kernel", "wfa_backtrace_kernel", "wfa_reset_kernel"],
    "wfa_lean.hip": ["wfa_forward_lean_kernel"],
    "wfa_biwfa.hip": ["wfa_breakpoint_kernel", "reverse_pack_kernel", "gather_ops_kernel"],
}


_ASM = {}


def assembly(name, tmp_path_factory):
    if name not in _ASM:
        out = tmp_path_factory.mktemp("isa") / (name + ".s")
        subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                               "-Xclang", "-packed-fp32-ops", "--offload-device-only", "-S",
                               "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, os.path.join(CSRC, name),
                               "-o", str(out)],
                              stderr=subprocess.DEVNULL)
        _ASM[name] = out.read_text()
    return _ASM[name]


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
@pytest.mark.parametrize("name", sorted(SOURCES))
def test_no_scalar_cache_writes(tmp_path_factory, name):
    asm = assembly(name, tmp_path_factory)
    for kernel in SOURCES[name]:
        assert kernel in asm, "%s: no %s in the assembly" % (name, kernel)
    m = BAD.search(asm)
    assert m is None, "%s: scalar-cache write: %s" % (name, m.group(0) if m else "")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_backtrace_writers_stay_in_registers(tmp_path_factory):
    asm = assembly("wfa_kernels.hip", tmp_path_factory)
    blocks = re.findall(r"\.amdhsa_kernel (\S*wfa_backtrace_kernel\S*)\n(.*?)\.end_amdhsa_kernel", asm, re.S)
    assert len(blocks) == 2, [b[0] for b in blocks]  # 2-bit and 8-bit sequences
    for kname, body in blocks:
        lds = int(re.search(r"\.amdhsa_group_segment_fixed_size (\d+)", body).group(1))
        priv = int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", body).group(1))
        assert (lds, priv) == (0, 0), (kname, lds, priv)
