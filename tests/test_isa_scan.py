"""The device code must write memory only through vector instructions: scan
the gfx950 assembly of the kernels for scalar-cache writes."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

BAD = re.compile(r"^\s+s_(store|atomic|dcache_wb|buffer_store|buffer_atomic|scratch_store)", re.M)


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_no_scalar_cache_writes(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    src = os.path.join(ROOT, "wfa2-lib_amd", "csrc", "wfa_kernels.hip")
    out = tmp_path / "k.s"
    subprocess.check_call([hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                           "-Xclang", "-packed-fp32-ops", "--offload-device-only", "-S",
                           "-I" + os.path.join(ROOT, "include"), src, "-o", str(out)],
                          stderr=subprocess.DEVNULL)
    asm = out.read_text()
    assert "wfa_forward_kernel" in asm
    assert not BAD.search(asm)
