"""Host-side checks of the C ABI that need no GPU: the library loads, exports
every function include/wfa_mi355x.h declares, keeps the reference's default
attributes (wavefront_attributes.c:38-100), and refuses configurations outside
the exact device path before touching the device."""
import os

import pytest

import wfa_amd
from conftest import ROOT
from wfa_amd import datagen


def test_library_exports_every_declared_function():
    L = wfa_amd.lib()
    names = wfa_amd.declared_functions(os.path.join(ROOT, "include", "wfa_mi355x.h"))
    assert "wavefront_align" in names and "wavefront_align_batch" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_default_attributes_match_reference():
    a = wfa_amd.default_attributes()
    assert a.distance_metric == 3 and a.alignment_scope == 1
    assert (a.affine_penalties.match, a.affine_penalties.mismatch, a.affine_penalties.gap_opening,
            a.affine_penalties.gap_extension) == (0, 4, 6, 2)
    p2 = a.affine2p_penalties
    assert (p2.match, p2.mismatch, p2.gap_opening1, p2.gap_extension1, p2.gap_opening2,
            p2.gap_extension2) == (0, 4, 6, 2, 24, 1)
    assert (a.linear_penalties.match, a.linear_penalties.mismatch, a.linear_penalties.indel) == (0, 4, 2)
    assert a.heuristic.strategy == 4 and a.heuristic.min_wavefront_length == 10
    assert a.heuristic.max_distance_threshold == 50 and a.heuristic.steps_between_cutoffs == 1
    assert a.memory_mode == 0
    assert a.system.max_alignment_steps == 2**31 - 1 and a.system.probe_interval_global == 3000
    assert a.system.max_memory_abort == 2**64 - 1 and a.system.max_num_threads == 1
    assert a.system.min_offsets_per_thread == 500 and a.plot.resolution_points == 2000


@pytest.mark.parametrize("kw", [
    dict(metric="edit"),
    dict(metric="affine", heuristic_none=False),
])
def test_unsupported_configurations_are_refused(kw):
    al = wfa_amd.Aligner(**kw)
    st, score, ops = al.align(b"ACGT", b"ACTT")
    assert st == wfa_amd.WF_STATUS_UNSUPPORTED
    assert ops == b""


def test_strerror_messages():
    L = wfa_amd.lib()
    assert L.wavefront_align_strerror(0) == b"[WFA] Alignment completed successfully"
    assert L.wavefront_align_strerror_short(-100) == b"FAILED.MaxSteps"


def test_generator_is_seeded_and_independent_per_pair():
    a = datagen.generate(1, 5, 50)
    b = datagen.generate(1, 5, 20, first=30)
    assert a.pattern(30) == b.pattern(0) and a.text(49) == b.text(19)
    c = datagen.generate(1, 6, 50)
    assert a.text(0) != c.text(0)
    assert all(100 <= a.tlen[i] <= 300 for i in range(a.n))
    assert set(a.buf) <= set(b"ACGT")
