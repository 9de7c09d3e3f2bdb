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
    dict(metric="edit", ends_free=(0, 0, 2, 2)),                       # free beginning, edit
    dict(metric="affine", penalties=(-1, 4, 6, 2), ends_free=(2, 0, 0, 0)),  # free beginning, match < 0
    dict(metric="affine", memory="ultralow", ends_free=(0, 0, 0, 0)),  # ends-free span in BiWFA
    dict(metric="edit", memory="med", extension=True),                 # indel/edit extension outside high
    dict(metric="affine", memory="med", heuristic=(0x20, 0, 0, 1, 0, 0, 0, 30)),   # z-drop, CIGAR, med
    dict(metric="affine", memory="low", heuristic=(0x24, 10, 50, 1, 0, 0, 0, 30)),  # wf-adaptive + z-drop, low
])
def test_unsupported_configurations_are_refused(kw):
    al = wfa_amd.Aligner(**kw)
    st, score, ops = al.align(b"ACGT", b"ACTT")
    assert st == wfa_amd.WF_STATUS_UNSUPPORTED
    assert ops == b""


@pytest.mark.parametrize("shares", [(1.0, 1.0, 1.0), (0.5, 0.625, 0.5), (0.0625, 0.375, 0.5), (0.125, 0.375, 1.5)])
def test_server_refuses_cu_shares_that_do_not_split_the_device(shares):
    """Short and medium take disjoint CU ranges of at least one die's share
    each; wfa_server_create refuses other splits before starting any lane."""
    kw = dict(zip(("short_cu_share", "medium_cu_share", "long_cu_share"), shares))
    with pytest.raises(RuntimeError, match="wfa_server_create failed"):
        wfa_amd.Server(**kw)


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


def test_cigar_scores_of_every_metric_match_oracle():
    """cigar_score_edit / _gap_linear / _gap_affine / _gap_affine2p of the C
    ABI (alignment/cigar.c:244-342) against the oracle's scorer on random
    operation strings."""
    import ctypes
    import random

    import oracle as O
    L = wfa_amd.lib()
    rng = random.Random(7)
    for _ in range(200):
        # at least one operation: the affine2p scorer exits on an empty CIGAR, as the reference
        ops = bytes(rng.choice(b"MMMXID") for _ in range(rng.randint(1, 40)))
        buf = ctypes.create_string_buffer(ops, len(ops) + 1)
        c = wfa_amd.cigar_t()
        c.operations = ctypes.cast(buf, ctypes.c_void_p)
        c.begin_offset, c.end_offset = 0, len(ops)
        ref = ctypes.byref(c)
        assert L.cigar_score_edit(ref) == O.cigar_score(O.Conf("edit", (0, 1, 1)), ops)
        lp = wfa_amd.linear_penalties_t(-1, 3, 2)
        assert L.cigar_score_gap_linear(ref, ctypes.byref(lp)) == O.cigar_score(O.Conf("linear", (-1, 3, 2)), ops)
        ap = wfa_amd.affine_penalties_t(0, 4, 6, 2)
        assert L.cigar_score_gap_affine(ref, ctypes.byref(ap)) == O.cigar_score(O.Conf("affine", (0, 4, 6, 2)), ops)
        a2 = wfa_amd.affine2p_penalties_t(0, 4, 6, 2, 24, 1)
        assert L.cigar_score_gap_affine2p(ref, ctypes.byref(a2)) == O.cigar_score(
            O.Conf("affine2p", (0, 4, 6, 2, 24, 1)), ops)


EXIT_CHILD = r"""
import sys
sys.path.insert(0, %r)
import wfa_amd
al = wfa_amd.Aligner(**%r)
al.align(b"ACGTACGT", b"ACGTTACGT")
print("returned")
"""


@pytest.mark.parametrize("kw,msg", [
    (dict(metric="affine", memory="ultralow", ends_free=(0, 2, 0, 2)), b"BiWFA ends-free has not been tested"),
    (dict(metric="affine", memory="ultralow", extension=True), b"BiWFA ends-free has not been tested"),
    (dict(metric="affine", ends_free=(0, 0, 20, 0)), b"Ends-free parameters must be not larger"),
    (dict(metric="edit", heuristic=(0x10, 0, 0, 1, 0, 0, 20)), b"Heuristics drops are not compatible"),
    (dict(metric="indel", heuristic=(0x20, 0, 0, 1)), b"Heuristics drops are not compatible"),
])
def test_configuration_errors_exit_like_the_reference(kw, msg):
    """wavefront_align_presets__checks (wavefront_align.c:46-106) ends the
    process with exit(1) and a message; so does the drop-in, before any
    device work."""
    import subprocess
    import sys
    code = EXIT_CHILD % (os.path.join(ROOT, "wfa2-lib_amd"), kw)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, timeout=120)
    assert r.returncode == 1 and b"returned" not in r.stdout
    assert msg in r.stderr


def test_heuristic_setters_accumulate_like_the_reference():
    """wavefront_aligner_set_heuristic_* add their strategy to those already
    set and overwrite their own parameters (wavefront_aligner.c:534-600,
    wavefront_heuristic.c:40-113); set_none clears the strategy only."""
    import ctypes
    L = wfa_amd.lib()
    attr = wfa_amd.default_attributes()
    wf = L.wavefront_aligner_new(ctypes.byref(attr))
    try:
        h = wf.contents.heuristic
        assert h.strategy == 4
        L.wavefront_aligner_set_heuristic_banded_static(wf, -7, 9)
        h = wf.contents.heuristic
        assert (h.strategy, h.min_k, h.max_k, h.min_wavefront_length) == (5, -7, 9, 10)
        L.wavefront_aligner_set_heuristic_banded_adaptive(wf, -3, 3, 7)
        h = wf.contents.heuristic
        assert (h.strategy, h.min_k, h.max_k, h.steps_between_cutoffs) == (7, -3, 3, 7)
        L.wavefront_aligner_set_heuristic_none(wf)
        h = wf.contents.heuristic
        assert (h.strategy, h.min_k, h.max_k) == (0, -3, 3)
    finally:
        L.wavefront_aligner_delete(wf)
