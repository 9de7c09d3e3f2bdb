"""Python host bindings of libwfa2_amd.so (the MI355X exact-WFA C ABI).

This is a thin ctypes mirror of ``include/wfa_mi355x.h`` used by the tests,
the benchmark and the serving loop.  It loads only the in-tree library and
fails loudly when it is missing: there is no CPU fallback.

    al = Aligner(metric="affine", penalties=(0, 4, 6, 2))
    status, score, ops = al.align(b"ACGT", b"ACTT")      # wavefront_align
    res = al.align_batch(patterns, texts)                 # wavefront_align_batch
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_bool, c_char_p, c_double, c_int, c_int32, c_int64,
                    c_uint, c_uint32, c_uint64, c_void_p)

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwfa2_amd.so")

# wavefront/wfa.h:45-55 (+ WF_STATUS_UNSUPPORTED, library-specific)
WF_STATUS_ALG_COMPLETED = 0
WF_STATUS_ALG_PARTIAL = 1
WF_STATUS_MAX_STEPS_REACHED = -100
WF_STATUS_OOM = -200
WF_STATUS_UNATTAINABLE = -300
WF_STATUS_UNSUPPORTED = -400

METRICS = {"indel": 0, "edit": 1, "linear": 2, "affine": 3, "affine2p": 4}
MEMORY = {"high": 0, "med": 1, "low": 2, "ultralow": 3}


class alignment_form_t(Structure):
    _fields_ = [("span", c_int), ("extension", c_bool), ("pattern_begin_free", c_int),
                ("pattern_end_free", c_int), ("text_begin_free", c_int), ("text_end_free", c_int)]


class linear_penalties_t(Structure):
    _fields_ = [("match", c_int), ("mismatch", c_int), ("indel", c_int)]


class affine_penalties_t(Structure):
    _fields_ = [("match", c_int), ("mismatch", c_int), ("gap_opening", c_int), ("gap_extension", c_int)]


class affine2p_penalties_t(Structure):
    _fields_ = [("match", c_int), ("mismatch", c_int), ("gap_opening1", c_int), ("gap_extension1", c_int),
                ("gap_opening2", c_int), ("gap_extension2", c_int)]


class wavefront_heuristic_t(Structure):
    _fields_ = [("strategy", c_uint)] + [(n, c_int) for n in (
        "steps_between_cutoffs", "min_k", "max_k", "min_wavefront_length", "max_distance_threshold",
        "xdrop", "zdrop", "steps_wait", "max_sw_score", "max_wf_score", "max_sw_score_offset",
        "max_sw_score_k")]


class timespec(Structure):
    _fields_ = [("tv_sec", c_int64), ("tv_nsec", c_int64)]


class profiler_counter_t(Structure):
    _fields_ = [("total", c_uint64), ("samples", c_uint64), ("min", c_uint64), ("max", c_uint64),
                ("m_oldM", c_double), ("m_newM", c_double), ("m_oldS", c_double), ("m_newS", c_double)]


class profiler_timer_t(Structure):
    _fields_ = [("begin_timer", timespec), ("time_ns", profiler_counter_t), ("accumulated", c_uint64)]


class alignment_system_t(Structure):
    _fields_ = [("max_alignment_steps", c_int), ("probe_interval_global", c_int),
                ("probe_interval_compact", c_int), ("max_partial_compacts", c_uint64),
                ("max_memory_compact", c_uint64), ("max_memory_resident", c_uint64),
                ("max_memory_abort", c_uint64), ("verbose", c_int), ("check_alignment_correct", c_bool),
                ("timer", profiler_timer_t), ("max_num_threads", c_int), ("min_offsets_per_thread", c_int)]


class wavefront_plot_attr_t(Structure):
    _fields_ = [("enabled", c_bool), ("resolution_points", c_int), ("align_level", c_int)]


class wavefront_aligner_attr_t(Structure):
    _fields_ = [("distance_metric", c_int), ("alignment_scope", c_int), ("alignment_form", alignment_form_t),
                ("linear_penalties", linear_penalties_t), ("affine_penalties", affine_penalties_t),
                ("affine2p_penalties", affine2p_penalties_t), ("heuristic", wavefront_heuristic_t),
                ("memory_mode", c_int), ("mm_allocator", c_void_p), ("plot", wavefront_plot_attr_t),
                ("system", alignment_system_t)]


class cigar_t(Structure):
    _fields_ = [("operations", c_void_p), ("max_operations", c_int), ("begin_offset", c_int),
                ("end_offset", c_int), ("score", c_int), ("end_v", c_int), ("end_h", c_int),
                ("has_misms", c_bool), ("cigar_buffer", c_void_p), ("cigar_length", c_int)]


class wavefront_align_status_t(Structure):
    _fields_ = [("status", c_int), ("score", c_int), ("dropped", c_bool), ("num_null_steps", c_int),
                ("memory_used", c_uint64), ("wf_align_compute", c_void_p), ("wf_align_extend", c_void_p)]


class wavefront_penalties_t(Structure):
    _fields_ = [("distance_metric", c_int), ("match", c_int), ("mismatch", c_int), ("gap_opening1", c_int),
                ("gap_extension1", c_int), ("gap_opening2", c_int), ("gap_extension2", c_int),
                ("linear_penalties", linear_penalties_t), ("affine_penalties", affine_penalties_t),
                ("affine2p_penalties", affine2p_penalties_t), ("internal_gap_e", c_int)]


class wavefront_pos_t(Structure):
    _fields_ = [("score", c_int), ("k", c_int), ("offset", c_int32)]


class wavefront_sequences_t(Structure):
    _fields_ = [("pattern", c_void_p), ("pattern_length", c_int), ("text", c_void_p), ("text_length", c_int)]


class wavefront_aligner_t(Structure):
    _fields_ = [("align_mode", c_int), ("align_mode_tag", c_char_p), ("align_status", wavefront_align_status_t),
                ("sequences", wavefront_sequences_t), ("alignment_scope", c_int),
                ("alignment_form", alignment_form_t), ("penalties", wavefront_penalties_t),
                ("heuristic", wavefront_heuristic_t), ("memory_mode", c_int), ("component_begin", c_int),
                ("component_end", c_int), ("alignment_end_pos", wavefront_pos_t), ("cigar", POINTER(cigar_t)),
                ("system", alignment_system_t), ("device", c_void_p)]


class wfa_pair_t(Structure):
    _fields_ = [("pattern", c_void_p), ("pattern_length", c_int), ("text", c_void_p), ("text_length", c_int)]


class wfa_result_t(Structure):
    _fields_ = [("status", c_int), ("score", c_int), ("end_v", c_int), ("end_h", c_int),
                ("cigar_offset", c_int64), ("cigar_length", c_int)]


class wfa_server_config_t(Structure):
    _fields_ = [("short_lanes", c_int), ("medium_lanes", c_int), ("long_lanes", c_int),
                ("short_max_length", c_int), ("long_min_length", c_int), ("max_batch_pairs", c_int),
                ("lane_memory_bytes", ctypes.c_size_t), ("error_prior", c_double),
                ("short_cu_share", c_double), ("medium_cu_share", c_double), ("long_cu_share", c_double)]


class wfa_server_result_t(Structure):
    _fields_ = [("id", c_int64), ("status", c_int), ("score", c_int), ("cigar_offset", c_int64),
                ("cigar_length", c_int), ("submit_time", c_double), ("start_time", c_double),
                ("finish_time", c_double)]


class wfa_run_stats_t(Structure):
    _fields_ = [("diagonal_steps", c_int64), ("forward_ms", c_double), ("backtrace_ms", c_double),
                ("forward_launches", c_int), ("record_arena_bytes", c_int64)]


_lib = None


def lib():
    """The loaded C-ABI library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libwfa2_amd.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.wavefront_aligner_new.restype = POINTER(wavefront_aligner_t)
        L.wavefront_aligner_new.argtypes = [POINTER(wavefront_aligner_attr_t)]
        L.wavefront_aligner_delete.argtypes = [POINTER(wavefront_aligner_t)]
        L.wavefront_align.argtypes = [POINTER(wavefront_aligner_t), c_char_p, c_int, c_char_p, c_int]
        L.wfa_align_seq_file.argtypes = [POINTER(wavefront_aligner_t), c_char_p, c_char_p, c_int, c_int, c_void_p]
        L.wavefront_align_batch.argtypes = [POINTER(wavefront_aligner_t), c_void_p, c_int, c_void_p,
                                            c_void_p, ctypes.c_size_t]
        L.wfa_device_batch_upload.restype = c_void_p
        L.wfa_device_batch_upload.argtypes = [POINTER(wavefront_aligner_t), c_void_p, c_int]
        L.wfa_device_batch_run.argtypes = [POINTER(wavefront_aligner_t), c_void_p]
        L.wfa_device_batch_run_noalloc.argtypes = [POINTER(wavefront_aligner_t), c_void_p]
        L.wfa_device_batch_download.argtypes = [POINTER(wavefront_aligner_t), c_void_p, c_void_p, c_void_p,
                                                ctypes.c_size_t]
        L.wfa_device_batch_free.argtypes = [c_void_p]
        L.wfa_device_batch_cigar_runs.restype = c_int64
        L.wfa_device_batch_cigar_runs.argtypes = [POINTER(wavefront_aligner_t), c_void_p]
        L.wfa_device_batch_download_runs.argtypes = [POINTER(wavefront_aligner_t), c_void_p, c_void_p, c_void_p,
                                                     ctypes.c_size_t]
        L.wfa_device_batch_stats.argtypes = [c_void_p, POINTER(wfa_run_stats_t)]
        L.wfa_server_config_default.argtypes = [POINTER(wfa_server_config_t)]
        L.wfa_server_create.restype = c_void_p
        L.wfa_server_create.argtypes = [POINTER(wavefront_aligner_attr_t), POINTER(wfa_server_config_t)]
        L.wfa_server_submit.restype = c_int64
        L.wfa_server_submit.argtypes = [c_void_p, c_char_p, c_int, c_char_p, c_int]
        L.wfa_server_replay.argtypes = [c_void_p, c_void_p, c_void_p, c_int]
        L.wfa_server_collect.argtypes = [c_void_p, c_void_p, c_int, c_void_p, ctypes.c_size_t, c_int]
        L.wfa_server_pending.restype = c_int64
        L.wfa_server_pending.argtypes = [c_void_p]
        L.wfa_server_clock.restype = c_double
        L.wfa_server_destroy.argtypes = [c_void_p]
        L.wavefront_align_strerror.restype = c_char_p
        L.wavefront_align_strerror_short.restype = c_char_p
        _lib = L
    return _lib


def default_attributes() -> wavefront_aligner_attr_t:
    """A copy of the library's wavefront_aligner_attr_default."""
    return wavefront_aligner_attr_t.from_buffer_copy(wavefront_aligner_attr_t.in_dll(lib(), "wavefront_aligner_attr_default"))


class BatchResult:
    """Per-pair results of one batch: numpy views of the wfa_result_t array
    and the CIGAR arena (filled in place by the library: no zero-fill, no
    second copy)."""

    def __init__(self, n, results, arena):
        self.n = n
        self._r = results
        self.arena = arena
        self.status = results["status"][:n]
        self.score = results["score"][:n]
        self.cigar_offset = results["cigar_offset"][:n]
        self.cigar_length = results["cigar_length"][:n]

    def ops(self, i) -> bytes:
        if self.arena is None:
            return b""
        o = int(self.cigar_offset[i])
        return self.arena[o:o + int(self.cigar_length[i])].tobytes()


# wfa_result_t as a numpy record (int status, score, end_v, end_h; int64 offset; int length; padding)
_RESULT_DTYPE = np.dtype([("status", np.int32), ("score", np.int32), ("end_v", np.int32), ("end_h", np.int32),
                          ("cigar_offset", np.int64), ("cigar_length", np.int32), ("pad", np.int32)])

_SAM_OPS = {7: "M", 8: "X", 1: "I", 2: "D"}


class RunsResult:
    """Results with run-length CIGARs (wfa_device_batch_download_runs)."""

    def __init__(self, n, res, runs, rc):
        self.n = n
        self.rc = rc
        self.status = res["status"]
        self.score = res["score"]
        self.cigar_offset = res["cigar_offset"]
        self.cigar_length = res["cigar_length"]
        self.runs = runs

    def pair_runs(self, i):
        o = int(self.cigar_offset[i])
        return self.runs[o:o + int(self.cigar_length[i])]

    def rle(self, i) -> str:
        """The CIGAR as cigar_sprint prints it with matches ("12M1X9M"), "-" when empty."""
        r = self.pair_runs(i)
        if len(r) == 0:
            return "-"
        return "".join("%d%s" % (int(x) >> 4, _SAM_OPS.get(int(x) & 15, "?")) for x in r)

    def ops(self, i) -> bytes:
        return b"".join(_SAM_OPS.get(int(x) & 15, "?").encode() * (int(x) >> 4) for x in self.pair_runs(i))


class _Pairs:
    """A wfa_pair_t array over one byte buffer (kept alive while in use).

    Accepts a PairSet-like object (buf, p_off, plen, t_off, tlen) or two
    lists of byte strings."""

    _dtype = np.dtype([("pattern", np.uint64), ("plen", np.int32), ("pad0", np.int32),
                       ("text", np.uint64), ("tlen", np.int32), ("pad1", np.int32)])

    def __init__(self, patterns, texts=None):
        if texts is None:
            ps = patterns
        else:
            from .datagen import PairSet
            ps = PairSet.from_lists(patterns, texts)
        self.n = int(ps.n)
        self.buf = ctypes.create_string_buffer(ps.buf, len(ps.buf) + 1)
        base = ctypes.addressof(self.buf)
        arr = np.zeros(max(self.n, 1), dtype=self._dtype)
        if self.n:
            arr["pattern"][:self.n] = base + ps.p_off.astype(np.uint64)
            arr["plen"][:self.n] = ps.plen
            arr["text"][:self.n] = base + ps.t_off.astype(np.uint64)
            arr["tlen"][:self.n] = ps.tlen
        self.arr = arr
        self.total = int(np.sum(ps.plen.astype(np.int64)) + np.sum(ps.tlen.astype(np.int64)))

    @property
    def address(self):
        return self.arr.ctypes.data


class DeviceBatch:
    """A batch resident in HBM (wfa_device_batch_upload / _run / _download)."""

    def __init__(self, aligner: "Aligner", patterns, texts=None):
        self.al = aligner
        self.pairs = _Pairs(patterns, texts)
        self.h = lib().wfa_device_batch_upload(aligner.wf, self.pairs.address, self.pairs.n)
        if not self.h:
            raise RuntimeError("wfa_device_batch_upload failed")

    def run(self) -> int:
        return lib().wfa_device_batch_run(self.al.wf, self.h)

    def run_noalloc(self) -> int:
        """wfa_device_batch_run_noalloc: the same run with no device allocation."""
        return lib().wfa_device_batch_run_noalloc(self.al.wf, self.h)

    def stats(self) -> wfa_run_stats_t:
        s = wfa_run_stats_t()
        lib().wfa_device_batch_stats(self.h, ctypes.byref(s))
        return s

    def download(self, cigars=True) -> BatchResult:
        n = self.pairs.n
        res = np.empty(max(n, 1), dtype=_RESULT_DTYPE)
        arena = np.empty(self.pairs.total + 1, dtype=np.uint8) if cigars else None
        lib().wfa_device_batch_download(self.al.wf, self.h, res.ctypes.data,
                                        arena.ctypes.data if arena is not None else None,
                                        self.pairs.total + 1 if arena is not None else 0)
        return BatchResult(n, res, arena)

    def download_runs(self, runs=None):
        """wfa_device_batch_download_runs: status, score, run offsets and
        counts per pair (numpy arrays) and the runs ((length << 4) | SAM code,
        uint32), into `runs` when it is a large enough uint32 array (a pinned
        or reused buffer avoids the allocation)."""
        total = lib().wfa_device_batch_cigar_runs(self.al.wf, self.h)
        if total < 0:
            raise RuntimeError("wfa_device_batch_cigar_runs returned %d" % total)
        n = self.pairs.n
        res = np.empty(max(n, 1), dtype=_RESULT_DTYPE)
        if runs is None or len(runs) < total:
            runs = np.empty(max(total, 1), dtype=np.uint32)
        rc = lib().wfa_device_batch_download_runs(self.al.wf, self.h, res.ctypes.data, runs.ctypes.data, len(runs))
        return RunsResult(n, res[:n], runs[:total], rc)

    def free(self):
        if self.h:
            lib().wfa_device_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def set_metric(a, metric, penalties):
    """Distance metric and penalties of attributes `a` (wavefront_aligner_attr_t).

    penalties: affine (match, mismatch, gap_opening, gap_extension); affine2p
    (match, mismatch, o1, e1, o2, e2); linear (match, mismatch, indel); indel
    and edit take none (their costs are fixed)."""
    a.distance_metric = METRICS[metric]
    if metric == "affine2p":
        a.affine2p_penalties = affine2p_penalties_t(*tuple(penalties))
    elif metric == "affine":
        a.affine_penalties = affine_penalties_t(*tuple(penalties)[:4])
    elif metric == "linear":
        a.linear_penalties = linear_penalties_t(*tuple(penalties)[:3])


class Aligner:
    """wavefront_aligner_new with attributes built from keyword arguments."""

    def __init__(self, metric="affine", penalties=(0, 4, 6, 2), memory="high", score_only=False,
                 max_steps=0, heuristic_none=True, ends_free=None, extension=False, heuristic=None):
        a = default_attributes()
        set_metric(a, metric, penalties)
        a.alignment_scope = 0 if score_only else 1
        a.memory_mode = MEMORY[memory]
        if ends_free is not None:  # (pattern_begin, pattern_end, text_begin, text_end)
            a.alignment_form.span = 1  # alignment_endsfree
            (a.alignment_form.pattern_begin_free, a.alignment_form.pattern_end_free,
             a.alignment_form.text_begin_free, a.alignment_form.text_end_free) = ends_free
        if extension:  # wavefront_aligner_set_alignment_extension
            a.alignment_form.span = 1
            a.alignment_form.extension = True
        if heuristic_none:
            a.heuristic.strategy = 0
        if heuristic is not None:
            # (strategy, min_wavefront_length, max_distance_threshold, steps_between_cutoffs
            #  [, min_k, max_k[, xdrop[, zdrop]]]); strategy bits as wf_heuristic_strategy: 1 banded
            # static, 2 banded adaptive, 4 wf-adaptive, 0x10 x-drop, 0x20 z-drop, 0x40 wfmash
            h = tuple(heuristic) + (0,) * (8 - len(heuristic))
            (a.heuristic.strategy, a.heuristic.min_wavefront_length, a.heuristic.max_distance_threshold,
             a.heuristic.steps_between_cutoffs, a.heuristic.min_k, a.heuristic.max_k, a.heuristic.xdrop,
             a.heuristic.zdrop) = h
            a.heuristic.steps_wait = h[3]
        if max_steps:
            a.system.max_alignment_steps = int(max_steps)
        self.attr = a
        self.score_only = score_only
        self.wf = lib().wavefront_aligner_new(ctypes.byref(a))
        if not self.wf:
            raise RuntimeError("wavefront_aligner_new failed")

    def close(self):
        if self.wf:
            lib().wavefront_aligner_delete(self.wf)
            self.wf = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def align(self, pattern: bytes, text: bytes):
        """wavefront_align: returns (status, cigar->score, CIGAR operations)."""
        st = lib().wavefront_align(self.wf, bytes(pattern), len(pattern), bytes(text), len(text))
        cg = self.wf.contents.cigar.contents
        ops = b""
        if cg.end_offset > cg.begin_offset and cg.operations:
            ops = ctypes.string_at(cg.operations + cg.begin_offset, cg.end_offset - cg.begin_offset)
        return st, cg.score, ops

    def align_batch(self, patterns, texts=None) -> BatchResult:
        """wavefront_align_batch over host buffers."""
        pr = _Pairs(patterns, texts)
        n = pr.n
        res = np.empty(max(n, 1), dtype=_RESULT_DTYPE)
        arena = None if self.score_only else np.empty(pr.total + 1, dtype=np.uint8)
        lib().wavefront_align_batch(self.wf, pr.address, n, res.ctypes.data,
                                    arena.ctypes.data if arena is not None else None,
                                    pr.total + 1 if arena is not None else 0)
        return BatchResult(n, res, arena)

    def upload(self, patterns, texts=None) -> DeviceBatch:
        return DeviceBatch(self, patterns, texts)

    def align_seq_file(self, input_path, output_path, batch_pairs=0, full=False):
        """wfa_align_seq_file: .seq in, .alg out; returns (status, wfa_io_stats_t)."""
        st = wfa_io_stats_t()
        rc = lib().wfa_align_seq_file(self.wf, os.fsencode(input_path), os.fsencode(output_path),
                                      int(batch_pairs), int(bool(full)), ctypes.byref(st))
        return rc, st


class wfa_io_stats_t(Structure):
    _fields_ = [("pairs", c_int64), ("bytes_in", c_int64), ("bytes_out", c_int64), ("seconds", ctypes.c_double),
                ("read_seconds", ctypes.c_double), ("align_seconds", ctypes.c_double),
                ("write_seconds", ctypes.c_double)]


class ServedResult:
    __slots__ = ("id", "status", "score", "ops", "submit", "start", "finish")

    def __init__(self, r, ops):
        self.id, self.status, self.score, self.ops = r.id, r.status, r.score, ops
        self.submit, self.start, self.finish = r.submit_time, r.start_time, r.finish_time

    @property
    def latency(self):
        return self.finish - self.submit


class Server:
    """wfa_server_*: a request loop over single-pair requests.

    Keyword arguments override the fields of wfa_server_config_default; the
    aligner attributes are built as for Aligner."""

    def __init__(self, metric="affine", penalties=(0, 4, 6, 2), **config):
        a = default_attributes()
        set_metric(a, metric, penalties)
        a.alignment_scope = 1
        a.heuristic.strategy = 0
        cfg = wfa_server_config_t()
        lib().wfa_server_config_default(ctypes.byref(cfg))
        for k, v in config.items():
            setattr(cfg, k, v)
        self.config = cfg
        self.attr = a
        self.h = lib().wfa_server_create(ctypes.byref(a), ctypes.byref(cfg))
        if not self.h:
            raise RuntimeError("wfa_server_create failed")
        self._replays = []
        self._arena = ctypes.create_string_buffer(1 << 24)

    def submit(self, pattern: bytes, text: bytes) -> int:
        return lib().wfa_server_submit(self.h, bytes(pattern), len(pattern), bytes(text), len(text))

    def replay(self, pairs, arrival_s):
        """Submit pair i at arrival_s[i] seconds from now (server thread)."""
        pr = pairs if isinstance(pairs, _Pairs) else _Pairs(pairs)
        at = np.ascontiguousarray(arrival_s, dtype=np.float64)
        assert len(at) == pr.n
        rc = lib().wfa_server_replay(self.h, pr.address, at.ctypes.data, pr.n)
        self._replays.append((pr, at))
        return rc

    def collect(self, max_results=4096, timeout_ms=100):
        res = (wfa_server_result_t * max_results)()
        n = lib().wfa_server_collect(self.h, ctypes.addressof(res), max_results, ctypes.addressof(self._arena),
                                     len(self._arena), int(timeout_ms))
        out = []
        for i in range(n):
            r = res[i]
            ops = ctypes.string_at(ctypes.addressof(self._arena) + r.cigar_offset, r.cigar_length) if r.cigar_length else b""
            out.append(ServedResult(r, ops))
        return out

    def pending(self) -> int:
        return lib().wfa_server_pending(self.h)

    @staticmethod
    def clock() -> float:
        return lib().wfa_server_clock()

    def close(self):
        if self.h:
            lib().wfa_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rle(ops: bytes) -> str:
    """CIGAR run-length form as cigar_sprint(print_matches=true) prints it."""
    if not ops:
        return "-"
    out = []
    i, n = 0, len(ops)
    while i < n:
        j = i
        while j < n and ops[j] == ops[i]:
            j += 1
        out.append("%d%c" % (j - i, ops[i]))
        i = j
    return "".join(out)


def declared_functions(header_path):
    """Function names declared in a C header of include/ (for export checks)."""
    import re
    text = open(header_path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "for", "while", "switch", "return", "sizeof", "defined"}
    decl = []
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M | re.S):
        if m.group(0).lstrip().startswith("typedef"):
            continue
        nm = m.group(1)
        if nm not in skip and not nm.endswith("_t"):
            decl.append(nm)
    return sorted(set(decl)) if decl else sorted(set(n for n in names if n not in skip))
