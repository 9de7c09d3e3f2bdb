"""ctypes bindings for the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Two checkers live under oracle/:

* ``liboracle.so``: the plain-C restatement in ``wfa_oracle.c`` of WFA2-lib
  v2.3 ``wavefront_align`` (gap-affine / gap-affine-2p, memory high and
  BiWFA), which also counts diagonal-steps.
* ``_ref/libwfa_ref_v{2,3,4}.so``: the reference itself, compiled from its own
  sources by ``oracle/Makefile`` together with ``ref_shim.c`` (one aligner per
  OpenMP thread, as tools/align_benchmark/align_benchmark.c:504-586), for
  x86-64-v2 (scalar kernels), -v3 (AVX2) and -v4 (AVX-512).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py
import this module.  The product library never calls into it.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

# Status codes, wavefront/wfa.h:45-55
WF_STATUS_ALG_COMPLETED = 0
WF_STATUS_ALG_PARTIAL = 1
WF_STATUS_MAX_STEPS_REACHED = -100
WF_STATUS_UNATTAINABLE = -300
ORACLE_UNSUPPORTED = -400

MEMORY_MODES = {"high": 0, "med": 1, "low": 2, "ultralow": 3}


@dataclass(frozen=True)
class Conf:
    """Aligner configuration (a subset of wavefront_aligner_attr_t)."""

    metric: str = "affine"            # "indel" | "edit" | "linear" | "affine" | "affine2p"
    penalties: tuple = (0, 4, 6, 2)   # (M,X,O,E), (M,X,O1,E1,O2,E2), linear (M,X,I); unused for indel/edit
    memory: str = "high"
    score_only: bool = False
    max_steps: int = 0                # 0 = unlimited
    ends_free: tuple = None           # (pattern_begin, pattern_end, text_begin, text_end) or None = end-to-end
    extension: bool = False           # alignment_form.extension (ends-free to both sequence ends + max-trim)
    heuristic: tuple = None           # (strategy, min_wavefront_length, max_distance_threshold,
                                      #  steps_between_cutoffs[, min_k, max_k[, xdrop[, zdrop]]]): 4 wf-adaptive,
                                      #  0x40 wfmash, 0x10 x-drop, 0x20 z-drop, 1 banded static, 2 banded adaptive
                                      #  (bits may combine)

    @property
    def form6(self):
        """{span, pattern_begin, pattern_end, text_begin, text_end, extension} or None."""
        if self.ends_free is None and not self.extension:
            return None
        ef = tuple(self.ends_free) if self.ends_free is not None else (0, 0, 0, 0)
        return (1,) + ef + (int(self.extension),)

    @property
    def metric_id(self) -> int:
        return {"indel": 0, "edit": 1, "linear": 2, "affine": 3, "affine2p": 4}[self.metric]

    @property
    def pen6(self) -> tuple:
        p = tuple(self.penalties)
        return p + (0,) * (6 - len(p))


class _CConf(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "metric", "match", "mismatch", "gap_opening1", "gap_extension1",
        "gap_opening2", "gap_extension2", "memory_mode", "score_only", "max_steps",
        "ends_free", "pattern_begin_free", "pattern_end_free", "text_begin_free", "text_end_free",
        "extension", "heuristic", "min_wavefront_length", "max_distance_threshold", "steps_between_cutoffs",
        "min_k", "max_k", "xdrop", "zdrop")]


def _cconf(conf: Conf) -> _CConf:
    p = conf.pen6
    f = conf.form6 or (0, 0, 0, 0, 0, 0)
    h = tuple(conf.heuristic or (0, 0, 0, 0))
    h = h + (0,) * (8 - len(h))
    return _CConf(conf.metric_id, p[0], p[1], p[2], p[3], p[4], p[5],
                  MEMORY_MODES[conf.memory], int(conf.score_only), int(conf.max_steps), *f, *h)


class Pairs:
    """Pairs packed into one byte buffer with int64 offsets."""

    def __init__(self, patterns, texts):
        assert len(patterns) == len(texts)
        pl = np.fromiter((len(p) for p in patterns), dtype=np.int32, count=len(patterns))
        tl = np.fromiter((len(t) for t in texts), dtype=np.int32, count=len(texts))
        inter = [None] * (2 * len(patterns))
        inter[0::2] = [bytes(p) for p in patterns]
        inter[1::2] = [bytes(t) for t in texts]
        self.buf = b"".join(inter)
        lens = np.empty(2 * len(patterns), dtype=np.int64)
        lens[0::2] = pl
        lens[1::2] = tl
        offs = np.zeros(2 * len(patterns), dtype=np.int64)
        if len(lens) > 1:
            offs[1:] = np.cumsum(lens[:-1])
        self.p_off = np.ascontiguousarray(offs[0::2])
        self.t_off = np.ascontiguousarray(offs[1::2])
        self.plen = pl
        self.tlen = tl
        self.n = len(patterns)
        cap = (pl.astype(np.int64) + tl.astype(np.int64))
        self.ops_off = np.zeros(self.n, dtype=np.int64)
        if self.n > 1:
            self.ops_off[1:] = np.cumsum(cap[:-1])
        self.ops_cap = int(cap.sum()) + 1

    def pattern(self, i):
        o = int(self.p_off[i])
        return self.buf[o:o + int(self.plen[i])]

    def text(self, i):
        o = int(self.t_off[i])
        return self.buf[o:o + int(self.tlen[i])]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


_oracle_lib = None


def oracle_lib():
    global _oracle_lib
    if _oracle_lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(ORACLE_SO)
        lib.oracle_align_batch.restype = None
        lib.oracle_cigar_score.restype = ctypes.c_int
        _oracle_lib = lib
    return _oracle_lib


class Result:
    def __init__(self, n, status, score, ops_raw, ops_off, ops_len, dsteps=None, end_score=None,
                 alg_score=None, seconds=None, stale=None):
        self.n = n
        self.status = status
        self.score = score
        self.ops_raw = ops_raw
        self.ops_off = ops_off
        self.ops_len = ops_len
        self.dsteps = dsteps
        self.end_score = end_score
        self.alg_score = alg_score
        self.seconds = seconds
        self.stale = stale  # oracle only: the backtrace met cells a cut-off left readable

    def ops(self, i) -> bytes:
        o = int(self.ops_off[i])
        return self.ops_raw[o:o + int(self.ops_len[i])]


def oracle_align(conf: Conf, pairs: Pairs) -> Result:
    lib = oracle_lib()
    n = pairs.n
    status = np.zeros(n, np.int32)
    score = np.zeros(n, np.int32)
    ops_len = np.zeros(n, np.int32)
    dsteps = np.zeros(n, np.int64)
    end_score = np.zeros(n, np.int32)
    stale = np.zeros(n, np.int32)
    ops = ctypes.create_string_buffer(pairs.ops_cap)
    cc = _cconf(conf)
    lib.oracle_align_batch(ctypes.byref(cc), ctypes.c_int(n), pairs.buf,
                           _ptr(pairs.p_off), _ptr(pairs.plen), _ptr(pairs.t_off), _ptr(pairs.tlen),
                           _ptr(status), _ptr(score), ops, _ptr(pairs.ops_off), _ptr(ops_len),
                           _ptr(dsteps), _ptr(end_score), _ptr(stale))
    r = Result(n, status, score, ops.raw, pairs.ops_off, ops_len, dsteps, end_score, stale=stale)
    r.alg_score = alg_scores(conf, r)
    return r


def cigar_score(conf: Conf, ops: bytes) -> int:
    cc = _cconf(conf)
    return oracle_lib().oracle_cigar_score(ctypes.byref(cc), ops, len(ops))


def alg_scores(conf: Conf, r: Result) -> np.ndarray:
    """Score column of align_benchmark's .alg output (benchmark_utils.c:143-179)."""
    out = np.zeros(r.n, np.int32)
    for i in range(r.n):
        if conf.score_only:
            out[i] = r.score[i]
        elif r.ops_len[i] > 0:
            out[i] = cigar_score(conf, r.ops(i))
        else:
            out[i] = -1
    return out


# --------------------------------------------------------------- reference

def _cpu_has_avx512() -> bool:
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return " avx512cd" in flags and " avx512vl" in flags


_ref_libs = {}


def ref_lib_path(isa=None) -> str:
    """isa: "v2" (scalar kernels), "v3" (AVX2), "v4" (AVX-512); default: the
    widest this host runs."""
    v = isa or ("v4" if _cpu_has_avx512() else "v3")
    return os.path.join(REF_DIR, "libwfa_ref_%s.so" % v)


def ref_available(isa=None) -> bool:
    return os.path.exists(ref_lib_path(isa))


def ref_lib(isa=None):
    path = ref_lib_path(isa)
    if path not in _ref_libs:
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (make -C oracle ref)" % path)
        lib = ctypes.CDLL(path)
        lib.ref_align_batch_ex.restype = ctypes.c_double
        lib.ref_align_time.restype = ctypes.c_double
        _ref_libs[path] = lib
    return _ref_libs[path]


def ref_align(conf: Conf, pairs: Pairs, threads: int = 4, isa=None) -> Result:
    lib = ref_lib(isa)
    n = pairs.n
    status = np.zeros(n, np.int32)
    score = np.zeros(n, np.int32)
    alg = np.zeros(n, np.int32)
    ops_len = np.zeros(n, np.int32)
    ops = ctypes.create_string_buffer(pairs.ops_cap)
    pen = (ctypes.c_int * 6)(*conf.pen6)
    f6 = conf.form6
    form = (ctypes.c_int * 6)(*f6) if f6 is not None else None
    heur = (ctypes.c_int * 8)(*(tuple(conf.heuristic) + (0,) * (8 - len(conf.heuristic)))) \
        if conf.heuristic is not None else None
    secs = lib.ref_align_batch_ex(conf.metric_id, pen, MEMORY_MODES[conf.memory], int(conf.score_only),
                                  int(conf.max_steps), form, heur, int(threads), n, pairs.buf,
                                  _ptr(pairs.p_off), _ptr(pairs.plen), _ptr(pairs.t_off), _ptr(pairs.tlen),
                                  _ptr(status), _ptr(score), _ptr(alg), ops, _ptr(pairs.ops_off), _ptr(ops_len))
    return Result(n, status, score, ops.raw, pairs.ops_off, ops_len, alg_score=alg, seconds=secs)


def ref_time(conf: Conf, pairs: Pairs, threads: int) -> tuple:
    """Wall seconds of the reference aligning all pairs on `threads` threads."""
    lib = ref_lib()
    pen = (ctypes.c_int * 6)(*conf.pen6)
    chk = ctypes.c_int64(0)
    secs = lib.ref_align_time(conf.metric_id, pen, MEMORY_MODES[conf.memory], int(threads), pairs.n,
                              pairs.buf, _ptr(pairs.p_off), _ptr(pairs.plen), _ptr(pairs.t_off),
                              _ptr(pairs.tlen), ctypes.byref(chk))
    return secs, chk.value


# ------------------------------------------------------------------ formats

def read_seq(path):
    """`>pattern` / `<text` line pairs (tools/align_benchmark/align_benchmark.c:313-339)."""
    pats, texts = [], []
    with open(path, "rb") as f:
        for line in f:
            line = line.rstrip(b"\r\n")
            if line.startswith(b">"):
                pats.append(line[1:])
            elif line.startswith(b"<"):
                texts.append(line[1:])
    return pats, texts


def rle(ops: bytes) -> str:
    """cigar_sprint with matches printed (alignment/cigar.c:436-470)."""
    if not ops:
        return "-"
    out = []
    i = 0
    n = len(ops)
    while i < n:
        j = i
        c = ops[i]
        while j < n and ops[j] == c:
            j += 1
        out.append("%d%c" % (j - i, c))
        i = j
    return "".join(out)


def alg_lines(r: Result, score_only=False):
    if score_only:
        return ["%d\t-" % r.alg_score[i] for i in range(r.n)]
    return ["%d\t%s" % (r.alg_score[i], rle(r.ops(i))) for i in range(r.n)]
