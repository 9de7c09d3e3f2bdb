"""Seeded synthetic workloads (ctypes front of libwfa_datagen.so).

Shapes: 0 = 1 kbp / 5% / 1-base indels; 1 = short reads, L ~ U{100..300},
e ~ U[0,10%]; 2 = 1-20 kbp log-uniform, e ~ U[1%,15%], geometric indels
(mean 3); 3 = 100 kbp / 15% / geometric indels (mean 2); 4 = serving mixture
(60% shape 0, 35% shape 2, 5% shape 3).  The text is uniform ACGT and the
pattern is the text with errors.  Pair i depends only on (seed, i).
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_float, c_int, c_int64, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwfa_datagen.so")


class _Gen(Structure):
    _fields_ = [("buf", c_void_p), ("bytes", c_int64), ("n", c_int), ("plen", POINTER(c_int)),
                ("tlen", POINTER(c_int)), ("p_off", POINTER(c_int64)), ("t_off", POINTER(c_int64)),
                ("err", POINTER(c_float)), ("shape", POINTER(c_int))]


_lib = None


def _L():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libwfa_datagen.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.wfa_gen.restype = POINTER(_Gen)
        L.wfa_gen.argtypes = [c_int, c_uint64, c_int64, c_int]
        L.wfa_gen_free.argtypes = [POINTER(_Gen)]
        _lib = L
    return _lib


class PairSet:
    """Pairs in one byte buffer: pattern i = buf[p_off[i]:p_off[i]+plen[i]]."""

    def __init__(self, buf, p_off, plen, t_off, tlen, err=None, shape=None):
        self.buf = buf
        self.p_off = np.ascontiguousarray(p_off, dtype=np.int64)
        self.plen = np.ascontiguousarray(plen, dtype=np.int32)
        self.t_off = np.ascontiguousarray(t_off, dtype=np.int64)
        self.tlen = np.ascontiguousarray(tlen, dtype=np.int32)
        self.err = err
        self.shape = shape
        self.n = len(self.plen)

    def pattern(self, i) -> bytes:
        o = int(self.p_off[i])
        return self.buf[o:o + int(self.plen[i])]

    def text(self, i) -> bytes:
        o = int(self.t_off[i])
        return self.buf[o:o + int(self.tlen[i])]

    def patterns(self):
        return [self.pattern(i) for i in range(self.n)]

    def texts(self):
        return [self.text(i) for i in range(self.n)]

    @staticmethod
    def from_lists(patterns, texts) -> "PairSet":
        pl = np.array([len(p) for p in patterns], dtype=np.int32)
        tl = np.array([len(t) for t in texts], dtype=np.int32)
        lens = np.empty(2 * len(pl), dtype=np.int64)
        lens[0::2] = pl
        lens[1::2] = tl
        offs = np.zeros_like(lens)
        if len(lens) > 1:
            offs[1:] = np.cumsum(lens[:-1])
        parts = [None] * (2 * len(pl))
        parts[0::2] = [bytes(p) for p in patterns]
        parts[1::2] = [bytes(t) for t in texts]
        return PairSet(b"".join(parts), offs[0::2], pl, offs[1::2], tl)

    def subset(self, idx) -> "PairSet":
        idx = np.asarray(idx)
        return PairSet.from_lists([self.pattern(i) for i in idx], [self.text(i) for i in idx])

    @property
    def cells(self) -> int:
        """Sum of n*m over the pairs (the GCUPS numerator)."""
        return int(np.sum(self.plen.astype(np.int64) * self.tlen.astype(np.int64)))


def generate(shape: int, seed: int, n: int, first: int = 0) -> PairSet:
    g = _L().wfa_gen(int(shape), int(seed), int(first), int(n))
    try:
        G = g.contents
        buf = ctypes.string_at(G.buf, G.bytes) if G.bytes else b""
        cnt = G.n
        p_off = np.ctypeslib.as_array(G.p_off, shape=(cnt,)).copy() if cnt else np.zeros(0, np.int64)
        plen = np.ctypeslib.as_array(G.plen, shape=(cnt,)).copy() if cnt else np.zeros(0, np.int32)
        t_off = np.ctypeslib.as_array(G.t_off, shape=(cnt,)).copy() if cnt else np.zeros(0, np.int64)
        tlen = np.ctypeslib.as_array(G.tlen, shape=(cnt,)).copy() if cnt else np.zeros(0, np.int32)
        err = np.ctypeslib.as_array(G.err, shape=(cnt,)).copy() if cnt else np.zeros(0, np.float32)
        shp = np.ctypeslib.as_array(G.shape, shape=(cnt,)).copy() if cnt else np.zeros(0, np.int32)
    finally:
        _L().wfa_gen_free(g)
    return PairSet(buf, p_off, plen, t_off, tlen, err, shp)


def skewed(seed: int, n: int, tlen=(600, 1200), insert=(1200, 2400), sub=0.03, unrelated_every=4) -> PairSet:
    """Pairs of very different lengths: the pattern is the text with one long
    random insertion and `sub` substitutions; every `unrelated_every`-th pair
    is two unrelated random sequences of the same lengths.  Their wavefronts
    grow wider than 1000 diagonals far from the end diagonal, which is where
    edit distance prunes (wavefront_compute_edit.c:219-278)."""
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b"ACGT", dtype=np.uint8)
    pats, texts = [], []
    for i in range(n):
        t = alphabet[rng.integers(0, 4, int(rng.integers(tlen[0], tlen[1] + 1)))]
        ins = alphabet[rng.integers(0, 4, int(rng.integers(insert[0], insert[1] + 1)))]
        at = int(rng.integers(0, len(t) + 1))
        p = np.concatenate([t[:at], ins, t[at:]])
        if unrelated_every and i % unrelated_every == unrelated_every - 1:
            p = alphabet[rng.integers(0, 4, len(p))]
        else:
            hit = rng.random(len(p)) < sub
            p = np.where(hit, alphabet[(np.searchsorted(alphabet, p) + rng.integers(1, 4, len(p))) % 4], p)
        pats.append(p.tobytes())
        texts.append(t.tobytes())
    return PairSet.from_lists(pats, texts)


def semiglobal(shape: int, seed: int, n: int, flank: int = 30) -> PairSet:
    """Pairs of the given shape with random flanks of up to `flank` bases on
    both sides of the text (odd pairs) or of the pattern (even pairs), the
    workload of ends-free alignment (free ends of either sequence)."""
    ps = generate(shape, seed, n)
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b"ACGT", dtype=np.uint8)
    pats, texts = [], []
    for i in range(ps.n):
        p, t = ps.pattern(i), ps.text(i)
        pre = alphabet[rng.integers(0, 4, int(rng.integers(0, flank)))].tobytes()
        suf = alphabet[rng.integers(0, 4, int(rng.integers(0, flank)))].tobytes()
        if i % 2:
            pats.append(p)
            texts.append(pre + t + suf)
        else:
            pats.append(pre + p + suf)
            texts.append(t)
    return PairSet.from_lists(pats, texts)
