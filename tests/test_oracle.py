"""Pins the CPU oracle (oracle/wfa_oracle.c) before it is trusted as checker.

1. Golden vectors: the reference's own regression outputs for the 305-pair
   suite (tests/wfa.utest.check/*.alg) in every exact mode the device path
   covers: memory high / med / score-only / BiWFA / BiWFA score-only,
   gap-affine with seven penalty sets (match < 0 included), gap-affine-2p,
   edit and indel.  (Gap-linear has no golden file: it is pinned by 2 and 3.)
2. The reference itself, compiled from its sources into oracle/_ref/, on
   seeded random pairs of the benchmark shapes (skipped if not built).
3. Mathematics: an independent Gotoh DP (tests/dp.py) for the optimal score,
   CIGAR validity and re-scoring, and closed forms.
"""
import random

import numpy as np
import pytest

import dp
import golden_suite as G
import oracle as O


def _pairs(pats, texts):
    return O.Pairs(pats, texts)


@pytest.fixture(scope="module")
def suite():
    return G.read_seq()


@pytest.mark.parametrize("case", G.cases(), ids=[c[0] for c in G.cases()])
def test_oracle_matches_golden(suite, case):
    name, metric, pen, mem, so, heur = case
    pats, texts = suite
    conf = O.Conf(metric, pen, mem, so, heuristic=heur)
    r = O.oracle_align(conf, _pairs(pats, texts))
    assert O.alg_lines(r, so) == G.golden_lines(name)


def _rand_seq(rng, n, alphabet=b"ACGT"):
    return bytes(rng.choice(alphabet) for _ in range(n))


def _mutate(rng, s, nerr, max_indel=1):
    s = bytearray(s)
    for _ in range(nerr):
        k = rng.randrange(3)
        if k == 0 and s:
            p = rng.randrange(len(s))
            s[p] = rng.choice([c for c in b"ACGT" if c != s[p]])
        elif k == 1:
            p = rng.randrange(len(s) + 1)
            L = rng.randint(1, max_indel)
            s[p:p] = bytes(rng.choice(b"ACGT") for _ in range(L))
        elif s:
            p = rng.randrange(len(s))
            L = rng.randint(1, max_indel)
            del s[p:p + L]
    return bytes(s)


def _random_pairs(seed, count, lo, hi, maxerr, max_indel=1):
    rng = random.Random(seed)
    pats, texts = [], []
    for _ in range(count):
        t = _rand_seq(rng, rng.randint(lo, hi))
        p = _mutate(rng, t, rng.randint(0, maxerr), max_indel)
        pats.append(p)
        texts.append(t)
    return pats, texts


PEN_SETS = [
    ("affine", (0, 4, 6, 2)),
    ("affine", (0, 1, 2, 1)),
    ("affine", (-3, 5, 3, 2)),
    ("affine", (-5, 1, 2, 1)),
    ("affine2p", (0, 4, 6, 2, 24, 1)),
    ("affine2p", (-2, 3, 4, 3, 12, 1)),
]


@pytest.mark.parametrize("metric,pen", PEN_SETS)
def test_oracle_score_is_dp_optimum(metric, pen):
    pats, texts = _random_pairs(hash((metric, pen)) & 0xFFFF, 60, 0, 36, 8, max_indel=4)
    for memory in ("high", "ultralow"):
        r = O.oracle_align(O.Conf(metric, pen, memory), _pairs(pats, texts))
        for i, (p, t) in enumerate(zip(pats, texts)):
            assert r.status[i] == 0
            best = dp.gotoh(p, t, *pen)
            ops = r.ops(i)
            assert dp.cigar_valid(p, t, ops), (p, t, ops)
            assert dp.cigar_score(ops, *pen) == best, (p, t, ops)
            if memory == "high" or (len(p) and len(t)):
                assert r.score[i] == best


# (metric, oracle penalties, Gotoh-equivalent (match, mismatch, open, extend), sign of the score)
LINEAR_SETS = [
    ("edit", (0, 1, 1), (0, 1, 0, 1), -1),   # edit/indel scores are distances (positive)
    ("indel", (0, 1, 1), (0, 2, 0, 1), -1),  # a mismatch costs as much as I+D: same optimum
    ("linear", (0, 4, 2), (0, 4, 0, 2), 1),
    ("linear", (-1, 3, 2), (-1, 3, 0, 2), 1),
]


@pytest.mark.parametrize("metric,pen,gpen,sign", LINEAR_SETS)
def test_oracle_linear_metrics_are_dp_optimum(metric, pen, gpen, sign):
    """Edit, indel and gap-linear: the score is the optimum of an independent
    DP with the equivalent costs, the CIGAR is valid and re-scores to it."""
    pats, texts = _random_pairs(hash((metric, pen)) & 0xFFFF, 60, 0, 36, 8, max_indel=4)
    for memory in ("high", "ultralow"):
        r = O.oracle_align(O.Conf(metric, pen, memory), _pairs(pats, texts))
        for i, (p, t) in enumerate(zip(pats, texts)):
            assert r.status[i] == 0
            best = dp.gotoh(p, t, *gpen)
            ops = r.ops(i)
            assert dp.cigar_valid(p, t, ops), (p, t, ops)
            assert dp.cigar_score(ops, *gpen) == best, (p, t, ops)
            if metric == "indel":
                assert b"X" not in ops
            if memory == "high" or (len(p) and len(t)):
                assert r.score[i] == sign * best


def test_oracle_closed_forms():
    conf = O.Conf("affine", (0, 4, 6, 2))
    t = b"ACGTTGCAACGTACGGATCCATGACCAGTACGATCG"
    sub = t[:10] + (b"A" if t[10:11] != b"A" else b"C") + t[11:]
    ins3 = t[:15] + b"GGG" + t[15:]
    r = O.oracle_align(conf, _pairs([t, sub, ins3, t[:20] + t[27:]], [t, t, t, t]))
    assert list(r.score) == [0, -4, -(6 + 3 * 2), -(6 + 7 * 2)]
    assert r.ops(0) == b"M" * len(t)
    # dual cost: a long gap takes the cheaper second piece
    conf2 = O.Conf("affine2p", (0, 4, 6, 2, 24, 1))
    long_del = t[:5] + b"A" * 30 + t[5:]
    r2 = O.oracle_align(conf2, _pairs([long_del, ins3], [t, t]))
    assert r2.score[0] == -min(6 + 30 * 2, 24 + 30)
    assert r2.score[1] == -min(6 + 3 * 2, 24 + 3)


def test_oracle_empty_and_trivial():
    conf = O.Conf("affine", (0, 4, 6, 2))
    r = O.oracle_align(conf, _pairs([b"", b"ACGT", b"", b"A"], [b"", b"", b"ACG", b"A"]))
    assert list(r.status) == [0, 0, 0, 0]
    assert r.ops(0) == b"" and r.ops(1) == b"DDDD" and r.ops(2) == b"III" and r.ops(3) == b"M"
    assert list(r.score) == [0, -(6 + 8), -(6 + 6), 0]


def test_oracle_max_steps():
    conf = O.Conf("affine", (0, 4, 6, 2), max_steps=10)
    t = b"ACGTACGTTTGACCA" * 4
    p = _mutate(random.Random(3), t, 12)
    r = O.oracle_align(conf, _pairs([p], [t]))
    assert r.status[0] == O.WF_STATUS_MAX_STEPS_REACHED and r.score[0] == -10


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("metric,pen,memory,shape", [
    ("affine", (0, 4, 6, 2), "high", "short"),
    ("affine", (0, 4, 6, 2), "high", "kbp"),
    ("affine2p", (0, 4, 6, 2, 24, 1), "high", "kbp"),
    ("affine", (0, 4, 6, 2), "ultralow", "kbp"),
    ("affine2p", (0, 4, 6, 2, 24, 1), "ultralow", "kbp"),
    ("affine", (-3, 5, 3, 2), "high", "short"),
    ("edit", (0, 1, 1), "high", "short"),
    ("indel", (0, 1, 1), "high", "short"),
    ("linear", (0, 4, 2), "high", "short"),
    ("linear", (-1, 3, 2), "high", "short"),
    ("edit", (0, 1, 1), "ultralow", "kbp"),
    ("indel", (0, 1, 1), "ultralow", "kbp"),
    ("linear", (0, 4, 2), "ultralow", "kbp"),
    ("edit", (0, 1, 1), "high", "kbp"),
    ("edit", (0, 1, 1), "high", "skew"),
    ("edit", (0, 1, 1), "ultralow", "skew"),
    ("indel", (0, 1, 1), "high", "skew"),
    ("linear", (-1, 3, 2), "ultralow", "skew"),
])
def test_oracle_matches_reference_build(metric, pen, memory, shape):
    """The skewed pairs (one long insertion) make edit distance prune its
    wavefronts (wavefront_compute_edit.c:219-278)."""
    from wfa_amd import datagen
    if shape == "skew":
        ps = datagen.skewed(5, 8)
    else:
        ps = datagen.generate(1, 11, 600) if shape == "short" else datagen.generate(2, 12, 6)
    pairs = _pairs(ps.patterns(), ps.texts())
    conf = O.Conf(metric, pen, memory)
    a = O.oracle_align(conf, pairs)
    b = O.ref_align(conf, pairs, threads=4)
    assert list(a.status) == list(b.status)
    assert list(a.score) == list(b.score)
    for i in range(pairs.n):
        assert a.ops(i) == b.ops(i), i


def test_oracle_edit_prune_cuts_diagonal_steps():
    # Edit and indel share the recurrence except for the mismatch source;
    # on skewed pairs only edit prunes, and its wavefronts stay far narrower.
    from wfa_amd import datagen
    ps = datagen.skewed(5, 8)
    pairs = _pairs(ps.patterns(), ps.texts())
    e = O.oracle_align(O.Conf("edit", (0, 1, 1)), pairs)
    i = O.oracle_align(O.Conf("indel", (0, 1, 1)), pairs)
    assert e.dsteps.sum() * 2 < i.dsteps.sum()


def test_oracle_counts_diagonal_steps():
    # W for a fixed pair is deterministic; two identical sequences need one
    # extension and no compute step at all.
    conf = O.Conf("affine", (0, 4, 6, 2))
    t = b"ACGTACGGTCA" * 9
    r = O.oracle_align(conf, _pairs([t, t[:50] + b"T" + t[51:]], [t, t]))
    assert r.dsteps[0] == 0
    assert r.dsteps[1] > 0 and r.end_score[1] == 4


# ------------------------------------------------------------- ends-free

def _semiglobal_pairs(seed, n, flank=30):
    from wfa_amd import datagen
    ps = datagen.semiglobal(1, seed, n, flank)
    return ps.patterns(), ps.texts()


# (metric, penalties, form): every ends-free form the device path accepts
ENDSFREE_SETS = [
    ("affine", (0, 4, 6, 2), dict(ends_free=(0, 0, 30, 30))),
    ("affine", (0, 4, 6, 2), dict(ends_free=(30, 30, 0, 0))),
    ("affine", (0, 4, 6, 2), dict(ends_free=(10, 10, 10, 10))),
    ("affine", (-2, 4, 6, 2), dict(ends_free=(0, 30, 0, 30))),
    ("affine", (0, 4, 6, 2), dict(extension=True)),
    ("affine", (-1, 3, 5, 2), dict(extension=True)),
    ("affine2p", (0, 4, 6, 2, 24, 1), dict(ends_free=(20, 20, 20, 20))),
    ("affine2p", (0, 4, 6, 2, 24, 1), dict(extension=True)),
    ("linear", (0, 4, 2), dict(ends_free=(30, 30, 30, 30))),
    ("linear", (-2, 3, 4), dict(extension=True)),
    ("edit", (0, 1, 1), dict(ends_free=(0, 30, 0, 30))),
    ("edit", (0, 1, 1), dict(extension=True)),
    ("indel", (0, 1, 1), dict(ends_free=(0, 0, 0, 30))),
]


@pytest.mark.skipif(not O.ref_available("v2"), reason="scalar reference build (oracle/_ref) not present")
@pytest.mark.parametrize("metric,pen,form", ENDSFREE_SETS)
def test_oracle_endsfree_matches_reference_build(metric, pen, form):
    """Against the reference's scalar build (x86-64-v2): its ends-free
    extension tests diagonals in ascending order (wavefront_extend_kernels.c:
    154-196); the AVX2 / AVX-512 kernels test the diagonals whose first four
    symbols match before the others (wavefront_extend_kernels_avx.c:296-392),
    so they can end on another diagonal of the same score."""
    pats, texts = _semiglobal_pairs(31, 1500)
    pairs = _pairs(pats, texts)
    # indel / edit extension in memory med reports score INT32_MIN (its
    # piggyback backtrace leaves cigar->score unset); the device path refuses it
    mems = ("high",) if metric in ("edit", "indel") and form.get("extension") else ("high", "med")
    for mem in mems:
        for so in (False, True):
            conf = O.Conf(metric, pen, mem, score_only=so, **form)
            a = O.oracle_align(conf, pairs)
            b = O.ref_align(conf, pairs, isa="v2")
            assert list(a.status) == list(b.status), (mem, so)
            assert list(a.score) == list(b.score), (mem, so)
            for i in range(pairs.n):
                assert a.ops(i) == b.ops(i), (mem, so, i)


@pytest.mark.parametrize("metric,pen,gpen,form", [
    ("affine", (0, 4, 6, 2), (0, 4, 6, 2), (0, 0, 5, 5)),
    ("affine", (0, 4, 6, 2), (0, 4, 6, 2), (5, 5, 0, 0)),
    ("affine", (0, 4, 6, 2), (0, 4, 6, 2), (3, 4, 2, 6)),
    ("affine2p", (0, 4, 6, 2, 24, 1), (0, 4, 6, 2, 24, 1), (4, 4, 4, 4)),
    ("linear", (0, 4, 2), (0, 4, 0, 2), (4, 4, 4, 4)),
    ("edit", (0, 1, 1), (0, 1, 0, 1), (0, 5, 0, 5)),
    ("indel", (0, 1, 1), (0, 2, 0, 1), (0, 4, 0, 6)),
])
def test_oracle_endsfree_score_is_dp_optimum(metric, pen, gpen, form):
    """Ends-free scores equal a semi-global DP's optimum; the CIGAR is valid
    and, without the gaps its free ends absorb, re-scores to it."""
    pats, texts = _random_pairs(hash((metric, form)) & 0xFFFF, 80, 0, 30, 6, max_indel=4)
    r = O.oracle_align(O.Conf(metric, pen, "high", ends_free=form), _pairs(pats, texts))
    sign = -1 if metric in ("edit", "indel") else 1
    for i, (p, t) in enumerate(zip(pats, texts)):
        if r.status[i] == O.ORACLE_UNSUPPORTED:
            assert form[0] > len(p) or form[1] > len(p) or form[2] > len(t) or form[3] > len(t)
            continue
        assert r.status[i] == 0
        best = dp.gotoh_endsfree(p, t, *gpen, pbf=form[0], pef=form[1], tbf=form[2], tef=form[3])
        assert r.score[i] == sign * best, (p, t, r.ops(i))
        ops = r.ops(i)
        assert dp.cigar_valid(p, t, ops), (p, t, ops)
        assert dp.cigar_score(dp.strip_free_ends(ops, *form), *gpen) == best, (p, t, ops)


@pytest.mark.parametrize("metric,pen", [("affine", (0, 4, 6, 2)), ("affine", (-1, 3, 5, 2)),
                                        ("affine2p", (0, 4, 6, 2, 24, 1)), ("linear", (-2, 3, 4)),
                                        ("linear", (0, 4, 2))])
def test_oracle_extension_is_maxtrimmed_endsfree(metric, pen):
    """Extension = free ends up to both sequence ends (wavefront_align.c:57-63)
    then the best-scoring CIGAR prefix (alignment/cigar_utils.c:111-278),
    status partial when it was cut."""
    pats, texts = _semiglobal_pairs(32, 400)
    pairs = _pairs(pats, texts)
    ext = O.oracle_align(O.Conf(metric, pen, "high", extension=True), pairs)
    trimmed_any = False
    for i in range(pairs.n):
        full = O.oracle_align(O.Conf(metric, pen, "high",
                                     ends_free=(0, len(pats[i]), 0, len(texts[i]))), _pairs([pats[i]], [texts[i]]))
        want_ops, want_score, trimmed = dp.maxtrim(full.ops(0), metric, *pen)
        assert ext.ops(i) == want_ops, i
        assert ext.status[i] == (O.WF_STATUS_ALG_PARTIAL if trimmed else 0), i
        assert ext.score[i] == (want_score if want_score is not None else -2**31), i
        trimmed_any |= trimmed
    assert trimmed_any


def test_oracle_endsfree_closed_forms():
    t = b"GATTACAGATTACACCGGTTAACC"
    inner = t[5:19]
    conf = O.Conf("affine", (0, 4, 6, 2), ends_free=(0, 0, 5, 5))
    r = O.oracle_align(conf, _pairs([inner], [t]))
    assert r.status[0] == 0 and r.score[0] == 0
    assert r.ops(0) == b"I" * 5 + b"M" * len(inner) + b"I" * 5
    # one free symbol short at the end: the last insertion is charged o + e
    r = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), ends_free=(0, 0, 5, 4)), _pairs([inner], [t]))
    assert r.score[0] == -(6 + 2)
    # the pattern's free ends
    r = O.oracle_align(O.Conf("edit", (0, 1, 1), ends_free=(0, 5, 0, 0)), _pairs([t], [t[:-5]]))
    assert r.score[0] == 0 and r.ops(0) == b"M" * (len(t) - 5) + b"D" * 5


def test_oracle_refuses_what_it_does_not_restate():
    pairs = _pairs([b"ACGTACGT"], [b"TTACGTACGTTT"])
    for conf in (O.Conf("edit", (0, 1, 1), ends_free=(0, 0, 2, 2)),          # free beginning, edit
                 O.Conf("affine", (-1, 4, 6, 2), ends_free=(0, 0, 2, 2)),    # free beginning, match < 0
                 O.Conf("affine", (0, 4, 6, 2), "ultralow", ends_free=(0, 0, 2, 2)),
                 O.Conf("affine", (0, 4, 6, 2), "ultralow", extension=True)):
        assert O.oracle_align(conf, pairs).status[0] == O.ORACLE_UNSUPPORTED


# ------------------------------------------------------------- heuristics

WFAPT_SETS = [
    ("affine", (0, 4, 6, 2), (4, 10, 50, 1)),
    ("affine", (-3, 5, 3, 2), (4, 10, 50, 10)),
    ("affine2p", (0, 4, 6, 2, 24, 1), (4, 10, 50, 1)),
    ("linear", (0, 4, 2), (4, 50, 10, 3)),
    ("edit", (0, 1, 1), (4, 10, 50, 1)),
    ("indel", (0, 1, 1), (4, 10, 50, 10)),
]


@pytest.mark.skipif(not O.ref_available("v2"), reason="scalar reference build (oracle/_ref) not present")
@pytest.mark.parametrize("metric,pen,heur", WFAPT_SETS)
def test_oracle_wfadaptive_matches_reference_build(metric, pen, heur):
    """wf-adaptive (wavefront_heuristic.c:229-290) in every memory mode, with
    CIGARs and score-only, short reads and 1-10 kbp pairs."""
    from wfa_amd import datagen
    a, b = datagen.generate(1, 41, 600), datagen.generate(2, 42, 6)
    keep = [i for i in range(b.n) if b.plen[i] <= 10000]
    pats = a.patterns() + [b.pattern(i) for i in keep]
    texts = a.texts() + [b.text(i) for i in keep]
    pairs = _pairs(pats, texts)
    for mem in ("high", "ultralow"):
        for so in (False, True):
            conf = O.Conf(metric, pen, mem, so, heuristic=heur)
            x = O.oracle_align(conf, pairs)
            y = O.ref_align(conf, pairs, isa="v2")
            assert list(x.status) == list(y.status), (mem, so)
            assert list(x.score) == list(y.score), (mem, so)
            if not so:
                for i in range(pairs.n):
                    assert x.ops(i) == y.ops(i), (mem, i)


@pytest.mark.skipif(not O.ref_available("v2"), reason="scalar reference build (oracle/_ref) not present")
def test_oracle_wfadaptive_endsfree_matches_reference_build():
    from wfa_amd import datagen
    ps = datagen.semiglobal(1, 43, 800)
    pairs = _pairs(ps.patterns(), ps.texts())
    for metric, pen, form in (("affine", (0, 4, 6, 2), dict(ends_free=(0, 0, 30, 30))),
                              ("affine", (0, 4, 6, 2), dict(extension=True)),
                              ("edit", (0, 1, 1), dict(ends_free=(0, 30, 0, 30)))):
        conf = O.Conf(metric, pen, "high", heuristic=(4, 10, 50, 1), **form)
        x = O.oracle_align(conf, pairs)
        y = O.ref_align(conf, pairs, isa="v2")
        assert list(x.status) == list(y.status) and list(x.score) == list(y.score)
        assert all(x.ops(i) == y.ops(i) for i in range(pairs.n))


BANDED = [
    (1, 0, 0, 0, -10, 10),      # banded static
    (1, 0, 0, 0, -3, 40),       # banded static, lopsided band
    (1, 0, 0, 0, 5, 10),        # banded static without diagonal 0: nothing is reachable
    (2, 0, 0, 1, -10, 10),      # banded adaptive
    (2, 0, 0, 3, -5, 5),        # banded adaptive, every third step
    (5, 10, 50, 1, -20, 20),    # wf-adaptive then banded static
    (6, 10, 50, 2, -8, 8),      # wf-adaptive then banded adaptive
]


def _banded_pairs():
    from wfa_amd import datagen
    a, b = datagen.generate(1, 41, 300), datagen.generate(2, 42, 4)
    ps = datagen.semiglobal(1, 43, 100)
    keep = [i for i in range(b.n) if b.plen[i] <= 10000]
    return (a.patterns() + [b.pattern(i) for i in keep] + ps.patterns(),
            a.texts() + [b.text(i) for i in keep] + ps.texts())


@pytest.mark.skipif(not O.ref_available("v2"), reason="scalar reference build (oracle/_ref) not present")
@pytest.mark.parametrize("metric,pen", [("affine", (0, 4, 6, 2)), ("affine", (-3, 5, 3, 2)),
                                        ("affine2p", (0, 4, 6, 2, 24, 1)), ("linear", (0, 4, 2)),
                                        ("edit", (0, 1, 1)), ("indel", (0, 1, 1))])
def test_oracle_banded_matches_reference_build(metric, pen):
    """Banded static and banded adaptive (wavefront_heuristic.c:451-500),
    alone and after wf-adaptive, in memory modes high and ultralow.  Bands
    that keep the end out of reach end in WF_STATUS_ALG_PARTIAL with an
    empty CIGAR for the gap metrics (END_UNREACHABLE after more null steps
    than the score scope, wavefront_extend.c:94-101) and WF_STATUS_UNATTAINABLE
    in BiWFA.  For indel and edit the reference's high-memory form crashes
    when a band empties the wavefront, so only pairs whose end diagonal lies
    in the band are compared there."""
    pats, texts = _banded_pairs()
    for h in BANDED:
        sel = list(range(len(pats)))
        if metric in ("edit", "indel"):
            # the oracle stops at once where the band empties the wavefront
            # (a step-limit result); those pairs crash the reference
            full = _pairs(pats, texts)
            r = O.oracle_align(O.Conf(metric, pen, "high", heuristic=h, max_steps=30000), full)
            sel = [i for i in sel if r.status[i] != O.WF_STATUS_MAX_STEPS_REACHED]
        if not sel:
            continue
        pairs = _pairs([pats[i] for i in sel], [texts[i] for i in sel])
        for mem in ("high", "ultralow"):
            for so in (False, True):
                conf = O.Conf(metric, pen, mem, so, heuristic=h, max_steps=30000)
                x = O.oracle_align(conf, pairs)
                y = O.ref_align(conf, pairs, isa="v2")
                assert list(x.status) == list(y.status), (h, mem, so)
                assert list(x.score) == list(y.score), (h, mem, so)
                if not so:
                    for i in range(pairs.n):
                        assert x.ops(i) == y.ops(i), (h, mem, i)


DROPS = [
    (0x40, 10, 50, 1),                   # wfmash
    (0x40, 10, 20, 3),
    (0x40 | 1, 10, 50, 1, -20, 20),      # wfmash then banded static
    (0x10, 0, 0, 1, 0, 0, 30),           # x-drop
    (0x10, 0, 0, 5, 0, 0, 100),
    (0x14, 10, 50, 1, 0, 0, 40),         # wf-adaptive then x-drop
    (0x11, 0, 0, 1, -10, 10, 40),        # x-drop then banded static
    (0x20, 0, 0, 1, 0, 0, 0, 30),        # z-drop
    (0x20, 0, 0, 3, 0, 0, 0, 100),
    (0x24, 10, 50, 1, 0, 0, 0, 40),      # wf-adaptive then z-drop
    (0x21, 0, 0, 1, -10, 10, 0, 40),     # z-drop then banded static
]


@pytest.mark.skipif(not O.ref_available("v2"), reason="scalar reference build (oracle/_ref) not present")
@pytest.mark.parametrize("metric,pen", [("affine", (-3, 5, 3, 2)), ("affine", (-1, 4, 6, 2)), ("affine", (0, 4, 6, 2)),
                                        ("affine2p", (-1, 4, 6, 2, 24, 1)), ("linear", (-2, 4, 2)),
                                        ("edit", (0, 1, 1))])
def test_oracle_wfmash_and_xdrop_match_reference_build(metric, pen):
    """wfmash (weighted distances in single-precision float,
    wavefront_heuristic.c:134-145, 190-208), x-drop and z-drop on
    Smith-Waterman scores (:294-431), alone and combined, in memory modes
    high and ultralow; a z-dropped alignment is traced from the best cell at
    the score it dropped at.  Drops with indel / edit are a configuration
    error in the reference (exit(1)), so edit runs wfmash only; an edit
    wavefront a band empties crashes the reference's high-memory form, so
    those pairs are left out there as in the banded test."""
    pats, texts = _banded_pairs()
    for h in DROPS:
        if metric == "edit" and h[0] & 0x30:
            continue
        sel = list(range(len(pats)))
        if metric == "edit":
            r = O.oracle_align(O.Conf(metric, pen, "high", heuristic=h, max_steps=30000), _pairs(pats, texts))
            sel = [i for i in sel if r.status[i] != O.WF_STATUS_MAX_STEPS_REACHED]
        pairs = _pairs([pats[i] for i in sel], [texts[i] for i in sel])
        for mem in ("high", "ultralow"):
            for so in (False, True):
                conf = O.Conf(metric, pen, mem, so, heuristic=h, max_steps=30000)
                x = O.oracle_align(conf, pairs)
                y = O.ref_align(conf, pairs, isa="v2")
                assert list(x.status) == list(y.status), (h, mem, so)
                assert list(x.score) == list(y.score), (h, mem, so)
                if not so:
                    for i in range(pairs.n):
                        assert x.ops(i) == y.ops(i), (h, mem, i)


def test_oracle_banded_static_is_banded_dynamic_programming():
    """A static band [min_k, max_k] that holds diagonals 0 and k_end limits
    the alignment to cells with min_k - 1 <= h - v <= max_k + 1 at most; the
    score equals the best over alignments inside [min_k, max_k] or improves
    on it only through the one-diagonal overhang, and never beats the DP
    optimum; a band wider than both sequences gives the exact result."""
    pats, texts = _random_pairs(91, 40, 30, 70, 10, max_indel=3)
    pairs = _pairs(pats, texts)
    exact = O.oracle_align(O.Conf("affine", (0, 4, 6, 2)), pairs)
    wide = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), heuristic=(1, 0, 0, 0, -200, 200)), pairs)
    tight = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), heuristic=(1, 0, 0, 0, -2, 2)), pairs)
    for i, (p, t) in enumerate(zip(pats, texts)):
        assert (wide.status[i], wide.score[i], wide.ops(i)) == (exact.status[i], exact.score[i], exact.ops(i))
        if not -2 <= len(t) - len(p) <= 2:
            continue
        assert tight.status[i] == 0 and dp.cigar_valid(p, t, tight.ops(i))
        assert dp.cigar_score(tight.ops(i), 0, 4, 6, 2) == tight.score[i] <= dp.gotoh(p, t, 0, 4, 6, 2)
        k, v, h = 0, 0, 0
        for op in tight.ops(i):
            v += op in b"MXD"
            h += op in b"MXI"
            assert -3 <= h - v <= 3


def test_oracle_wfadaptive_is_a_valid_alignment_no_better_than_optimal():
    """What a heuristic may change is which alignment is found, never its
    validity: the CIGAR transforms pattern into text, re-scores to the
    reported score, and that score never beats the DP optimum; with a
    threshold nothing can exceed it equals the exact result."""
    pats, texts = _random_pairs(77, 60, 40, 90, 12, max_indel=3)
    pairs = _pairs(pats, texts)
    r = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), heuristic=(4, 10, 5, 1)), pairs)
    exact = O.oracle_align(O.Conf("affine", (0, 4, 6, 2)), pairs)
    loose = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), heuristic=(4, 10, 100000, 1)), pairs)
    for i, (p, t) in enumerate(zip(pats, texts)):
        assert r.status[i] == 0
        assert dp.cigar_valid(p, t, r.ops(i))
        assert dp.cigar_score(r.ops(i), 0, 4, 6, 2) == r.score[i]
        assert r.score[i] <= dp.gotoh(p, t, 0, 4, 6, 2)
        assert (loose.score[i], loose.ops(i)) == (exact.score[i], exact.ops(i))
