"""CPU checks of tests/golden/long_pairs.json (the reference's results for
long pairs, used by the GPU tests): the seeded generator still produces the
pairs the fixtures were written for, every CIGAR is a valid alignment of its
pair that re-scores to the recorded score, and the oracle's restatement
agrees with the reference on the pairs it finishes in seconds."""
import hashlib
import json
import os

import pytest

import oracle as O
from conftest import GOLDEN
from wfa_amd import datagen

FIX = json.load(open(os.path.join(GOLDEN, "long_pairs.json")))["sets"]


def unrle(s):
    out, num = [], ""
    for ch in s:
        if ch.isdigit():
            num += ch
        else:
            out.append(ch * int(num))
            num = ""
    return "".join(out).encode()


def pairs_of(name):
    s = FIX[name]
    full = datagen.generate(s["shape"], s["seed"], s["first_n"])
    return full, s


def transforms(pattern, text, ops):
    """The CIGAR turns the pattern into the text (alignment/cigar.c:347-418)."""
    i = j = 0
    for op in ops:
        c = chr(op)
        if c == "M":
            if pattern[i] != text[j]:
                return False
            i += 1
            j += 1
        elif c == "X":
            if pattern[i] == text[j]:
                return False
            i += 1
            j += 1
        elif c == "D":
            i += 1
        elif c == "I":
            j += 1
        else:
            return False
    return i == len(pattern) and j == len(text)


@pytest.mark.parametrize("name", sorted(FIX))
def test_fixture_pairs_regenerate_and_cigars_are_valid(name):
    full, s = pairs_of(name)
    pen = tuple(s["penalties"])
    for p in s["pairs"]:
        i = p["index"]
        assert (int(full.plen[i]), int(full.tlen[i])) == (p["plen"], p["tlen"])
        ops = unrle(p["cigar"])
        assert hashlib.sha256(ops).hexdigest() == p["sha256"]
        assert p["status"] == 0
        assert transforms(full.pattern(i), full.text(i), ops), i
        conf = O.Conf(s["metric"], pen, "high")
        assert O.cigar_score(conf, ops) == p["score"], i


def test_c3_pairs_are_full_size():
    full, s = pairs_of("c3")
    for p in s["pairs"]:
        assert p["tlen"] == 100000 and p["memory"] == "ultralow"


def test_oracle_equals_reference_on_serving_mix_memory_high_pairs():
    full, s = pairs_of("c4_mix")
    sel = [p for p in s["pairs"] if p["memory"] != "ultralow"]
    sub = full.subset([p["index"] for p in sel])
    r = O.oracle_align(O.Conf("affine", (0, 4, 6, 2), "high"), O.Pairs(sub.patterns(), sub.texts()))
    for j, p in enumerate(sel):
        assert (int(r.status[j]), int(r.score[j])) == (p["status"], p["score"]), p["index"]
        assert hashlib.sha256(r.ops(j)).hexdigest() == p["sha256"], p["index"]


def test_oracle_equals_reference_on_affine2p_tail_pairs():
    """The tail pairs the oracle aligns in memory mode high in about a second
    each (the others take up to 5 GB of stored wavefronts)."""
    full, s = pairs_of("c2_tail")
    sel = [p for p in s["pairs"] if abs(p["score"]) <= 7000]
    assert len(sel) >= 4
    sub = full.subset([p["index"] for p in sel])
    r = O.oracle_align(O.Conf("affine2p", (0, 4, 6, 2, 24, 1), "high"), O.Pairs(sub.patterns(), sub.texts()))
    for j, p in enumerate(sel):
        assert (int(r.status[j]), int(r.score[j])) == (p["status"], p["score"]), p["index"]
        assert hashlib.sha256(r.ops(j)).hexdigest() == p["sha256"], p["index"]
