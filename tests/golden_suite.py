"""The reference's regression suite as data: tests/golden/wfa.utest.seq (305
pairs) and the golden .alg outputs copied from tests/wfa.utest.check/ of
WFA2-lib v2.3 (generated there by tests/wfa.utest.sh:35-64)."""
import os

from conftest import GOLDEN

PENALTIES = {
    "": (0, 4, 6, 2),
    ".p0": (0, 1, 2, 1),
    ".p1": (0, 3, 1, 4),
    ".p2": (0, 5, 3, 2),
    ".p3": (-5, 1, 2, 1),
    ".p4": (-2, 3, 1, 4),
    ".p5": (-3, 5, 3, 2),
}
AFFINE2P = (0, 4, 6, 2, 24, 1)
# prefix -> (memory mode, score_only), tests/wfa.utest.sh:39-43
MODES = {
    "test": ("high", False),
    "test.pb": ("med", False),
    "test.score": ("high", True),
    "test.biwfa": ("ultralow", False),
    "test.biwfa.score": ("ultralow", True),
}


def read_seq(path=None):
    path = path or os.path.join(GOLDEN, "wfa.utest.seq")
    pats, texts = [], []
    with open(path, "rb") as f:
        for line in f:
            line = line.rstrip(b"\r\n")
            if line.startswith(b">"):
                pats.append(line[1:])
            elif line.startswith(b"<"):
                texts.append(line[1:])
    return pats, texts


def golden_lines(name):
    with open(os.path.join(GOLDEN, name + ".alg")) as f:
        return [l.rstrip("\n") for l in f]


# wf-adaptive (strategy 4) with (min_wavefront_length, max_distance_threshold,
# steps_between_cutoffs), tests/wfa.utest.sh:57-58
HEURISTICS = {
    ".wfapt0": (4, 10, 50, 1),
    ".wfapt1": (4, 10, 50, 10),
}


def cases(metrics=("affine", "affine2p", "edit", "indel", "wfapt")):
    """(file stem, metric, penalties, memory, score_only, heuristic) for
    every golden file; heuristic is None (exact) or the wf-adaptive tuple."""
    out = []
    for pre, (mem, so) in MODES.items():
        if "affine" in metrics:
            for suf, pen in PENALTIES.items():
                out.append(("%s.affine%s" % (pre, suf), "affine", pen, mem, so, None))
        if "affine2p" in metrics:
            out.append(("%s.affine2p" % pre, "affine2p", AFFINE2P, mem, so, None))
        for m in ("edit", "indel"):
            if m in metrics:
                out.append(("%s.%s" % (pre, m), m, (0, 1, 1), mem, so, None))
        if "wfapt" in metrics:
            for suf, h in HEURISTICS.items():
                out.append(("%s.affine%s" % (pre, suf), "affine", (0, 4, 6, 2), mem, so, h))
    return out
