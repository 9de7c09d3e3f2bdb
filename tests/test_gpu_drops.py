"""wfmash (wf-adaptive on weighted distances, wavefront_heuristic.c:134-145,
190-208: single-precision float on the device as on the reference's host),
x-drop and z-drop (Smith-Waterman scores, :294-431; a z-dropped alignment is
traced from its best cell) on the GPU, alone and combined
with wf-adaptive and a band, in memory modes high and ultralow, bit-exact
with the oracle (which is pinned to the reference build).  As for the bands,
only the pairs the oracle flags as a backtrace with taken-back operations
are declined (WF_STATUS_UNSUPPORTED)."""
import pytest

from test_gpu_banded import check, pairs, say

pytestmark = pytest.mark.gpu

DROPS = [
    (0x40, 10, 50, 1),
    (0x40, 10, 20, 3),
    (0x40 | 1, 10, 50, 1, -20, 20),
    (0x10, 0, 0, 1, 0, 0, 30),
    (0x10, 0, 0, 5, 0, 0, 100),
    (0x14, 10, 50, 1, 0, 0, 40),
    (0x11, 0, 0, 1, -10, 10, 40),
    (0x20, 0, 0, 1, 0, 0, 0, 30),
    (0x20, 0, 0, 3, 0, 0, 0, 100),
    (0x24, 10, 50, 1, 0, 0, 0, 40),
    (0x21, 0, 0, 1, -10, 10, 0, 40),
]
METRICS = [("affine", (-3, 5, 3, 2)), ("affine", (-1, 4, 6, 2)), ("affine", (0, 4, 6, 2)),
           ("affine2p", (-1, 4, 6, 2, 24, 1)), ("linear", (-2, 4, 2)), ("edit", (0, 1, 1))]


@pytest.mark.parametrize("metric,pen", METRICS, ids=["%s%s" % m for m in METRICS])
def test_drops_vs_oracle(metric, pen):
    ps = pairs(4401)
    for heur in DROPS:
        if metric == "edit" and heur[0] & 0x30:
            continue  # drops with edit: exit(1) in the reference (wavefront_align.c:82-89)
        for mem in ("high", "ultralow"):
            for so in (False, True):
                say("drops %s %s %s %s so=%s" % (metric, pen, heur, mem, so))
                check(metric, pen, heur, ps, mem, so)


def test_drops_long_pairs_vs_oracle():
    """1-10 kbp pairs on the HBM-ring and BiWFA kernels."""
    from wfa_amd import datagen
    ps = datagen.generate(2, 4402, 10)
    ps = ps.subset([i for i in range(ps.n) if ps.plen[i] <= 10000][:6])
    for heur in ((0x40, 10, 50, 1), (0x10, 0, 0, 1, 0, 0, 200), (0x14, 10, 50, 1, 0, 0, 300),
                 (0x20, 0, 0, 1, 0, 0, 0, 200)):
        for mem in ("high", "ultralow"):
            say("drops long %s %s" % (heur, mem))
            check("affine", (-1, 4, 6, 2), heur, ps, mem, False)
