"""The compiled host stages of the CPU baseline (oracle/cpu_path.cpp) against
the oracle modules they restate, on the same inputs:

* column scores and anchor columns bit for bit against libMUSCLE's own
  LetterObjScoreXP / FindAnchorColsPP (tests/golden/refine_golden.txt);
* the profile pass, the refinement pass and both against
  refine_oracle.refine_alignment / gapped_alignment (wildcards, lower case,
  one-sided stretches over 5000 columns, dense and sparse pieces);
* the gaps of the recursive anchor search against pyoracle.recursion_gaps;
* the XMFA text read back by index_oracle.parse_xmfa, and the LUT against
  index_oracle's column walk and its three post-passes;
* the NWSmall batch against one or_nwsmall call per job.
"""
import os

import numpy as np
import pytest

from oracle import index_oracle as io
from oracle import pyoracle as po
from oracle import refine_oracle as ro
from refine_cases import alignment

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden_app():
    lines = open(os.path.join(HERE, "golden", "refine_golden.txt")).read().splitlines()
    out = []
    k = 1
    while k < len(lines):
        f = lines[k].split()
        if f[0] in ("APP", "WILD"):
            s = lines[k + 1].split()
            a = lines[k + 2].split()
            bits = np.array([int(x, 16) for x in s[2:]], dtype=np.uint32)
            out.append((f[1], f[2], bits, [int(x) for x in a[2:]]))
            k += 4
        else:
            k += 2
    return out


GOLD = _golden_app()


@pytest.mark.parametrize("i", range(0, len(GOLD), 3))
def test_column_scores_and_anchors_match_libmuscle(i):
    r1, r2, bits, anchors = GOLD[i]
    assert po.c_column_scores(r1, r2).view(np.uint32).tolist() == bits.tolist()
    assert po.c_anchor_cols(r1, r2) == anchors


def _wild(rng, r, rate=0.01):
    b = bytearray(r.encode())
    for i in np.nonzero(rng.random(len(b)) < rate)[0]:
        if b[i] != ord("-"):
            b[i] = ord(rng.choice(list("NNXRYUunZ")))
    return b.decode()


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.choice([60, 400, 2500, 9000, 24000]))
        r1, r2 = alignment(rng, n, div=float(rng.choice([0.02, 0.1, 0.25])),
                           marble=float(rng.choice([0.0, 0.2, 0.6])),
                           big_gap=int(rng.choice([0, 0, 5600])) if n > 5000 else 0, lower=bool(k % 4 == 3))
        if k % 3 == 2:
            r1, r2 = _wild(rng, r1), _wild(rng, r2)
        out.append((r1, r2))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_refine_passes_match_refine_oracle(seed):
    for r1, r2 in _cases(seed, 6):
        assert po.c_refine_pass(r1, r2, True) == ro.refine_alignment(r1, r2, True)
        assert po.c_refine_pass(r1, r2, False) == ro.refine_alignment(r1, r2, False)
        g1, g2, cells = po.c_refine_gapped(r1, r2)
        assert (g1, g2) == ro.gapped_alignment(r1, r2)
        assert cells > 0


def test_refine_one_genome_only():
    for r1, r2 in (("A" * 30, "-" * 30), ("-" * 30, "ACGTU" * 6), ("ACGT" * 2000, "-" * 8000)):
        assert po.c_refine_gapped(r1, r2)[:2] == ro.gapped_alignment(r1, r2)


def _lut_oracle(text, L, s1, s2):
    lut = io.lut_from_groups(io.parse_xmfa(text), L)
    io.fix_zero_idx(lut, s1, s2)
    io.fill_gaps(lut, 300)
    io.smooth_edges(lut, 20)
    return lut


@pytest.mark.parametrize("seed", [4, 5])
def test_xmfa_lut_match_index_oracle(tmp_path, seed):
    rng = np.random.default_rng(seed)
    for n in (500, 3000, 12000):
        r1, r2 = alignment(rng, n, div=0.1, marble=0.3)
        s1, s2 = r1.replace("-", ""), r2.replace("-", "")
        path = str(tmp_path / "x.xmfa")
        lut, size = po.c_xmfa_lut(path, ("g1.fas", "g2.fas"), r1, r2, len(s1), len(s2), s1, s2)
        text = open(path).read()
        assert size == len(text.encode())
        groups = io.parse_xmfa(text)
        assert len(groups) == 1
        (e1, e2), _ = groups[0]
        assert (e1[0], e1[1], e1[2], e1[3], e1[5]) == (1, 1, len(s1), "+", r1)
        assert (e2[0], e2[1], e2[2], e2[3], e2[5]) == (2, 1, len(s2), "+", r2)
        assert lut.tolist() == _lut_oracle(text, len(s1), s1, s2)


def test_xmfa_lut_fills_and_smooths(tmp_path):
    """Unaligned stretches of 5..200 columns: fill_gaps and smooth_edges both
    change entries (the walk alone leaves -1 and off-by-shift runs)."""
    rng = np.random.default_rng(9)
    r1, r2 = alignment(rng, 20000, div=0.2, marble=0.6)
    s1, s2 = r1.replace("-", ""), r2.replace("-", "")
    lut, _ = po.c_xmfa_lut(str(tmp_path / "y.xmfa"), ("a", "b"), r1, r2, len(s1), len(s2), s1, s2)
    text = open(tmp_path / "y.xmfa").read()
    walk = io.lut_from_groups(io.parse_xmfa(text), len(s1))
    want = _lut_oracle(text, len(s1), s1, s2)
    assert walk != want
    assert lut.tolist() == want


def test_nwsmall_batch_matches_single_calls():
    rng = np.random.default_rng(11)
    la = rng.integers(1, 300, 40).astype(np.uint32)
    lb = rng.integers(1, 300, 40).astype(np.uint32)
    ao = np.concatenate([[0], np.cumsum(la)[:-1]]).astype(np.uint64)
    bo = np.concatenate([[0], np.cumsum(lb)[:-1]]).astype(np.uint64)
    A = rng.integers(0, 6, int(la.sum())).astype(np.uint8)
    B = rng.integers(0, 6, int(lb.sum())).astype(np.uint8)
    sc = np.zeros(40, dtype=np.float32)
    cells = po.c_nwsmall_batch(A, B, ao, la, bo, lb, 0, 40, sc)
    assert cells == int((la.astype(np.int64) * lb).sum())
    for k in range(40):
        _, want = po.nwsmall(A[ao[k]:ao[k] + la[k]], B[bo[k]:bo[k] + lb[k]])
        assert sc[k] == np.float32(want)


@pytest.mark.parametrize("seed", [12, 13])
def test_recursion_gaps_match_pyoracle(seed):
    rng = np.random.default_rng(seed)
    for n, mg in ((3000, 20), (30000, 200), (30000, 50)):
        r1, r2 = alignment(rng, n, div=0.05, marble=0.8)
        want = po.recursion_gaps(r1, r2, 1, 1, mg)
        assert po.c_recursion_gaps(r1, r2, 1, 1, mg) == want
        assert po.c_recursion_gaps(r1, r2, 7, 11, mg) == po.recursion_gaps(r1, r2, 7, 11, mg)
    assert any(po.recursion_gaps(*alignment(np.random.default_rng(s), 30000, div=0.05, marble=0.8), 1, 1, 50)
               for s in range(3))
