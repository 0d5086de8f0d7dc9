"""Pins of the refinement oracle (oracle/refine_oracle.py), CPU only.

Column scores are worked by hand from MU/anchoredpp.cpp:16-316 (NUC_SP =
HOXD + 60 on letter pairs; gap opens -400, terminal -200, extension 0; a
gap's total spread evenly over its columns); the smoothing is checked on
integer scores against one division of the exact window sum; MergeBestCols,
ColsToRanges, the pieces of removeLargeGapsPP and the density rule against
cases computed by hand; the re-aligned pieces against the oracle NWSmall.
"""
import numpy as np
import pytest

from oracle import refine_oracle as ro
from oracle import pyoracle as po

F = np.float32


@pytest.mark.parametrize("r1,r2,want", [
    ("ACGT", "ACGT", [151, 160, 160, 151]),
    ("AC-GT", "ACTGT", [151, 160, -400, 160, 151]),        # open inside: the column takes -400
    ("-CGT", "ACGT", [-200, 160, 160, 151]),               # open at the first column: terminal -200
    ("ACG-", "ACGT", [151, 160, 160, -200]),               # open at the end: -400 replaced by -200
    ("A--T", "ACGT", [151, -200, -200, 151]),              # -400 over two columns
    ("A-CT", "AG-T", [151, 0, -800, 151]),                 # gap changes rows: both opens on the later column
    ("A-T", "A-T", [151, 0, 151]),                         # all-gap column: skipped
    ("ANT", "ACT", [151, 0, 151]),                         # wildcard: no letter score
    ("aCgT", "AcGt", [151, 160, 160, 151]),                # case does not matter
    ("AAAA", "TTTT", [-63, -63, -63, -63]),                # -123 + 60
])
def test_column_scores_known(r1, r2, want):
    assert [float(x) for x in ro.column_scores(r1, r2)] == want


def test_window_smooth_is_window_sum_over_21():
    rng = np.random.default_rng(3)
    for n in (5, 21, 22, 40, 300):
        s = [F(v) for v in rng.integers(-900, 200, n)]
        got = ro.window_smooth(s)
        if n <= 21:
            assert all(v == 0 for v in got)
            continue
        for i in range(n):
            if i < 10 or i >= n - 10:
                assert got[i] == 0
            else:
                exact = F(sum(int(v) for v in s[i - 10:i + 11]))
                assert got[i] == exact / F(21)


def test_merge_best_cols_known():
    score = [F(0)] * 400
    score[5], score[50] = F(150), F(160)
    assert ro.merge_best_cols(score, [5, 50, 200]) == [50, 200]          # two within 96: higher score
    score[5] = F(161)
    assert ro.merge_best_cols(score, [5, 50, 200]) == [5, 200]
    score[5] = F(160)
    assert ro.merge_best_cols(score, [5, 50, 200]) == [50, 200]          # tie: the later one
    assert ro.merge_best_cols(score, [5, 10, 20, 300]) == [10, 300]      # three within 96: the second
    assert ro.merge_best_cols(score, [0, 96, 192]) == [0, 96, 192]       # exactly 96 apart: separate
    assert ro.merge_best_cols(score, []) == []


def test_cols_to_ranges():
    assert ro.cols_to_ranges([3, 10], 15) == [(0, 3), (3, 10), (10, 15)]
    assert ro.cols_to_ranges([], 7) == [(0, 7)]
    assert ro.cols_to_ranges([0], 7) == [(0, 0), (0, 7)]


def test_best_cols_identity_block():
    # identical rows: every column scores >= 151, so every column whose full
    # window lies inside scores (avg >= 151) -> best cols are 10 .. n-11
    seq = "".join(np.random.default_rng(1).choice(list("ACGT"), 300))
    s = ro.column_scores(seq, seq)
    b = ro.best_cols(s, ro.window_smooth(s))
    assert b == list(range(10, 290))
    a = ro.merge_best_cols(s, b)
    assert a[0] in (10, 11) and all(y - x >= 1 for x, y in zip(a, a[1:]))


def test_remove_large_gaps_known():
    r1 = "A" * 100 + "C" * 6000 + "G" * 100
    r2 = "A" * 100 + "-" * 6000 + "G" * 100
    # first aligned column after the 6000-column gap is 6100: the aligned part
    # keeps 300 columns of the gap, the unaligned piece ends 300 before 6100
    assert ro.remove_large_gaps(r1, r2) == [(0, 399, False), (399, 5800, True), (5800, 6200, False)]
    r2b = "-" * 5500 + "A" * 10
    assert ro.remove_large_gaps("A" * 5510, r2b) == [(0, 5200, True), (5200, 5510, False)]
    assert ro.remove_large_gaps("A" * 10 + "C" * 5001, "A" * 10 + "-" * 5001) == [(0, 309, False), (309, 5011, True)]
    assert ro.remove_large_gaps("ACGT" * 10, "ACGT" * 10) == [(0, 40, False)]


def test_density_known():
    r = "A" * 1000
    assert ro.dense_enough(r, r, 0, 1000) == 2                          # density 1 > 0.9
    half = "A" * 500 + "-" * 500
    # density = 1500 / 2000 = 0.75; threshold(1000) = 0.4/-19800*(1000-20000)+0.5 = 0.8838...
    assert ro.dense_enough(r, half, 0, 1000) == 0
    # at 19000 columns the threshold is 0.5202...: density 0.75 -> 1
    assert ro.dense_enough("A" * 19000, "A" * 9500 + "-" * 9500, 0, 19000) == 1


def test_pieces_halving():
    # 7000 columns of density 0 are halved (7000 > 20000/3), 3500 + 3500
    r1 = "A" * 7000
    r2 = "A" * 2000 + "-" * 4800 + "A" * 200
    assert ro.dense_enough(r1, r2, 0, 7000) == 0
    assert ro.pieces_of(r1, r2) == [(0, 3500, False), (3500, 7000, False)]


def test_anchored_pp_realigns_ranges():
    rng = np.random.default_rng(5)
    a = "".join(rng.choice(list("ACGT"), 400))
    b = a[:150] + "TTGCA" + a[150:]
    # marbled layout of the insertion: A's residues against gaps first
    r1 = a[:150] + "-----" + a[150:]
    r2 = b
    o1, o2 = ro.anchored_profile_profile(r1, r2)
    assert o1.replace("-", "") == a and o2.replace("-", "") == b
    anchors = ro.anchor_cols(r1, r2)
    want1, want2 = [], []
    for lo, hi in ro.cols_to_ranges(anchors, len(r1)):
        if hi > lo:
            path, _ = po.nwsmall(r1[lo:hi].replace("-", ""), r2[lo:hi].replace("-", ""))
            x, y = po.rows_from_path(r1[lo:hi].replace("-", ""), r2[lo:hi].replace("-", ""), path)
            want1.append(x)
            want2.append(y)
    assert (o1, o2) == ("".join(want1), "".join(want2))


def test_refine_w_windows():
    rng = np.random.default_rng(6)
    a = "".join(rng.choice(list("ACGT"), 450))
    b = "".join(rng.choice(list("ACGT"), 430))
    r1 = a + "-" * 30
    r2 = b[:200] + "-" * 50 + b[200:]
    o = ro.refine_w([r1, r2], 200)
    parts = []
    for f in range(0, 480, 200):
        s1, s2 = r1[f:f + 200].replace("-", ""), r2[f:f + 200].replace("-", "")
        parts.append(ro.nwsmall_rows(s1, s2))
    assert o == ["".join(p[0] for p in parts), "".join(p[1] for p in parts)]
    # a window with residues on one side only: that side against gaps
    o = ro.refine_w(["A" * 100 + "C" * 100, "G" * 100 + "-" * 100], 100)
    assert o[0].endswith("C" * 100) and o[1].endswith("-" * 100)
    assert o[0][:-100].replace("-", "") == "A" * 100 and o[1][:-100].replace("-", "") == "G" * 100


def test_anchored_range_one_sided_is_fatal():
    with pytest.raises(ValueError):
        ro.nwsmall_rows("", "ACGT")
