"""The refinement oracle (oracle/refine_oracle.py) and the library's host-side
anchor columns against libMUSCLE's own code (tests/golden/refine_golden.txt,
written by oracle/_ref/ref_harness from MU/anchoredpp.cpp, MU/anchors.cpp,
MU/refinew.cpp and the rest of libMUSCLE under the globals MuscleInterface
sets):

* LetterObjScoreXP column scores, float bit for bit: gaps of 1 to 9 columns
  spread -400 in fractional shares, terminal gaps -200, wildcard letters;
* FindAnchorColsPP (smoothing in float, best columns, MergeBestCols);
* AnchoredProfileProfile's rows (ranges re-aligned by ProfileProfile, in
  float where a range holds wildcard letters);
* RefineW's rows for windows of 200 and 500 columns.
"""
import os

import numpy as np
import pytest

import mauvehip
from oracle import refine_oracle as ro

HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    lines = open(os.path.join(HERE, "golden", "refine_golden.txt")).read().splitlines()
    app, rw = [], []
    k = 1
    while k < len(lines):
        f = lines[k].split()
        if f[0] in ("APP", "WILD"):
            s = lines[k + 1].split()
            a = lines[k + 2].split()
            r = lines[k + 3].split()
            assert s[0] == "S" and a[0] == "A" and r[0] == "R"
            bits = np.array([int(x, 16) for x in s[2:]], dtype=np.uint32)
            app.append((f[0], f[1], f[2], bits.view(np.float32), [int(x) for x in a[2:]], r[1], r[2]))
            k += 4
        else:
            assert f[0] == "RW"
            o = lines[k + 1].split()
            rw.append((int(f[1]), f[2], f[3], o[1], o[2]))
            k += 2
    return app, rw


APP, RW = _cases()


def test_golden_counts():
    assert len(APP) >= 60 and len(RW) >= 20
    assert sum(1 for c in APP if c[0] == "WILD") >= 5
    # fractional gap shares occur (not every column score is an integer)
    assert any((s != np.round(s)).any() for _, _, _, s, _, _, _ in APP)


@pytest.mark.parametrize("i", range(len(APP)))
def test_column_scores_bit_exact(i):
    tag, r1, r2, want, _, _, _ = APP[i]
    got = np.array(ro.column_scores(r1, r2), dtype=np.float32)
    assert got.view(np.uint32).tolist() == want.view(np.uint32).tolist()


@pytest.mark.parametrize("i", range(len(APP)))
def test_anchor_columns(i):
    tag, r1, r2, _, want, _, _ = APP[i]
    assert ro.anchor_cols(r1, r2) == want
    assert mauvehip.refine_anchor_cols(r1, r2) == want


@pytest.mark.parametrize("i", range(len(APP)))
def test_anchored_profile_profile_rows(i):
    _, r1, r2, _, _, o1, o2 = APP[i]
    assert ro.anchored_profile_profile(r1, r2) == (o1, o2)


@pytest.mark.parametrize("i", range(len(RW)))
def test_refine_w_rows(i):
    w, r1, r2, o1, o2 = RW[i]
    assert ro.refine_w([r1, r2], w) == [o1, o2]
