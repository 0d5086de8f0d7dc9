"""The oracle's NWSmall on operands with wildcard and other letters against
libMUSCLE's own ProfileProfile / NWSmall (tests/golden/dp_wild_golden.txt,
written by oracle/_ref/ref_harness): libMUSCLE scores such positions with
fractional profile counts in float (or_score6), so the path of every pair must
come out the same.  Paths are compared column by column (M both rows hold a
letter, D only row A, I only row B); libMUSCLE writes letters it does not know
back as 'N', which does not change a path."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    out = []
    for line in open(os.path.join(HERE, "golden", "dp_wild_golden.txt")):
        if line.startswith("#"):
            continue
        a, b, ra, rb = line.split()
        path = "".join("M" if x != "-" and y != "-" else ("D" if x != "-" else "I") for x, y in zip(ra, rb))
        out.append((a, b, path))
    return out


CASES = _cases()


def test_score_table():
    s = po.score6()
    hoxd = np.array([[91, -114, -31, -123], [-114, 100, -125, -31], [-31, -125, 100, -114], [-123, -31, -114, 91]])
    assert np.array_equal(s[:4, :4], (hoxd + 60).astype(np.float32))
    # X counts half G, half A (the DNA branch of GetFractionalWeightedCounts tests AX_R == NX_X)
    assert s[4, 0] == np.float32(0.5) * np.float32(151) + np.float32(0.5) * np.float32(29)
    assert s[4, 2] == np.float32(0.5) * np.float32(29) + np.float32(0.5) * np.float32(160)
    # any other wildcard 1/20 of each base, not normalised
    f = np.float32(1.0) / np.float32(20)
    want = np.float32(0)
    for j in range(4):
        want = np.float32(want + np.float32(f * np.float32(hoxd[0, j] + 60)))
    assert s[0, 5] == want


@pytest.mark.parametrize("i", range(len(CASES)))
def test_nwsmall_wildcards_match_libmuscle(i):
    a, b, path = CASES[i]
    got, _ = po.nwsmall(a, b)
    assert got == path
