"""Pins of the recursive-anchoring oracle (pyoracle.eliminate_overlaps,
pyoracle.anchor_gap), CPU only.

EliminateOverlaps_v2 cases are worked by hand from
LM/ProgressiveAligner.h:301-394 and the crop rules of
LM/UngappedLocalAlignment.h:138-184; the gap search is checked against what
the mathematics fixes: every anchor pairs similar stretches of the two
genomes (forward, or reverse complement for a negative start) inside its
gap, at least 12 long, and no two anchors overlap in either genome.  The
matches themselves come from the oracle's pinned pairwise search."""
import numpy as np
import pytest

from mauvehip import synth
from oracle import pyoracle as po


def _el(matches, min_len=0):
    st = np.array([m[:2] for m in matches], dtype=np.int64).reshape(len(matches), 2)
    ln = np.array([m[2] for m in matches], dtype=np.uint64)
    s, l = po.eliminate_overlaps(st, ln, min_len)
    return [(int(a), int(b), int(c)) for (a, b), c in zip(s, l)]


def test_shorter_loses_overlap():
    # overlap of 5 in genome 0; the later match is longer: the earlier loses its right end
    assert _el([(10, 100, 20), (25, 200, 30)]) == [(10, 100, 15), (25, 200, 30)]
    # the earlier is longer: the later loses its left end (forward: start moves right)
    assert _el([(10, 100, 30), (35, 200, 20)]) == [(10, 100, 30), (40, 205, 15)]


def test_equal_lengths_crop_the_later():
    assert _el([(10, 100, 20), (25, 200, 20)]) == [(10, 100, 20), (30, 205, 15)]


def test_contained_match_is_dropped():
    assert _el([(10, 100, 50), (20, 300, 10)]) == [(10, 100, 50)]
    assert _el([(20, 300, 10), (10, 100, 50)]) == [(10, 100, 50)]


def test_reverse_strand_crops():
    # earlier match reverse in genome 1; cropping its right end in genome 0 is
    # CropEnd: the reverse start moves by the overlap (-100 -> -105)
    assert _el([(10, -100, 20), (25, 200, 30)]) == [(10, -105, 15), (25, 200, 30)]
    # reverse in genome 0 itself: CropRight in genome 0 is CropStart there,
    # which moves the forward start of genome 1
    assert _el([(-10, 100, 20), (25, 300, 30)]) == [(-10, 105, 15), (25, 300, 30)]


def test_second_genome_pass_and_order():
    # no overlap in genome 0; in genome 1 (100..133 and 130..) they overlap by
    # 4 and the earlier one is longer, so the later loses 4 from its left end;
    # the result comes out sorted by genome 1
    got = _el([(500, 130, 20), (10, 100, 34)])
    assert got == [(10, 100, 34), (504, 134, 16)]


def test_length_filter():
    assert _el([(10, 100, 11), (50, 300, 12)], min_len=12) == [(50, 300, 12)]


def _check_anchor(a, b, st, ln, lend, rend):
    for (s0, s1), n in zip(st, ln):
        n = int(n)
        assert n >= 12
        x = a[abs(s0) - 1:abs(s0) - 1 + n]
        y = b[abs(s1) - 1:abs(s1) - 1 + n]
        same = (s0 > 0) == (s1 > 0)
        if not same:
            y = synth.revcomp(np.frombuffer(y.encode(), dtype=np.uint8)).tobytes().decode()
        # spaced seeds extend across mismatches at their free positions
        assert sum(p == q for p, q in zip(x, y)) >= 0.75 * n
        for g, s in enumerate((s0, s1)):
            assert lend[g] <= abs(s) and abs(s) + n - 1 <= rend[g] - 2
    for g in range(2):
        iv = sorted((abs(s[g]), abs(s[g]) + int(n)) for s, n in zip(st, ln))
        assert all(p[1] <= q[0] for p, q in zip(iv, iv[1:]))


def test_anchor_gap_properties():
    rng = np.random.default_rng(12)
    a = synth.BASES[rng.integers(0, 4, 20000)]
    b = a.copy()
    hit = rng.random(len(b)) < 0.05
    b[hit] = synth.BASES[rng.integers(0, 4, int(hit.sum()))]
    b[8000:9000] = synth.revcomp(b[8000:9000])
    A, B = a.tobytes().decode(), b.tobytes().decode()
    total = 0
    for lo, hi in [(1, 3000), (5000, 12000), (100, 150), (15000, 20001)]:
        lend, rend = (lo, lo + 7), (hi, hi + 3)
        st, ln = po.anchor_gap(A, B, lend, rend)
        _check_anchor(A, B, st, ln, lend, rend)
        total += len(ln)
    assert total > 3
    # pieces too short for a weight-5 seed: nothing
    st, ln = po.anchor_gap(A, B, (10, 10), (14, 14))
    assert len(ln) == 0
