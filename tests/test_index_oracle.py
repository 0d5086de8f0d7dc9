"""Pins of the buildIndex oracle (oracle/index_oracle.py), CPU only.

* grammar: the reference's own fixture T/example.xmfa (tests/golden/), whose
  groups, entry fields and notes are read off the file by eye;
* the three passes: hand-worked known answers (each case says why);
* the column walk: alignments generated from a known edit script, where the
  expected map comes from the script's own coordinates, not from counting
  characters in the aligned rows.  The reference's walk starts both indices
  at start-1 and advances them before their first use
  (PY/buildindex.py:119-129), so a column pairing the 0-based positions
  (g, r) sets lut[g + 1] = r + 1 (index 0 is left to fixZeroIdx, and the
  last base of genome 1 falls outside the table).
"""
import os

import numpy as np
import pytest

from oracle import index_oracle as io

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_example_xmfa_grammar():
    groups = io.parse_xmfa(open(os.path.join(GOLDEN, "example.xmfa")).read())
    assert len(groups) == 2
    assert [n for _, n in groups] == [" score = 111", " score = 11"]
    e0 = [e[:5] for e in groups[0][0]]
    assert e0 == [(1, 1, 598, "+", "chrY"), (2, 5000, 5598, "-", "chr17"), (3, 19000, 19598, "-", "chr7")]
    e1 = [e[:5] for e in groups[1][0]]
    assert e1 == [(1, 1000, 1060, "+", "chrY"), (2, 6000, 6060, "+", "chr17"), (3, 20000, 20060, "-", "chr12")]
    # rows of a group have equal length, whitespace removed
    for entries, _ in groups:
        assert len({len(e[5]) for e in entries}) == 1
        assert all(set(e[5]) <= set("ACGTN-") for e in entries)
    assert groups[0][0][0][5].startswith("TCCAAGTCGGCTTTATG")
    assert groups[1][0][1][5].count("-") == 4


@pytest.mark.parametrize("lut,want,width", [
    ([5, 6, -1, -1, 9], [5, 6, 7, 8, 9], 300),                 # commensurate flanks: filled
    ([5, 6, -1, -1, 10], [5, 6, -1, -1, 10], 300),             # 10-6 != 4-1: left alone
    ([10, -1, -1, 13, 14], [10, 11, 12, 13, 14], 300),         # first commensurate upper wins
    ([0, -1, 2, -1, -1, 5], [0, 1, 2, 3, 4, 5], 300),          # fills chain left to right
    ([-1, -1, 3, 4], [-1, -1, 3, 4], 300),                     # no mapped left neighbour
    ([0, -1, -1, 3], [0, -1, -1, 3], 1),                       # upper 2 away from the gap start > width 1
    ([0, -1, 2], [0, 1, 2], 1),                                # upper 1 away: within width
    ([7, -1, 20, -1, 9], [7, -1, 20, -1, 9], 300),             # lower 20 at idx 2 -> 9 never fits
])
def test_fill_gaps_known(lut, want, width):
    lut = list(lut)
    io.fill_gaps(lut, width)
    assert lut == want


def test_fill_gaps_width_boundary():
    for gap, filled in ((300, True), (301, False)):
        lut = [0] + [-1] * gap + [gap + 1]
        io.fill_gaps(lut, 300)
        # the upper entry sits gap+1 - 1 = gap positions after the gap start
        assert (lut == list(range(gap + 2))) == filled


@pytest.mark.parametrize("lut,want,radius", [
    ([0, 1, 2, 50, 4, 5], [0, 1, 2, 3, 4, 5], 20),             # one wrong entry in a run
    ([0, 1, 2, 50, 60, 70, 6], [0, 1, 2, 3, 4, 5, 6], 20),     # three wrong entries
    ([0, 1, 2, 10, 11, 12, 13], [0, 1, 2, 10, 11, 12, 13], 20),  # a real shift (indel) stays
    ([0, 1, -1, 3], [0, 1, 2, 3], 20),                         # a hole is repaired too
    ([0, 1, 9, 9, 4], [0, 1, 9, 9, 4], 1),                     # upper 2 away from the break > radius 1
    ([-1, 5, 6, 0, 8], [-1, 5, 6, 7, 8], 20),                  # break after a mapped entry only
])
def test_smooth_edges_known(lut, want, radius):
    lut = list(lut)
    io.smooth_edges(lut, radius)
    assert lut == want


def test_fix_zero_idx_known():
    lut = [-1] * 10 + [15, 16]
    io.fix_zero_idx(lut, "A" * 12, "CCCCCA" + "C" * 20)       # ref[15-10] = 'A' == genome[0]
    assert lut[0] == 5
    lut = [-1] * 10 + [15, 16]
    io.fix_zero_idx(lut, "A" * 12, "C" * 26)                   # bases differ: unchanged
    assert lut[0] == -1
    lut = [-1, 3, 4, 5, 6, 7, 8, 9, 10, 11, -1, 8]             # mapped entries below 10 are passed over
    io.fix_zero_idx(lut, "G" * 12, "T" * 11 + "GG")            # k = 11: 8 - 11 < 0: unchanged
    assert lut[0] == -1
    lut = [3] + [-1] * 20
    io.fix_zero_idx(lut, "A", "A")                             # mapped already: unchanged
    assert lut[0] == 3
    with pytest.raises(IndexError):
        io.fix_zero_idx([-1, 1, 2, 3, 4, 5, 6, 7, 8, 9, -1, -1], "A" * 12, "A" * 12)


def _edit_script_alignment(rng, n):
    """Reference row R, genome row G and their aligned strings, with the
    expected map {genome pos: reference pos} of equal columns taken from the
    script's coordinates."""
    bases = np.array(list("ACGT"))
    R = list(bases[rng.integers(0, 4, n)])
    G, ga, ra, expect = [], [], [], {}
    i = 0
    while i < n:
        u = rng.random()
        if u < 0.05:                                  # deletion from genome: ref char against a gap
            ra.append(R[i]); ga.append("-"); i += 1
        elif u < 0.10:                                # insertion in genome
            c = str(bases[rng.integers(0, 4)])
            ga.append(c); ra.append("-"); G.append(c)
        else:
            c = R[i] if u < 0.9 else str(bases[(np.searchsorted(bases, R[i]) + 1) % 4])
            if c == R[i]:
                expect[len(G)] = i
            ga.append(c); ra.append(R[i]); G.append(c); i += 1
    return "".join(G), "".join(R), "".join(ga), "".join(ra), expect


def _wrap(s, w=80):
    return "\n".join(s[k:k + w] for k in range(0, len(s), w)) + "\n"


def test_column_walk_edit_script():
    rng = np.random.default_rng(21)
    G, R, ga, ra, expect = _edit_script_alignment(rng, 700)
    off_g, off_r = 37, 120            # the aligned block starts at these 0-based positions
    text = ("#FormatVersion Mauve1\n> 1:%d-%d + g.fa\n%s> 2:%d-%d + r.fa\n%s= score=1\n"
            % (off_g + 1, off_g + len(G), _wrap(ga), off_r + 1, off_r + len(R), _wrap(ra)))
    lut = io.lut_from_groups(io.parse_xmfa(text), off_g + len(G) + 10)
    want = [-1] * (off_g + len(G) + 10)
    for g, r in expect.items():
        want[off_g + g + 1] = off_r + r + 1
    assert lut == want


def test_column_walk_group_rules():
    # a group without sequence 2 maps nothing; entry order within a group does
    # not matter; sequence 3 is ignored; the first entry of a number counts
    text = ("> 1:1-4 + a\nACGT\n> 3:1-4 + c\nACGT\n=\n"
            "> 2:11-14 + b\nACGA\n> 1:5-8 + a\nACGT\n> 1:1-4 + a\nAAAA\n= x\n")
    lut = io.lut_from_groups(io.parse_xmfa(text), 10)
    assert lut == [-1, -1, -1, -1, -1, 11, 12, 13, -1, -1]


def test_seq_from_fasta():
    assert io.seq_from_fasta(">chr1 desc\nACGT\nAC GT\n\n>chr2\nTTTT\n") == "ACGTACGT"
    assert io.seq_from_fasta("no record") == ""
