"""Known answers for ExtendMatch with the spaced seed 0x16df6d (weight 15,
length 21, the CLI's weight-15 seed), worked by hand from
LM/MatchFinder.h:218-374, pinning the oracle's or_extend_seed.

The seed selects window slots {0,2,3,5,6,8,9,10,11,12,14,15,17,18,20}; slots
{1,4,7,13,16,19} are masked.  The pattern is palindromic and has odd weight,
so no selected string is its own reverse complement, and for these cases a
window "matches" (canonical keys equal and parities equal, 290-318) exactly
when the selected characters of the two windows are equal (forward seeds), or
equal after reverse-complementing genome 2's window (reverse seeds).  Every
derivation below uses only that rule and the four phases:
  0  jumps of 21 to the left while the next window matches;
  1  jumps of 21 to the right (after Invert);
  2  single steps to the left, t = 1..21, keeping the furthest t whose window
     matched (extend_limit), the others rolled back;
  3  the same to the right;
and a restart at phase 0 when phase 2 or 3 moved (extend_again).  maxlen clips
every phase at the sequence ends.  Positions are 0-based in the text, the
match starts 1-based as ExtendMatch keeps them.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

SEED = 0x16DF6D
SEL = [0, 2, 3, 5, 6, 8, 9, 10, 11, 12, 14, 15, 17, 18, 20]
COMP = bytes.maketrans(b"ACGT", b"TGCA")


def _core(n, seed):
    rng = np.random.default_rng(seed)
    return "".join("ACGT"[x] for x in rng.integers(0, 4, n))


def _sub(s, *pos):
    """s with a different base at each position (A<->C, G<->T)."""
    b = bytearray(s.encode())
    for p in pos:
        b[p] = {65: 67, 67: 65, 71: 84, 84: 71}[b[p]]
    return b.decode()


def _rc(s):
    return s.encode().translate(COMP)[::-1].decode()


def _extend(a, b, s0, s1):
    starts, length = po.extend_seed([a, b], SEED, [s0, s1], 21)
    return starts[0], starts[1], length


def test_pattern_slots():
    assert po.seed_length(SEED) == 21 and po.seed_weight(SEED) == 15
    assert [p for p in range(21) if (SEED >> (20 - p)) & 1] == SEL


def test_flanks_differ_everywhere():
    """A = A*40 + C + A*40, B = C*40 + C + C*40, |C| = 120, seed window at 90.
    Phase 0: windows at 69 and 48 match (inside C), 27 holds flank at slot 0:
    stop, left end 48.  Phase 1: windows at 111 and 132 match, 153 holds flank
    at slot 20: stop, right end 153.  Phase 2: steps t = 1..8 reach 40 (no
    flank); from t = 9 slot 0 is flank: limit 8, left end 40.  Phase 3: steps
    1..7 reach 160; from 8 slot 20 is flank: right end 160.  The restart moves
    nothing.  Result: exactly C, starts 41, length 120."""
    c = _core(120, 1)
    a = "A" * 40 + c + "A" * 40
    b = "C" * 40 + c + "C" * 40
    assert _extend(a, b, 91, 91) == (41, 41, 120)


def test_substitution_under_masked_slot_jumped_over():
    """A random 200, B = A with one substitution at 62, seed window at 100.
    Phase 0 windows 79, 58, 37, 16: 62 sits at slot 4 of window 58 (masked),
    so every jump matches and the left end is 16 after 4 jumps (maxlen 100 ->
    16).  Phase 1 reaches 184, phases 2 and 3 step to the ends (16 and 16
    steps, all inside identical sequence).  Result: the whole genome (a solid
    seed would stop at 62)."""
    a = _core(200, 2)
    b = _sub(a, 62)
    assert _extend(a, b, 101, 101) == (1, 1, 200)


def test_rollback_then_restart():
    """A random 200, B = A with one substitution at 60, seed window at 100.
    Phase 0: window 79 matches, window 58 holds 60 at slot 2 (selected): stop
    at 79.  Phase 1: right end 184.  Phase 2 from 79: t = 1..18 match, t = 19
    (window 60, slot 0) fails, t = 20 (window 59, slot 1, masked) matches,
    t = 21 (slot 2) fails: extend_limit 20, the failed 21st step rolled back:
    left end 59.  Phase 3 reaches 200.  extend_again: phase 0 from 59 jumps to
    38 and 17, phase 2 steps 17 to 0.  Result: the whole genome."""
    a = _core(200, 3)
    b = _sub(a, 60)
    assert _extend(a, b, 101, 101) == (1, 1, 200)


def test_substitution_next_to_flank():
    """A = A*40 + C + A*40, B = C*40 + C' + C*40, C' = C with a substitution at
    core offset 3 (position 43), seed window at 100.  Phase 0: 79, 58 match;
    37 holds flank.  Phase 1: right end 142 (window 142 holds flank at slot
    18).  Phase 2 from 58: t <= 14 match (windows from 44), t = 15 (window 43,
    slot 0) fails, t = 16 (window 42, the substitution at masked slot 1)
    matches, t = 17 and 18 (slots 2 and 3) fail, t >= 19 hold flank: left end
    42.  Phase 3: 18 steps to 160.  The restart finds nothing: windows 41 and
    40 hold 43 at slots 2 and 3, further left is flank.  Result: starts 43,
    length 118 (a solid seed would start at 45)."""
    c = _core(120, 4)
    a = "A" * 40 + c + "A" * 40
    b = "C" * 40 + _sub(c, 3) + "C" * 40
    assert _extend(a, b, 101, 101) == (43, 43, 118)


def test_walls_and_single_substitutions():
    """A random 300, B = A with substitutions at 40, 41 (a wall: the masked
    slots 1,4,7,13,16,19 hold no two neighbours, so every window over both
    fails, and windows 20 and 42 hold one of them at slot 20 or 0), 100 and
    180 (single: each lies at a masked slot of some window, which phase 2 or 3
    reaches, and the restart carries on past it) and 250, 251 (a wall).  Seed
    window at 130.  The left end stops at 42, the first window clear of the
    wall; the right end at 250.  Result: starts 43, length 208."""
    a = _core(300, 5)
    b = _sub(a, 40, 41, 100, 180, 250, 251)
    assert _extend(a, b, 131, 131) == (43, 43, 208)


def test_reverse_strand():
    """Genome 2 is the reverse complement of B of the previous case (length
    300).  A reverse seed pairs window w of A with the reverse complement of
    genome 2's window 300 - 21 - w, so every window test is the forward test
    of A against B: the match is the same stretch [42, 250) of A, which is
    [300 - 250, 300 - 42) = [50, 258) of genome 2, start 51, negated.
    Result: starts (43, -51), length 208."""
    a = _core(300, 5)
    b = _sub(a, 40, 41, 100, 180, 250, 251)
    g2 = _rc(b)
    assert _extend(a, g2, 131, -(300 - 21 - 130 + 1)) == (43, -51, 208)


@pytest.mark.parametrize("w", [0, 5, 20, 21, 279])
def test_sequence_ends(w):
    """Identical genomes of 300: from any seed window, jumps and steps are
    clipped by maxlen at both ends (start - 1 to the left, length - end to the
    right); the match is the whole genome."""
    a = _core(300, 6)
    assert _extend(a, a, w + 1, w + 1) == (1, 1, 300)


@pytest.mark.parametrize("w", [0, 7, 150, 279])
def test_sequence_ends_reverse(w):
    """Genome 2 = reverse complement of genome 1: the reverse match covers
    both whole genomes, starts (1, -1)."""
    a = _core(300, 7)
    assert _extend(a, _rc(a), w + 1, -(300 - 21 - w + 1)) == (1, -1, 300)


def test_walls_on_both_genome_ends_offset():
    """Genome 2 is genome 1's middle shifted: B = X(37) + A[60:260] + Y(25), with
    X, Y random.  A seed at A window 150 <-> B window 127.  The shared stretch
    is A[60:260] = B[37:237]; the bases around it differ at the first position
    outside (checked), and, as for the flanks, any window over a differing
    selected slot fails: slot 0 and slot 20 are selected, so the match is the
    shared stretch unless an outside window matches on its selected slots,
    which the checks below exclude.  Result: starts (61, 38), length 200."""
    a = _core(320, 8)
    rng = np.random.default_rng(9)
    x = "".join("ACGT"[i] for i in rng.integers(0, 4, 37))
    y = "".join("ACGT"[i] for i in rng.integers(0, 4, 25))
    b = x + a[60:260] + y
    # premises: the windows that straddle the stretch's ends differ at slot 0
    # (left) or slot 20 (right)
    assert a[59] != b[36] and a[260] != b[237]
    for t in range(1, 21):
        wa, wb = 60 - t, 37 - t
        if wa >= 0 and wb >= 0:
            assert any(a[wa + p] != b[wb + p] for p in SEL if wa + p < 60)
        wa, wb = 260 - 21 + t, 237 - 21 + t
        if wa + 21 <= len(a) and wb + 21 <= len(b):
            assert any(a[wa + p] != b[wb + p] for p in SEL if wa + p >= 260)
    assert _extend(a, b, 151, 128) == (61, 38, 200)
