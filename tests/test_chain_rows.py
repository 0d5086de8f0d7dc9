"""The stand-in interval rows of a chain (mh_chain_rows) and the choice of
gaps for the recursive anchor search (mh_recursion_gaps) on the CPU.

Recursion gaps are checked against the oracle restatement of
LM/ProgressiveAligner.cpp:838-905 and a case worked by hand."""
import numpy as np
import pytest

import mauvehip
from oracle import pyoracle as po
from refine_cases import alignment


def _matches(rows):
    m = np.zeros(len(rows), dtype=mauvehip.MATCH_DTYPE)
    for i, (a, b, n) in enumerate(rows):
        m[i]["start"][:2] = (a, b)
        m[i]["length"] = n
    return m


def test_chain_rows_layout():
    A = b"AAAACCCCGGGGTTTT"
    B = b"TTAAAACCXGGGGT"
    m = _matches([(1, 3, 4), (9, 10, 4)])       # AAAA and GGGG
    ra, rb = mauvehip.chain_rows(m, [0, 1], A, B)
    assert ra == b"--AAAACCCC---GGGGTTTT-"
    assert rb == b"TTAAAA----CCXGGGG----T"
    assert ra.replace(b"-", b"") == A and rb.replace(b"-", b"") == B


def test_chain_rows_long_chain():
    """A chain of 3000 anchors (the rows are written in blocks of anchors in
    parallel): each stretch is genome 1's residues against gaps, then gaps
    against genome 2's, then the anchor; the last stretch runs to the ends."""
    rng = np.random.default_rng(5)
    A = rng.choice(list(b"ACGT"), 400_000).astype(np.uint8).tobytes()
    B = rng.choice(list(b"ACGT"), 410_000).astype(np.uint8).tobytes()
    rows, pa, pb = [], 0, 0
    for _ in range(3000):
        a = pa + int(rng.integers(0, 40))
        b = pb + int(rng.integers(0, 40))
        n = int(rng.integers(1, 60))
        rows.append((a + 1, b + 1, n))
        pa, pb = a + n, b + n
    m = _matches(rows)
    ra, rb = mauvehip.chain_rows(m, np.arange(len(rows)), A, B)
    wa, wb, pa, pb = [], [], 0, 0
    for a, b, n in rows + [(len(A) + 1, len(B) + 1, 0)]:
        a, b = a - 1, b - 1
        wa += [A[pa:a], b"-" * (b - pb), A[a:a + n]]
        wb += [b"-" * (a - pa), B[pb:b], B[b:b + n]]
        pa, pb = a + n, b + n
    want_a, want_b = b"".join(wa), b"".join(wb)
    assert ra == want_a and rb == want_b


def test_chain_rows_rejects_bad_chains():
    m = _matches([(5, 5, 10), (8, 30, 4)])      # overlapping in genome 1
    with pytest.raises(mauvehip.MhError):
        mauvehip.chain_rows(m, [0, 1], b"A" * 100, b"A" * 100)
    m = _matches([(5, -5, 10)])
    with pytest.raises(mauvehip.MhError):
        mauvehip.chain_rows(m, [0], b"A" * 100, b"A" * 100)


def test_recursion_gaps_known():
    # aligned 3 columns, 250 + 260 unaligned residues, aligned 3 columns
    ra = "AAA" + "C" * 250 + "-" * 260 + "GGG"
    rb = "AAA" + "-" * 250 + "T" * 260 + "GGG"
    # at the aligned column after the gap: 253 residues of A and 263 of B
    # before it, 2 and 2 at the previous aligned column (col 2)
    want = [((10 + 2, 20 + 2), (10 + 253, 20 + 263))]
    assert po.recursion_gaps(ra, rb, 10, 20) == want
    got = mauvehip.recursion_gaps(ra, rb, 10, 20)
    assert [(tuple(g["lend"]), tuple(g["rend"])) for g in got] == want
    # the counts include the residue of the previous aligned column: 199
    # residues between the aligned columns make 200, which is not more than 200
    rb2 = "AAA" + "-" * 250 + "T" * 199 + "-" * 61 + "GGG"
    assert len(mauvehip.recursion_gaps(ra, rb2, 1, 1)) == 0 == len(po.recursion_gaps(ra, rb2, 1, 1))


@pytest.mark.parametrize("seed", range(16))
def test_recursion_gaps_match_oracle(seed):
    rng = np.random.default_rng(70 + seed)
    ra, rb = alignment(rng, int(rng.integers(500, 20000)), div=0.1, marble=float(rng.choice([0.3, 1.0, 3.0])))
    la, lb = int(rng.integers(1, 1000)), int(rng.integers(1, 1000))
    mg = int(rng.choice([0, 7, 8, 9, 50, 200]))
    got = mauvehip.recursion_gaps(ra, rb, la, lb, mg)
    want = po.recursion_gaps(ra, rb, la, lb, mg)
    assert [(tuple(int(x) for x in g["lend"]), tuple(int(x) for x in g["rend"])) for g in got] == want


def test_recursion_gaps_across_chunks():
    """Rows longer than the library's 2^18-column scan chunks: gaps that open
    in one chunk and close in another, one spanning a whole chunk, and an
    aligned column right at a chunk border."""
    n = 700_000
    ra = bytearray(b"A" * n)
    rb = bytearray(b"A" * n)
    for lo, hi, row in [(262_000, 262_500, rb), (262_500, 263_000, ra), (300_000, 524_400, rb),
                        (524_400, 524_700, ra), (600_000, 600_250, ra), (600_250, 600_500, rb)]:
        row[lo:hi] = b"-" * (hi - lo)
    ra[524_288] = ord("-")
    ra, rb = bytes(ra), bytes(rb)
    got = mauvehip.recursion_gaps(ra, rb, 5, 9, 200)
    want = po.recursion_gaps(ra.decode(), rb.decode(), 5, 9, 200)
    assert len(want) >= 3
    assert [((g["lend"][0], g["lend"][1]), (g["rend"][0], g["rend"][1])) for g in got] == want
