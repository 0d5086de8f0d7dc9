"""The host parts of mh_refine_pairwise (anchor columns, pieces) against the
refinement oracle, bit-exact, on the CPU."""
import numpy as np
import pytest

import mauvehip
from oracle import refine_oracle as ro
from refine_cases import alignment


@pytest.mark.parametrize("seed", range(10))
def test_anchor_cols_match_oracle(seed):
    rng = np.random.default_rng(seed)
    r1, r2 = alignment(rng, int(rng.integers(30, 3000)), div=float(rng.choice([0.02, 0.1, 0.3])),
                       marble=float(rng.choice([0.0, 0.1, 0.5])), lower=bool(seed % 3 == 0))
    assert mauvehip.refine_anchor_cols(r1, r2) == ro.anchor_cols(r1, r2)


def test_anchor_cols_corners():
    for r1, r2 in [("A", "A"), ("A" * 21, "A" * 21), ("A" * 22, "A" * 22), ("-A-", "AA-"),
                   ("AC--GT" * 50, "ACGTGT" * 50), ("N" * 40, "A" * 40), ("A.C" * 30, "AGC" * 30)]:
        assert mauvehip.refine_anchor_cols(r1, r2) == ro.anchor_cols(r1, r2), (r1[:10], r2[:10])


@pytest.mark.parametrize("seed,big", [(0, 0), (1, 6000), (2, 5001), (3, 12000), (4, 0)])
def test_pieces_match_oracle(seed, big):
    rng = np.random.default_rng(40 + seed)
    r1, r2 = alignment(rng, int(rng.integers(2000, 30000)), div=0.2, marble=0.5, big_gap=big)
    got = mauvehip.refine_pieces(r1, r2)
    assert got == ro.pieces_of(r1, r2)
    assert got[0][0] == 0 and got[-1][1] == len(r1)
    assert all(p[1] == q[0] for p, q in zip(got, got[1:]))


def test_pieces_across_scan_chunks():
    # unaligned runs placed across the library's 2^18-column scan chunks
    rng = np.random.default_rng(91)
    n = 700_000
    r1 = np.full(n, ord("A"), dtype=np.uint8)
    r2 = np.full(n, ord("A"), dtype=np.uint8)
    for lo, ln in [(3000, 6000), (262144 - 2500, 5200), (524288 - 5001, 5001), (524288 + 9000, 4999),
                   (n - 5100, 5100)]:
        side = r2 if rng.random() < 0.5 else r1
        side[lo:lo + ln] = ord("-")
    a, b = r1.tobytes().decode(), r2.tobytes().decode()
    assert mauvehip.refine_pieces(a, b) == ro.pieces_of(a, b)


def test_pieces_across_prefix_chunks():
    """An alignment longer than the library's residue-count chunks (2^12
    blocks of 64 columns), with long one-sided gaps placed across a chunk
    border: the cuts and the density halving read counts at arbitrary columns."""
    rng = np.random.default_rng(91)
    r1, r2 = alignment(rng, 600_000, div=0.05, marble=0.3)
    c = 262_144 - 3000
    r1 = r1[:c] + "-" * 6000 + r1[c:]
    r2 = r2[:c] + "".join(rng.choice(list("ACGT"), 6000)) + r2[c:]
    got = mauvehip.refine_pieces(r1, r2)
    assert got == ro.pieces_of(r1, r2)
    assert len(got) > 20
