"""Seed patterns: known answers from the reference table (LM/SeedMasks.h) and
the library's CPU-side seed helpers against the oracle."""
import math

import pytest

from oracle import pyoracle as po


def test_known_seeds():
    assert po.get_seed(15, 3) == 0x16df6d          # LM/SeedMasks.h:132, the CLI's CODING_SEED w15
    assert po.get_seed(17, 3) == (1 << 17) - 1     # no rank-3 mask for 17 -> solid
    assert po.get_seed(19, 3) == (1 << 19) - 1
    assert po.get_seed(17, 0) == 0x6dbedb
    assert po.get_seed(21, 2) == 0x7eb76bf
    assert po.get_seed(11, 3) == 0x5bed
    assert po.get_seed(12, 0) == 0x7954f           # the table's weight-12 row starts with a weight-11 mask
    assert po.get_seed(40, 0) == (1 << 32) - 1     # weight > 31 -> solid 32
    assert po.get_seed(15, 7) == (1 << 15) - 1     # rank > 5 -> solid
    assert po.get_seed(2, 0) == 3                  # no_seeds rows -> solid


def test_length_weight_closed_form():
    for w in range(3, 32):
        for r in range(0, 6):
            s = po.get_seed(w, r)
            bits = bin(s)[2:]
            assert po.seed_length(s) == len(bits.strip("0")) if s else 0
            assert po.seed_weight(s) == bits.count("1")
    assert po.seed_length(0x16df6d) == 21 and po.seed_weight(0x16df6d) == 15
    # the CLI's coding seed is palindromic, so a window and its reverse
    # complement select the same characters (GetDnaSeedMer relies on it)
    b = bin(0x16df6d)[2:]
    assert b == b[::-1]


@pytest.mark.parametrize("avg", [0, 1, 10, 100, 200, 1000, 5000, 123456, 4_600_000, 10**9])
def test_default_weight(avg):
    if avg == 0:
        assert po.default_seed_weight(0) == 0
        return
    m = math.ceil((math.log(avg) / math.log(2.0)) / 1.5)
    if m % 2 == 0:
        m += 1
    m = 0 if m < 5 else m
    assert po.default_seed_weight(avg) == min(m, 31)


def test_library_seed_helpers_match_oracle():
    import mauvehip
    for w in range(0, 40):
        for r in range(0, 8):
            s = po.get_seed(w, r)
            assert mauvehip.get_seed(w, r) == s
            assert mauvehip.get_seed_length(s) == po.seed_length(s)
            assert mauvehip.get_seed_weight(s) == po.seed_weight(s)
    for avg in (0, 1, 200, 777, 10**6):
        assert mauvehip.get_default_seed_weight(avg) == po.default_seed_weight(avg)
