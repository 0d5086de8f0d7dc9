"""Oracle NWSmall against the reference libMUSCLE (golden vectors written by
tests/golden/gen_golden.py) and against brute-force enumeration of all paths."""
import itertools
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dp_golden.txt")


def _gold():
    for line in open(GOLD):
        if line.startswith("#"):
            continue
        a, b, ra, rb = line.split()
        yield a, b, ra, rb


def test_oracle_equals_reference_muscle():
    n = 0
    for a, b, ra, rb in _gold():
        p, _ = po.nwsmall(a, b)
        assert po.rows_from_path(a, b, p) == (ra, rb), (a, b)
        n += 1
    assert n == 260


def all_paths(la, lb, prev=""):
    """Every MUSCLE-valid edge sequence (a gap in A is never next to a gap in B)."""
    if la == 0 and lb == 0:
        yield ""
        return
    for e in "MDI":
        if e == "M" and (la == 0 or lb == 0):
            continue
        if e == "D" and (la == 0 or prev == "I"):
            continue
        if e == "I" and (lb == 0 or prev == "D"):
            continue
        for rest in all_paths(la - (e != "I"), lb - (e != "D"), e):
            yield e + rest


@pytest.mark.parametrize("la,lb", [(1, 1), (1, 2), (2, 1), (1, 5), (5, 1), (2, 2), (3, 4), (4, 3), (5, 5), (2, 6)])
def test_brute_force_optimum(la, lb):
    rng = np.random.default_rng(la * 31 + lb)
    for _ in range(25):
        a = rng.integers(0, 4, la, dtype=np.uint8)
        b = rng.integers(0, 4, lb, dtype=np.uint8)
        p, sc = po.nwsmall(a, b)
        # the path stays inside the model and its score is the reported terminal score
        assert "DI" not in p and "ID" not in p
        assert po.path_score(a, b, p) == sc
        best = max(po.path_score(a, b, q) for q in all_paths(la, lb))
        assert sc == best


def test_scores_are_integers_and_sentinel_free():
    rng = np.random.default_rng(9)
    for _ in range(200):
        la, lb = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        p, sc = po.nwsmall(rng.integers(0, 4, la, dtype=np.uint8), rng.integers(0, 4, lb, dtype=np.uint8))
        assert sc == int(sc) and abs(sc) < 1e7
        assert p.count("M") + p.count("D") == la and p.count("M") + p.count("I") == lb
