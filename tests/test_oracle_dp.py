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


# ---- banded region DP (no reference counterpart) ----------------------------

def _region_rows(la, lb, band, sw, s):
    if band == 0:
        return 1, la
    c0, c1 = s * sw + 1, min((s + 1) * sw, lb)
    lo = c0 - band
    r0 = 1 if lo <= 0 else max(1, -(-la * lo // lb))
    r1 = min(la, la * (c1 + band) // lb)
    return r0, r1


def _in_region(la, lb, band, sw, path):
    i = j = 0
    for e in path:
        i += e != "I"
        j += e != "D"
        if i >= 1 and j >= 1:
            r0, r1 = _region_rows(la, lb, band, sw, (j - 1) // sw)
            if not r0 <= i <= r1:
                return False
    return True


def test_region_whole_matrix_equals_nwsmall():
    """Band 0 is the whole matrix: the virtual-boundary integer DP reproduces
    the literal NWSmall restatement (and so libMUSCLE) path for path."""
    for a, b, ra, rb in _gold():
        p, sc = po.nwsmall(a, b)
        q, sq, dv = po.nwsmall_region(a, b, 0, 512)
        assert q == p and sq == int(sc), (a, b)
        assert dv == po.path_dev(len(a), len(b), p)
    rng = np.random.default_rng(11)
    for _ in range(40):
        a = rng.integers(0, 4, int(rng.integers(1, 200)), dtype=np.uint8)
        b = rng.integers(0, 4, int(rng.integers(1, 200)), dtype=np.uint8)
        p, sc = po.nwsmall(a, b)
        q, sq, _ = po.nwsmall_region(a, b, 0, 64)
        assert q == p and sq == int(sc)


@pytest.mark.parametrize("la,lb", [(1, 4), (4, 1), (3, 3), (4, 5), (5, 3), (2, 6), (6, 6)])
def test_region_brute_force_optimum(la, lb):
    """The banded result is the best path that stays in the region, and no
    path when the region holds none."""
    rng = np.random.default_rng(la * 7 + lb)
    for _ in range(6):
        a = rng.integers(0, 4, la, dtype=np.uint8)
        b = rng.integers(0, 4, lb, dtype=np.uint8)
        for band in (1, 2, 3):
            for sw in (1, 2, 3):
                inside = [q for q in all_paths(la, lb) if _in_region(la, lb, band, sw, q)]
                p, sc, dv = po.nwsmall_region(a, b, band, sw)
                if not inside:
                    assert p is None
                    continue
                assert p in inside
                assert po.path_score(a, b, p) == sc
                assert sc == max(po.path_score(a, b, q) for q in inside)
                assert dv == po.path_dev(la, lb, p)


def test_region_contains_full_path():
    """Whenever the full DP's path stays within distance < W of the line, the
    banded result is the full result exactly (score and path); otherwise its
    score can only be lower."""
    rng = np.random.default_rng(12)
    for _ in range(30):
        la = int(rng.integers(20, 300))
        a = rng.integers(0, 4, la, dtype=np.uint8)
        b = a.copy()
        hit = rng.random(la) < 0.15
        b[hit] = rng.integers(0, 4, int(hit.sum()), dtype=np.uint8)
        cut = int(rng.integers(0, la))
        b = np.concatenate([b[:cut], rng.integers(0, 4, int(rng.integers(0, 12)), dtype=np.uint8), b[cut + 3:]])
        if len(b) == 0:
            continue
        p, sc = po.nwsmall(a, b)
        d = po.path_dev(la, len(b), p)
        for sw in (4, 16, 256):
            q, sq, dq = po.nwsmall_region(a, b, d + 1, sw)
            assert q == p and sq == int(sc) and dq == d
            if d > 1:
                q2, sq2, _ = po.nwsmall_region(a, b, max(1, d // 2), sw)
                assert q2 is None or sq2 <= int(sc)


def _in_window(la, lb, band, width, path):
    i = j = 0
    for e in path:
        i += e != "I"
        j += e != "D"
        if i >= 1 and j >= 1:
            b = po.window_base(la, lb, band, width, i)
            if not b + 1 <= j <= b + width:
                return False
    return True


@pytest.mark.parametrize("la,lb", [(1, 4), (4, 1), (3, 3), (4, 5), (5, 3), (2, 6), (6, 6), (6, 2)])
def test_window_brute_force_optimum(la, lb):
    """The sliding-window result is the best path inside the window."""
    rng = np.random.default_rng(la * 13 + lb)
    for _ in range(6):
        a = rng.integers(0, 4, la, dtype=np.uint8)
        b = rng.integers(0, 4, lb, dtype=np.uint8)
        for band in (1, 2):
            for width in (2, 3, 4):
                inside = [q for q in all_paths(la, lb) if _in_window(la, lb, band, width, q)]
                p, sc, dv = po.nwsmall_window(a, b, band, width)
                if not inside:
                    assert p is None
                    continue
                assert p in inside
                assert po.path_score(a, b, p) == sc
                assert sc == max(po.path_score(a, b, q) for q in inside)
                assert dv == po.path_dev(la, lb, p)


def test_window_covers_band_and_contains_full_path():
    rng = np.random.default_rng(14)
    # coverage: every cell within h = (width-1)/2 of the line is in the window
    for la, lb, band in [(50, 80, 3), (80, 50, 5), (200, 230, 10), (30, 120, 4), (7, 9, 2)]:
        width = 4 * band + 2
        for i in range(1, la + 1):
            b = po.window_base(la, lb, band, width, i)
            for j in range(1, lb + 1):
                if abs(i * lb - j * la) <= ((width - 1) // 2) * la:
                    assert b + 1 <= j <= b + width, (la, lb, band, i, j)
    for _ in range(30):
        la = int(rng.integers(20, 300))
        a = rng.integers(0, 4, la, dtype=np.uint8)
        bb = a.copy()
        hit = rng.random(la) < 0.15
        bb[hit] = rng.integers(0, 4, int(hit.sum()), dtype=np.uint8)
        cut = int(rng.integers(0, la))
        bb = np.concatenate([bb[:cut], rng.integers(0, 4, int(rng.integers(0, 12)), dtype=np.uint8), bb[cut + 3:]])
        if len(bb) == 0:
            continue
        p, sc = po.nwsmall(a, bb)
        d = po.path_dev(la, len(bb), p)
        for width in (2 * d + 3, 2 * d + 40, 64):
            if (width - 1) // 2 < d + 1:
                continue
            q, sq, dq = po.nwsmall_window(a, bb, d + 1, width)
            assert q == p and sq == int(sc) and dq == d
