"""GPU parity of mh_anchor_gaps (pairwiseAnchorSearch over a list of gaps)
with the oracle: the same anchors, in the same order, for every gap."""
import numpy as np
import pytest

import mauvehip
from mauvehip import synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _pair(seed, n):
    rng = np.random.default_rng(seed)
    a = synth.BASES[rng.integers(0, 4, n)]
    b = a.copy()
    hit = rng.random(n) < 0.04
    b[hit] = synth.BASES[rng.integers(0, 4, int(hit.sum()))]
    s = n // 3
    b[s:s + 2000] = synth.revcomp(b[s:s + 2000])
    return rng, a.tobytes().decode(), b.tobytes().decode()


def _gaps(rng, n, count):
    g = np.zeros(count, dtype=mauvehip.ANCHOR_GAP_DTYPE)
    for i in range(count):
        lo = int(rng.integers(1, n - 100))
        size = int(rng.choice([3, 40, 250, 1500, 6000]))
        shift = int(rng.integers(-30, 30))
        g[i]["lend"] = (lo, max(1, lo + shift))
        g[i]["rend"] = (min(n + 1, lo + size), min(n + 1, max(1, lo + shift) + size + int(rng.integers(-50, 50))))
    return g


def test_anchor_gaps_match_oracle(ctx):
    rng, A, B = _pair(31, 60000)
    gaps = _gaps(rng, len(A), 60)
    st, ln, first = mauvehip.anchor_gaps(ctx, A, B, gaps)
    total = 0
    for i, gp in enumerate(gaps):
        ws, wl = po.anchor_gap(A, B, tuple(int(x) for x in gp["lend"]), tuple(int(x) for x in gp["rend"]))
        s, e = int(first[i]), int(first[i + 1])
        assert np.array_equal(st[s:e], ws), i
        assert np.array_equal(ln[s:e], wl), i
        total += e - s
    assert total == len(ln) and total > 20


def test_anchor_gaps_empty_and_short(ctx):
    rng, A, B = _pair(32, 5000)
    g = np.zeros(3, dtype=mauvehip.ANCHOR_GAP_DTYPE)
    g[0]["lend"], g[0]["rend"] = (10, 10), (10, 10)          # nothing between the anchors
    g[1]["lend"], g[1]["rend"] = (100, 100), (120, 4000)      # genome 1 piece of 19 bases
    g[2]["lend"], g[2]["rend"] = (1, 1), (5001, 5001)         # whole genomes
    st, ln, first = mauvehip.anchor_gaps(ctx, A, B, g)
    for i, gp in enumerate(g):
        ws, wl = po.anchor_gap(A, B, tuple(int(x) for x in gp["lend"]), tuple(int(x) for x in gp["rend"]))
        assert np.array_equal(st[first[i]:first[i + 1]], ws)
        assert np.array_equal(ln[first[i]:first[i + 1]], wl)
