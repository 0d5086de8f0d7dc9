"""SearchRange's repeat-limit restart (LM/MatchFinder.cpp:166-340, GetBreakpoint
89-126): the library's restart planner (mh_search_live_ranges, host only)
against the oracle's literal merge with its buffers and restarts.

The planner turns the restarts into the index ranges of each sorted list that
the search reads; the seed tuples are then every masked mer with exactly one
live entry in two lists.  The oracle runs the reference's loop entry by entry
(10,000-entry buffers, the idmer list, the restart and GetBreakpoint quirks)
and records its HashMatch calls; both must give the same tuples in the same
order and the same restart count.  The oracle's restart itself is pinned by
the hand-worked cases in test_oracle_repeat.py.
"""
import numpy as np
import pytest

import mauvehip
from mauvehip import synth
from oracle import pyoracle as po


def _live_tuples(seqs, seed):
    k = len(seqs)
    w = po.seed_weight(seed)
    keys, pos, lens = [], [], []
    for s in seqs:
        _, kk, pp = po.sml_build(s, seed)
        keys.append(kk)
        pos.append(pp)
        lens.append(len(s))
    live, restarts = mauvehip.search_live_ranges(keys, lens, seed)
    ent = []
    for g in range(k):
        mask = np.zeros(len(keys[g]), dtype=bool)
        for lo, hi in live[g]:
            assert 0 <= lo < hi <= len(keys[g])
            mask[lo:hi] = True
        idx = np.nonzero(mask)[0]
        for i in idx:
            ent.append((int(keys[g][i]) >> (64 - 2 * w), g, int(pos[g][i])))
    ent.sort(key=lambda e: (e[0], e[1]))
    out = []
    i = 0
    while i < len(ent):
        j = i
        while j < len(ent) and ent[j][0] == ent[i][0]:
            j += 1
        cnt = {}
        for e in ent[i:j]:
            cnt.setdefault(e[1], []).append(e[2])
        uniq = sorted((g, p[0]) for g, p in cnt.items() if len(p) == 1)
        for a in range(len(uniq)):
            for b in range(a + 1, len(uniq)):
                t = [0] * k
                t[uniq[a][0]] = uniq[a][1] + 1
                t[uniq[b][0]] = uniq[b][1] + 1
                out.append(t)
        i = j
    return np.array(out, dtype=np.int64).reshape(-1, k), restarts, live


def _check(seqs, seed, min_restarts=1):
    seqs = [np.asarray(s).tobytes() if not isinstance(s, bytes) else s for s in seqs]
    got, restarts, live = _live_tuples(seqs, seed)
    want, _ = po.seed_tuples(seqs, seed)
    _, _, ost = po.find_matches(seqs, seed)
    assert restarts == ost.restarts
    assert restarts >= min_restarts
    assert got.shape == want.shape
    assert np.array_equal(got, np.abs(want))
    return restarts, live


def _rand(rng, n):
    return synth.BASES[rng.integers(0, 4, n)]


def _hp(base, n):
    return np.full(n, ord(base), dtype=np.uint8)


def _tandem(unit, copies):
    return np.frombuffer((unit * copies).encode(), dtype=np.uint8).copy()


SEEDS = [po.get_seed(15, 3), po.get_seed(17, 3), po.get_seed(11, 0), po.get_seed(9, 0)]


@pytest.mark.parametrize("seed", SEEDS)
def test_poly_a_one_genome(seed):
    """A 1200 bp poly-A run in one genome: masked mer 0 leads every list; the
    run is one block that crosses no buffer end, so it is enumerated (no
    restart), unless the other genome holds some of it."""
    rng = np.random.default_rng(1)
    a = np.concatenate([_rand(rng, 4000), _hp("A", 1200), _rand(rng, 4000)])
    b = a.copy()
    b[:4000] = _rand(rng, 4000)
    _check([a, b], seed, min_restarts=0)


@pytest.mark.parametrize("seed", SEEDS)
def test_poly_a_both_genomes(seed):
    """Poly-A in both genomes (split 700 / 600): the second block's check sees
    more than 1000 entries only through buffer ends, or the first block alone."""
    rng = np.random.default_rng(2)
    base = _rand(rng, 20000)
    a = np.concatenate([base[:9000], _hp("A", 1400), base[9000:]])
    b = np.concatenate([base[:5000], _hp("A", 700), base[5000:15000], _hp("T", 650), base[15000:]])
    _check([a, b], seed, min_restarts=0)


@pytest.mark.parametrize("seed", SEEDS)
def test_tandem_repeats_restart(seed):
    """Long (AC)n, (ACG)n and (AGCT)n tandems: heavy runs in the middle of the
    mer order, several restarts, GetBreakpoint jumps in both lists."""
    rng = np.random.default_rng(3)
    a = np.concatenate([_rand(rng, 6000), _tandem("AC", 1500), _rand(rng, 3000), _tandem("ACG", 1200),
                        _rand(rng, 3000), _tandem("AGCT", 1100), _rand(rng, 6000)])
    b = np.concatenate([_rand(rng, 2000), a[:6000], _tandem("AC", 900), a[9000:18000], _tandem("CGT", 1300),
                        a[21600:]])
    r, _ = _check([a, b], seed)
    assert r >= 2


@pytest.mark.parametrize("trial", range(12))
def test_runs_at_buffer_ends(trial):
    """Homopolymer runs placed so that their place in the sorted order falls
    across the 10,000-entry buffer ends of either list: the restart fires, or
    not, on the refill checks."""
    rng = np.random.default_rng(100 + trial)
    seed = SEEDS[trial % len(SEEDS)]
    la, lb = int(rng.integers(20000, 40000)), int(rng.integers(20000, 40000))
    a, b = _rand(rng, la), _rand(rng, lb)
    for s in (a, b):
        for base in "ACGT":
            n = int(rng.integers(300, 900))
            at = int(rng.integers(0, len(s) - n))
            s[at:at + n] = ord(base)
    shared = int(rng.integers(2000, 6000))
    b[:shared] = a[:shared]
    _check([a, b], seed, min_restarts=0)


# placements whose poly-C run straddles a buffer end of a list: in the first
# four the restart fires on a refill check inside the second block (each block
# alone holds fewer than 1000 entries), in the other four it does not
_FIRE = [32, 60, 260, 266]
_NO_FIRE = [2, 16, 20, 26]


@pytest.mark.parametrize("trial", _FIRE + _NO_FIRE)
def test_runs_at_buffer_ends_fire(trial):
    rng = np.random.default_rng(10000 + trial)
    seed = SEEDS[trial % 2]
    a, b = _rand(rng, int(rng.integers(15000, 45000))), _rand(rng, int(rng.integers(15000, 45000)))
    for s in (a, b):
        n = int(rng.integers(450, 700))
        at = int(rng.integers(0, len(s) - n))
        s[at:at + n] = ord("C")
    r, _ = _check([a, b], seed, min_restarts=0)
    assert (r > 0) == (trial in _FIRE)


@pytest.mark.parametrize("k,seed", [(3, po.get_seed(15, 3)), (3, po.get_seed(11, 0)), (4, po.get_seed(17, 3)),
                                    (4, po.get_seed(9, 0))])
def test_three_and_four_genomes(k, seed):
    """Heavy runs over three and four lists: block orders with shared and
    different previous mers, initial heads, lists closed by a restart."""
    rng = np.random.default_rng(40 + k)
    base = _rand(rng, 30000)
    seqs = []
    for g in range(k):
        s = base.copy()
        hit = rng.random(len(s)) < 0.02
        s[hit] = _rand(rng, int(hit.sum()))
        at = int(rng.integers(1000, 25000))
        rep = ["AC", "AGT", "A", "CG"][g % 4]
        ins = _tandem(rep, int(rng.integers(300, 700)) * (2 if len(rep) == 1 else 1))
        seqs.append(np.concatenate([s[:at], ins, s[at:]]))
    _check(seqs, seed, min_restarts=0)


def test_identical_genomes_with_repeat():
    """Two identical genomes: every run holds both lists with the same previous
    mer, so the block order alternates along the whole list."""
    rng = np.random.default_rng(9)
    a = np.concatenate([_rand(rng, 15000), _tandem("ACGT", 700), _rand(rng, 15000)])
    _check([a, a.copy()], po.get_seed(15, 3), min_restarts=1)


def test_block_order_decides():
    """1100 copies of one masked mer in one list, 80 in the other: the restart
    fires when the large block is processed first (the check before the second
    block sees more than 1000), which the idmer list's insertion order decides
    from the mers just before the run in each list."""
    fired = []
    for trial in range(24):
        rng = np.random.default_rng(20000 + trial)
        seed = SEEDS[trial % 4]
        a, b = _rand(rng, int(rng.integers(6000, 9000))), _rand(rng, int(rng.integers(6000, 9000)))
        big, small = (a, b) if trial % 2 else (b, a)
        at = int(rng.integers(0, len(big) - 1100))
        big[at:at + 1100] = ord("G")
        at = int(rng.integers(0, len(small) - 80))
        small[at:at + 80] = ord("C")
        r, _ = _check([a, b], seed, min_restarts=0)
        fired.append(r > 0)
    assert 4 <= sum(fired) <= 20
