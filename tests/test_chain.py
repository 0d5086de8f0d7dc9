"""The anchor-chain gap-job producer (host code, no GPU).

mh_chain_gap_jobs has no counterpart in the reference: it stands in for the
LCB / anchored profile-profile job producer so the path runs end to end.  Its
contract is checked against an O(n^2) chain DP on small random match sets:
the chain weight is the maximum over collinear, non-overlapping
forward-strand chains, consecutive anchors are ordered and disjoint in both
genomes, and the jobs are exactly the non-empty gaps of at most max_gap on
both sides.
"""
import numpy as np

import mauvehip


def _brute_best(s0, s1, ln):
    order = np.lexsort((np.arange(len(s0)), s1, s0))
    best = {}
    for i in order:
        b = ln[i]
        for j in order:
            if j in best and s0[j] + ln[j] <= s0[i] and s1[j] + ln[j] <= s1[i]:
                b = max(b, best[j] + ln[i])
        best[i] = b
    return max(best.values()) if best else 0


def _check(s0, s1, ln, La, Lb, max_gap):
    starts = np.stack([s0 + 1, s1 + 1], 1).astype(np.int64)
    chain, jobs = mauvehip.chain_gap_jobs(starts, ln.astype(np.uint64), La, Lb, max_gap)
    w = int(ln[chain].sum()) if len(chain) else 0
    assert w == _brute_best(s0, s1, ln)
    pa = pb = 0
    want = []
    for q in chain:
        assert s0[q] >= pa and s1[q] >= pb
        ga, gb = s0[q] - pa, s1[q] - pb
        if 0 < ga <= max_gap and 0 < gb <= max_gap:
            want.append((pa, pb, ga, gb))
        pa, pb = s0[q] + ln[q], s1[q] + ln[q]
    ga, gb = La - pa, Lb - pb
    if 0 < ga <= max_gap and 0 < gb <= max_gap:
        want.append((pa, pb, ga, gb))
    got = [(int(j["a_off"]), int(j["b_off"]), int(j["a_len"]), int(j["b_len"])) for j in jobs]
    assert got == want


def test_chain_matches_brute_force():
    rng = np.random.default_rng(5)
    for trial in range(60):
        n = int(rng.integers(0, 40))
        La, Lb = 2000, 2100
        ln = rng.integers(5, 60, size=n)
        s0 = np.array([rng.integers(0, La - l) for l in ln], dtype=np.int64)
        if trial % 2:   # mostly collinear, like real anchors
            s1 = np.clip(s0 + rng.integers(-30, 30, size=n), 0, None)
            s1 = np.minimum(s1, Lb - ln)
        else:
            s1 = np.array([rng.integers(0, Lb - l) for l in ln], dtype=np.int64)
        _check(s0, s1, ln, La, Lb, max_gap=int(rng.choice([50, 500, 5000])))


def test_chain_empty_and_reverse_strand_ignored():
    starts = np.array([[10, -50], [-5, 20]], dtype=np.int64)
    chain, jobs = mauvehip.chain_gap_jobs(starts, np.array([20, 20], dtype=np.uint64), 100, 100, 5000)
    assert len(chain) == 0
    assert [(int(j["a_off"]), int(j["b_off"]), int(j["a_len"]), int(j["b_len"])) for j in jobs] == [(0, 0, 100, 100)]


def _vector_best(s0, s1, ln):
    """The O(n^2) chain DP of _brute_best, one numpy pass per anchor."""
    order = np.lexsort((np.arange(len(s0)), s1, s0))
    best = np.zeros(len(s0), dtype=np.int64)
    e0, e1 = s0 + ln, s1 + ln
    for i in order:
        ok = (e0 <= s0[i]) & (e1 <= s1[i])
        best[i] = ln[i] + (best[ok].max() if ok.any() else 0)
    return int(best.max()) if len(best) else 0


def test_chain_large_genomes_and_repeated_starts():
    """Starts of more than 20 bits in both genomes (the sort keys then take
    several radix passes) and anchors sharing starts or ends."""
    rng = np.random.default_rng(11)
    for La, Lb, n in [(3_000_000, 3_100_000, 1500), (40_000_000, 70_000, 900), (900, 30_000_000, 700)]:
        ln = rng.integers(5, 400, size=n)
        s0 = (np.sort(rng.integers(0, La - 400, size=n)) + rng.integers(-2000, 2000, size=n)).clip(0, La - 400)
        s1 = (s0 * (Lb / La)).astype(np.int64) + rng.integers(-500, 500, size=n)
        s1 = s1.clip(0, Lb - 400)
        s0[rng.integers(0, n, size=n // 10)] = s0[0]      # shared starts in genome 1
        s1[rng.integers(0, n, size=n // 10)] = s1[1]      # and in genome 2
        starts = np.stack([s0 + 1, s1 + 1], 1).astype(np.int64)
        chain, _ = mauvehip.chain_gap_jobs(starts, ln.astype(np.uint64), La, Lb, 5000)
        assert int(ln[chain].sum()) == _vector_best(s0, s1, ln)
        pa = pb = 0
        for q in chain:
            assert s0[q] >= pa and s1[q] >= pb
            pa, pb = s0[q] + ln[q], s1[q] + ln[q]
