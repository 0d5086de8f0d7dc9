"""The oracle's SearchRange repeat-limit restart (LM/MatchFinder.cpp:166-340,
GetBreakpoint 89-126), pinned by hand-worked cases, and its FindMer pinned by
the reference's own SortedMerList::FindMer (tests/golden/findmer_golden.npz,
written by oracle/_ref/ref_harness from LM/SortedMerList.cpp).

Every case is small enough to reason about by hand: a poly-A stretch makes
the all-A mer (masked value 0, first in every sorted list) a run of
L - 16 entries (seed weight 17, solid), the rest is random sequence with no
other long A stretch, and the answer is written as "every masked mer with one
entry in each list, in mer order" minus what the restart drops, derived step
by step in each docstring.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))


def _seed():
    return po.get_seed(17, 3)          # weight 17 has no rank-3 mask: solid 17-mer


def _rand(rng, n):
    """n random bases ending in 'C' (so a following poly-A run keeps its length)."""
    return "".join("ACGT"[x] for x in rng.integers(0, 4, n - 1)) + "C"


def _lists(seqs, seed):
    w = po.seed_weight(seed)
    out = []
    for s in seqs:
        _, k, p = po.sml_build(s, seed)
        out.append((k, p, k >> np.uint64(64 - 2 * w)))
    return out


def _unique_pairs(seqs, seed, drop=()):
    """(pos0 + 1, pos1 + 1) of every masked mer with exactly one entry in each
    list, ascending masked mer; drop = (list, sorted index) entries removed."""
    L = _lists(seqs, seed)
    dropped = set(drop)
    per = []
    for g, (k, p, m) in enumerate(L):
        d = {}
        for i in range(len(k)):
            if (g, i) in dropped:
                continue
            d.setdefault(int(m[i]), []).append(int(p[i]))
        per.append(d)
    out = []
    for mer in sorted(set(per[0]) & set(per[1])):
        if len(per[0][mer]) == 1 and len(per[1][mer]) == 1:
            out.append((per[0][mer][0] + 1, per[1][mer][0] + 1))
    return out


def _oracle(seqs, seed):
    st, _ = po.seed_tuples(seqs, seed)
    _, _, stats = po.find_matches(seqs, seed)
    return [tuple(abs(int(x)) for x in r) for r in st], stats.restarts


def _count(L, g, mer):
    return int((L[g][2] == mer).sum())


def test_findmer_against_reference_golden():
    """FindMer's bsearch probe order: the found flag and the index it stops at
    (on a miss too) equal the reference's on every query, including lists
    shorter than the seed (result left unset)."""
    z = np.load(os.path.join(HERE, "golden", "findmer_golden.npz"), allow_pickle=False)
    total = 0
    for c in range(int(z["n_cases"][0])):
        seq = z[f"c{c}_seq"].tobytes()
        seed = int(z[f"c{c}_seed"][0])
        _, keys, _ = po.sml_build(seq, seed)
        for q, f, r in zip(z[f"c{c}_query"], z[f"c{c}_found"], z[f"c{c}_result"]):
            got_f, got_r = po.find_mer(keys, len(seq), seed, int(q))
            if got_f < 0:
                assert f == 0 and r == 0xFFFFFFFFFFFFFFFF
            else:
                assert (got_f, got_r) == (int(f), int(r)), (c, hex(int(q)))
            total += 1
    assert total > 5000


def test_restart_found_break_mer():
    """Both genomes: random 4000 + A*1200 + 'C' + shared random 4000.
    The all-A mer is 1184 entries in each list; both lists start with it, so
    both heads come from the initial (stable) sort, list 0 first.  After list
    0's block the run holds 1184 > 1000 entries when list 1's head (same mer)
    comes up: restart.  next_mer = masked 1 = A^16 C, present once in each
    list (the window ending at the 'C'); FindMer finds it in list 0 at index
    1184, so GetBreakpoint starts list 0 at the first entry of masked 1 (1184)
    and list 1 one past FindMer's entry (cur_matchI + 1).  The dropped run
    had no pair anyway; the A^16 C mer loses its list-1 entry, so its pair is
    the only one missing."""
    rng = np.random.default_rng(11)
    tail = "C" + _rand(rng, 4000)
    s0 = _rand(rng, 4000) + "A" * 1200 + tail
    s1 = _rand(rng, 4000) + "A" * 1200 + tail
    seed = _seed()
    L = _lists([s0, s1], seed)
    assert _count(L, 0, 0) == _count(L, 1, 0) == 1184
    assert _count(L, 0, 1) == _count(L, 1, 1) == 1
    j0 = int(np.nonzero(L[1][2] == 1)[0][0])
    got, restarts = _oracle([s0, s1], seed)
    assert restarts == 1
    want = _unique_pairs([s0, s1], seed, drop=[(1, j0)])
    assert got == want
    assert len(want) == len(_unique_pairs([s0, s1], seed)) - 1 > 3000


def test_single_list_run_is_enumerated():
    """A*1200 in genome 0 only: the all-A run is one block (1184 entries, one
    buffer), no iteration checks it while it is the current run, so no
    restart; it gives no pair (genome 1 has no entry)."""
    rng = np.random.default_rng(12)
    tail = _rand(rng, 4000)
    s0 = _rand(rng, 3000) + "A" * 1200 + tail
    s1 = _rand(rng, 3000) + "C" + tail
    got, restarts = _oracle([s0, s1], _seed())
    assert restarts == 0
    assert got == _unique_pairs([s0, s1], _seed())


@pytest.mark.parametrize("n0,n1,fires", [(700, 700, False), (700, 1200, False), (1200, 700, True)])
def test_check_counts_blocks_before_the_last(n0, n1, fires):
    """All-A runs of n0 - 16 and n1 - 16 entries, both heads initial, list 0's
    block first.  The only check while the run is current comes before list
    1's block and counts list 0's block: more than 1000 only when n0 = 1200
    (1184).  684 + 684 and 684 + 1184 both exceed 1000 in total and do not
    restart.  When it fires, list 1 skips its A^16 C entry as above."""
    rng = np.random.default_rng(13)
    tail = "C" + _rand(rng, 4000)
    s0 = _rand(rng, 3000) + "A" * n0 + tail
    s1 = _rand(rng, 3500) + "A" * n1 + tail
    seed = _seed()
    L = _lists([s0, s1], seed)
    assert _count(L, 0, 0) == n0 - 16 and _count(L, 1, 0) == n1 - 16
    got, restarts = _oracle([s0, s1], seed)
    assert restarts == (1 if fires else 0)
    drop = [(1, int(np.nonzero(L[1][2] == 1)[0][0]))] if fires else []
    assert got == _unique_pairs([s0, s1], seed, drop=drop)


def test_restart_break_mer_not_found():
    """As the first case, with 'G' after the poly-A and no A^16 C anywhere:
    FindMer(A^16 C) misses in list 0 and stops at index r (r is pinned by the
    reference golden above); next_pos = r + 1.  GetBreakpoint takes the
    entry at r + 1 as break_mer, starts list 0 at the first entry of its
    masked mer and list 1 one past FindMer's hit of break_mer's key.  Every
    list-0 entry between the dropped run and that start is skipped, and the
    list-1 entry of break_mer."""
    rng = np.random.default_rng(14)
    tail = "G" + _rand(rng, 4000)
    s0 = _rand(rng, 4000) + "A" * 1200 + tail
    s1 = _rand(rng, 4000) + "A" * 1200 + tail
    seed = _seed()
    w = po.seed_weight(seed)
    L = _lists([s0, s1], seed)
    assert _count(L, 0, 1) == 0 and _count(L, 1, 1) == 0
    q = 1 << (64 - 2 * w)
    f, r = po.find_mer(L[0][0], len(s0), seed, q)
    assert f == 0 and r in (1183, 1184)
    nxt = r + 1
    brk = L[0][0][nxt]
    start0 = int(np.nonzero(L[0][2] == L[0][2][nxt])[0][0])
    f1, j = po.find_mer(L[1][0], len(s1), seed, int(brk))
    assert f1 == 1
    drop = [(0, i) for i in range(1184, start0)] + [(1, j)]
    got, restarts = _oracle([s0, s1], seed)
    assert restarts == 1
    assert got == _unique_pairs([s0, s1], seed, drop=drop)
