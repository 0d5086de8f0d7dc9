"""The oracle's UniqueMatchFinder (MA/UniqueMatchFinder.cpp:36-60 over
LM/MatchFinder.cpp:172-340 and LM/MemHash.cpp:167-251), pinned by brute
force with solid seeds of odd length (no palindromic windows):

* seed tuples: for every canonical k-mer, in key order, the genomes that hold
  it exactly once, hashed together when there are at least two; the first of
  them leads, and a genome whose window is the other strand's is negative
  (SetDirection);
* extension: ExtendMatch over the present genomes with a solid seed is the
  longest stretch around the seed where every present genome agrees with the
  first, i.e. the smallest of the pairwise maximal exact matches on each side;
* the list: each distinct (genomes present, extension) once, in bucket order
  of the k-genome offset (CalculateOffset), collisions = seeds - entries.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

COMP = str.maketrans("ACGT", "TGCA")


def _canon(w):
    rc = w[::-1].translate(COMP)
    return (w, 0) if w <= rc else (rc, 1)


def brute_tuples(seqs, k):
    idx = []
    for s in seqs:
        d = {}
        for p in range(len(s) - k + 1):
            key, flag = _canon(s[p:p + k])
            d.setdefault(key, []).append((p, flag))
        idx.append(d)
    out = []
    for key in sorted(set().union(*[set(d) for d in idx])):
        uniq = [(g, d[key][0]) for g, d in enumerate(idx) if len(d.get(key, ())) == 1]
        if len(uniq) < 2:
            continue
        f0 = uniq[0][1][1]
        st = [0] * len(seqs)
        for g, (p, f) in uniq:
            st[g] = p + 1 if f == f0 else -(p + 1)
        out.append(tuple(st))
    return out


def _rc_pos(n, s, k):
    """0-based window start in the reverse complement for a negative start."""
    return n - (-s - 1) - k


def brute_extend(seqs, st, k):
    gs = [g for g, v in enumerate(st) if v]
    f = gs[0]
    a = seqs[f]
    p0 = st[f] - 1
    left = right = None
    views = {}
    for g in gs[1:]:
        b = seqs[g] if st[g] > 0 else seqs[g][::-1].translate(COMP)
        p1 = st[g] - 1 if st[g] > 0 else _rc_pos(len(seqs[g]), st[g], k)
        views[g] = (b, p1)
        lf = 0
        while p0 - lf - 1 >= 0 and p1 - lf - 1 >= 0 and a[p0 - lf - 1] == b[p1 - lf - 1]:
            lf += 1
        rt = 0
        while p0 + k + rt < len(a) and p1 + k + rt < len(b) and a[p0 + k + rt] == b[p1 + k + rt]:
            rt += 1
        left = lf if left is None else min(left, lf)
        right = rt if right is None else min(right, rt)
    L = k + left + right
    out = [0] * len(st)
    out[f] = p0 - left + 1
    for g in gs[1:]:
        b, p1 = views[g]
        out[g] = p1 - left + 1 if st[g] > 0 else -(len(seqs[g]) - (p1 - left + L) + 1)
    return tuple(out), L


def _family(rng, n, ng, snp=0.03, dup=True):
    base = "".join("ACGT"[x] for x in rng.integers(0, 4, n))
    seqs = [base]
    for g in range(ng - 1):
        t = list(base)
        for i in np.nonzero(rng.random(len(t)) < snp)[0]:
            t[i] = "ACGT"[("ACGT".index(t[i]) + int(rng.integers(1, 4))) % 4]
        t = "".join(t)
        at = int(rng.integers(0, n // 2))
        t = t[:at] + t[at:at + n // 6][::-1].translate(COMP) + t[at + n // 6:]
        if dup and g % 2 == 0:
            # a duplicated stretch: its mers occur twice in this genome
            d0 = int(rng.integers(0, n - 300))
            t = t + t[d0:d0 + 250]
        seqs.append(t)
    return seqs


@pytest.mark.parametrize("k,ng", [(11, 3), (13, 5), (11, 6), (13, 8)])
def test_seed_tuples_brute_force(k, ng):
    rng = np.random.default_rng(500 + k + ng)
    seqs = _family(rng, 2500, ng)
    seed = (1 << k) - 1
    st, off = po.seed_tuples_unique(seqs, seed)
    want = brute_tuples(seqs, k)
    assert [tuple(int(v) for v in x) for x in st] == want
    # the duplicated stretches leave genomes out of some seeds
    assert any(sum(1 for v in t if v) < ng for t in want)
    # CalculateOffset with the seed length
    for t, o in zip(want, off):
        f = next(i for i, v in enumerate(t) if v)
        assert o == sum(v - t[f] - (k if v < 0 else 0) for v in t[f + 1:] if v)


@pytest.mark.parametrize("k,ng", [(11, 3), (13, 5), (11, 6), (13, 8)])
def test_match_list_brute_force(k, ng):
    rng = np.random.default_rng(700 + k + ng)
    seqs = _family(rng, 3000, ng)
    seed = (1 << k) - 1
    starts, lengths, stats = po.find_matches_unique(seqs, seed)
    tup = brute_tuples(seqs, k)
    want = {brute_extend(seqs, t, k) for t in tup}
    got = [(tuple(int(v) for v in s), int(L)) for s, L in zip(starts, lengths)]
    assert len(got) == len(set(got))
    assert set(got) == want
    assert stats.hash_calls == len(tup)
    assert stats.collisions == len(tup) - len(want)

    def bucket(m):
        st, L = m
        f = next(i for i, v in enumerate(st) if v)
        off = sum(v - st[f] - (L if v < 0 else 0) for v in st[f + 1:] if v)
        return ((off % 40000) + 40000) % 40000
    bs = [bucket(m) for m in got]
    assert bs == sorted(bs)


def test_genomes_with_two_copies_left_out():
    """Genome 1 holds the shared stretch twice: its mers' seeds hold genomes 0
    and 2 only, and the stretch's match spans genomes 0 and 2."""
    rng = np.random.default_rng(9)
    r = lambda n: "".join("ACGT"[x] for x in rng.integers(0, 4, n))
    m = r(200)
    g0 = r(300) + m + r(300)
    g1 = r(250) + m + r(100) + m + r(250)
    g2 = r(400) + m + r(200)
    seed = (1 << 13) - 1
    starts, lengths, _ = po.find_matches_unique([g0, g1, g2], seed)
    hit = [(tuple(int(v) for v in s), int(L)) for s, L in zip(starts, lengths) if int(L) >= 200]
    assert hit == [((301, 0, 401), 200)]
