"""Oracle pairwise match finding pinned by brute force.

MatchFinder/MemHash need boost, which is absent, so there is no reference
build to compare with.  For solid seeds the reference semantics reduce to
closed forms that a plausible slip would break:
  * seed tuples = k-mers whose canonical key occurs exactly once in each
    genome, in ascending key order, genome-1 start negated when the strands differ;
  * ExtendMatch of such a seed = the maximal exact match containing it
    (forward, or against the reverse complement);
  * the match list = each distinct extended match once, in bucket
    (offset mod 40000) order, by start within a bucket.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

COMP = str.maketrans("ACGT", "TGCA")


def canon(w):
    rc = w[::-1].translate(COMP)
    return (w, 0) if w < rc or w == rc else (rc, 1)


def brute_tuples(a, b, k):
    def index(s):
        d = {}
        for p in range(len(s) - k + 1):
            key, flag = canon(s[p:p + k])
            d.setdefault(key, []).append((p, flag))
        return d
    ia, ib = index(a), index(b)
    out = []
    for key in sorted(set(ia) & set(ib)):
        if len(ia[key]) == 1 and len(ib[key]) == 1:
            (p0, f0), (p1, f1) = ia[key][0], ib[key][0]
            s1 = p1 + 1 if f0 == f1 else -(p1 + 1)
            out.append((p0 + 1, s1))
    return out


def brute_extend(a, b, s0, s1, k):
    bb = b if s1 > 0 else b[::-1].translate(COMP)
    p0 = s0 - 1
    p1 = s1 - 1 if s1 > 0 else len(b) - (-s1 - 1) - k
    left = 0
    while p0 - left - 1 >= 0 and p1 - left - 1 >= 0 and a[p0 - left - 1] == bb[p1 - left - 1]:
        left += 1
    right = 0
    while p0 + k + right < len(a) and p1 + k + right < len(bb) and a[p0 + k + right] == bb[p1 + k + right]:
        right += 1
    n0 = p0 - left + 1
    L = k + left + right
    if s1 > 0:
        n1 = p1 - left + 1
    else:
        n1 = -(len(b) - (p1 - left + L) + 1)
    return n0, n1, L


def _pair(rng, L, snp=0.03):
    a = "".join("ACGT"[x] for x in rng.integers(0, 4, L))
    b = list(a)
    for i in np.nonzero(rng.random(L) < snp)[0]:
        b[i] = "ACGT"[(("ACGT".index(b[i])) + int(rng.integers(1, 4))) % 4]
    b = "".join(b)
    s = int(rng.integers(0, L // 2))
    b = b[:s] + b[s:s + L // 5][::-1].translate(COMP) + b[s + L // 5:]
    return a, b


@pytest.mark.parametrize("k", [9, 11, 13])
def test_seed_tuples_brute_force(k):
    rng = np.random.default_rng(k)
    a, b = _pair(rng, 3000)
    seed = (1 << k) - 1
    st, off = po.seed_tuples([a, b], seed)
    assert [tuple(x) for x in st] == brute_tuples(a, b, k)
    for (s0, s1), o in zip(st, off):
        assert o == (s1 - s0 - (k if s1 < 0 else 0))


@pytest.mark.parametrize("k", [9, 11, 15])
def test_extension_is_maximal_exact_match(k):
    rng = np.random.default_rng(100 + k)
    a, b = _pair(rng, 4000)
    seed = (1 << k) - 1
    st, _ = po.seed_tuples([a, b], seed)
    for s0, s1 in st[:: max(1, len(st) // 300)]:
        got = po.extend_seed([a, b], seed, [s0, s1], k)
        n0, n1, L = brute_extend(a, b, int(s0), int(s1), k)
        assert (got[0][0], got[0][1], got[1]) == (n0, n1, L), (s0, s1)


@pytest.mark.parametrize("k", [11, 13])
def test_match_list_brute_force(k):
    rng = np.random.default_rng(200 + k)
    a, b = _pair(rng, 5000)
    seed = (1 << k) - 1
    starts, lengths, stats = po.find_matches([a, b], seed)
    tup = brute_tuples(a, b, k)
    ext = {brute_extend(a, b, s0, s1, k) for s0, s1 in tup}
    got = [(int(s[0]), int(s[1]), int(L)) for s, L in zip(starts, lengths)]
    assert len(got) == len(set(got)) == len(ext)
    assert set(got) == ext
    assert stats.hash_calls == len(tup)
    assert stats.collisions == len(tup) - len(ext)

    def bucket(m):
        s0, s1, L = m
        off = s1 - s0 - (L if s1 < 0 else 0)   # CalculateOffset on the extended match
        return ((off % 40000) + 40000) % 40000
    assert [bucket(m) for m in got] == sorted(bucket(m) for m in got)


def test_spaced_seed_invariants():
    rng = np.random.default_rng(5)
    a, b = _pair(rng, 6000, snp=0.02)
    seed = po.get_seed(15, 3)
    sl = po.seed_length(seed)
    starts, lengths, stats = po.find_matches([a, b], seed)
    assert len(lengths) > 0 and np.all(lengths >= sl)
    assert stats.hash_calls == stats.collisions + stats.matches
    # every match covers at least one seed tuple on the same diagonal
    st, _ = po.seed_tuples([a, b], seed)
    for (s0, s1), L in zip(starts, lengths):
        if s1 > 0:
            assert any(s0 <= t0 and t0 + sl <= s0 + L and t1 - t0 == s1 - s0 for t0, t1 in st if t1 > 0)


@pytest.mark.parametrize("k", [11, 13, 17])
def test_clean_bucket_closed_form(k):
    """A bucket whose seeds all lie on one diagonal or one anti-diagonal
    (same offset, same strand) ends as its seeds' distinct extensions in
    genome-0 start order -- the closed form the GPU resolves such buckets
    with -- for both strands, checked against the literal replay."""
    rng = np.random.default_rng(300 + k)
    a, b = _pair(rng, 6000, snp=0.02)
    seed = (1 << k) - 1
    starts, lengths, _ = po.find_matches([a, b], seed)
    st, off = po.seed_tuples([a, b], seed)

    def bucket(o):
        return ((int(o) % 40000) + 40000) % 40000
    lines = {}
    for (s0, s1), o in zip(st, off):
        lines.setdefault(bucket(o), set()).add((int(o), s1 < 0))
    got = {}
    for s, L in zip(starts, lengths):
        s0, s1 = int(s[0]), int(s[1])
        o = s1 - s0 - (int(L) if s1 < 0 else 0)
        got.setdefault(bucket(o), []).append((s0, s1, int(L)))
    clean_rc = 0
    for bk, ls in lines.items():
        if len(ls) != 1:
            continue
        members = [(int(s0), int(s1)) for (s0, s1), o in zip(st, off) if bucket(o) == bk]
        ext = sorted({brute_extend(a, b, s0, s1, k) for s0, s1 in members})
        assert got[bk] == ext, bk
        clean_rc += next(iter(ls))[1]
    assert clean_rc > 0   # reverse-strand clean buckets were exercised


@pytest.mark.parametrize("k,ng", [(11, 3), (13, 3), (11, 4)])
def test_multi_genome_brute_force(k, ng):
    """Three or four genomes (PairwiseMatchFinder): every run hashes each pair
    of genomes holding the mer exactly once, so with a solid seed the seed
    tuples are the union over genome pairs of the two-genome closed form (the
    first genome of the pair leading, the others absent), the list is each
    distinct extension of each pair once (entries of different pairs never
    compare equal: MheCompare orders genome presence first), in bucket order,
    and the collisions are counted per pair."""
    rng = np.random.default_rng(300 + k + ng)
    base, _ = _pair(rng, 4000)
    seqs = [base]
    for _ in range(ng - 1):
        # a relative of base: its own SNPs and inversion, on base's sequence
        t = list(base)
        for i in np.nonzero(rng.random(len(t)) < 0.03)[0]:
            t[i] = "ACGT"[("ACGT".index(t[i]) + int(rng.integers(1, 4))) % 4]
        t = "".join(t)
        at = int(rng.integers(0, 2000))
        seqs.append(t[:at] + t[at:at + 700][::-1].translate(COMP) + t[at + 700:])
    seed = (1 << k) - 1
    starts, lengths, stats = po.find_matches(seqs, seed)
    want, tuples = set(), 0
    for x in range(ng):
        for y in range(x + 1, ng):
            tup = brute_tuples(seqs[x], seqs[y], k)
            tuples += len(tup)
            for s0, s1 in tup:
                n0, n1, L = brute_extend(seqs[x], seqs[y], s0, s1, k)
                st = [0] * ng
                st[x], st[y] = n0, n1
                want.add((tuple(st), L))
    got = [(tuple(int(v) for v in s), int(L)) for s, L in zip(starts, lengths)]
    assert len(got) == len(set(got))
    assert set(got) == want
    assert stats.hash_calls == tuples
    assert stats.collisions == tuples - len(want)

    def bucket(m):
        st, L = m
        first = next(i for i, v in enumerate(st) if v)
        off = sum(v - st[first] - (L if v < 0 else 0) for v in st[first + 1:] if v)
        return ((off % 40000) + 40000) % 40000
    assert [bucket(m) for m in got] == sorted(bucket(m) for m in got)
