"""GPU parity of pairwise match finding (join + extension + hash replay) against the oracle.

Bar: identical match lists (starts, lengths, order) and identical seed/collision
counts.  The oracle's SearchRange/HashMatch/ExtendMatch/AddHashEntry
restatement is pinned by brute force in tests/test_oracle_match.py.
"""
import numpy as np
import pytest

import mauvehip
from mauvehip import synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _compare(ctx, a, b, seed):
    sa = mauvehip.SortedMerList(ctx, a, seed)
    sb = mauvehip.SortedMerList(ctx, b, seed)
    gs, gl, gst = mauvehip.find_matches(ctx, [sa, sb])
    os_, ol, ost = po.find_matches([a, b], seed)
    assert gst.seed_tuples == ost.hash_calls
    assert gst.collisions == ost.collisions
    assert len(gl) == len(ol)
    assert np.array_equal(gs, os_)
    assert np.array_equal(gl, ol)
    _compare.last_stats = gst
    return len(gl)


def _mutant(rng, ref, snp=0.01, indel=0.001, inv=5000):
    m = ref.copy()
    hit = rng.random(len(m)) < snp
    m[hit] = synth.BASES[(np.searchsorted(synth.BASES, m[hit]) + rng.integers(1, 4, hit.sum())) % 4]
    out, prev = [], 0
    for p in np.nonzero(rng.random(len(m)) < indel)[0]:
        if p < prev:
            continue
        out.append(m[prev:p])
        if rng.random() < 0.5:
            out.append(synth.BASES[rng.integers(0, 4, rng.integers(1, 21))])
            prev = p
        else:
            prev = p + int(rng.integers(1, 21))
    out.append(m[prev:])
    m = np.concatenate(out)
    if inv and len(m) > 2 * inv:
        s = int(rng.integers(0, len(m) - inv))
        m[s:s + inv] = synth.revcomp(m[s:s + inv])
    return m


@pytest.mark.parametrize("seed", [po.get_seed(15, 3), po.get_seed(17, 3), po.get_seed(11, 0), po.get_seed(13, 3)])
def test_match_small_pairs(ctx, seed):
    rng = np.random.default_rng(seed & 0xFFF)
    for L in (300, 3000, 30000):
        ref = synth.BASES[rng.integers(0, 4, size=L)]
        mut = _mutant(rng, ref, inv=L // 6)
        _compare(ctx, ref.tobytes(), mut.tobytes(), seed)


def test_match_iid_200kb(ctx):
    ref, mut = synth.pair_iid()
    n = _compare(ctx, ref.tobytes(), mut.tobytes(), po.get_seed(15, 3))
    assert n > 0


def test_match_recoded_4_6Mb(ctx):
    ref, mut, _ = synth.pair_recoded()
    n = _compare(ctx, ref.tobytes(), mut.tobytes(), po.get_seed(17, 3))
    assert n > 10000
    st = _compare.last_stats          # a solid seed: every seed walks on its own
    assert st.walks == st.seed_tuples and st.long_walks == 0


@pytest.mark.parametrize("seed", [po.get_seed(15, 0), po.get_seed(11, 0)])
def test_match_spaced_long_walks(ctx, seed):
    """Spaced seeds over long identical and near-identical stretches, forward
    and inverted: the seeds of one chain (one seed length apart on a diagonal)
    share one walk, and walks past the step budget finish per diagonal."""
    rng = np.random.default_rng(seed & 0xFFFF)
    a = synth.BASES[rng.integers(0, 4, 120_000)]
    b = a.copy()
    snp = rng.random(len(b)) < 0.02
    snp[:30_000] = False                                  # identical
    snp[60_000:90_000] = False                            # identical, inverted below
    b[snp] = synth.BASES[(np.searchsorted(synth.BASES, b[snp]) + rng.integers(1, 4, snp.sum())) % 4]
    b[60_000:120_000] = synth.revcomp(b[60_000:120_000])
    b = np.concatenate([b[:45_000], b[45_012:100_000], synth.BASES[rng.integers(0, 4, 7)], b[100_000:]])
    n = _compare(ctx, a.tobytes(), b.tobytes(), seed)
    assert n >= 5
    st = _compare.last_stats
    assert 0 < st.long_walks <= st.walks < st.seed_tuples


@pytest.mark.parametrize("seed", [po.get_seed(17, 3), po.get_seed(15, 3)])
def test_match_bucket_alias(ctx, seed):
    """Diagonals 40000 apart share a MemHash bucket: the second half of the
    derivative sits 40000 bp further along, so both halves hash to bucket 0,
    next to an inverted block (reverse-strand seeds) and SNP-split runs."""
    rng = np.random.default_rng(77)
    ref = synth.BASES[rng.integers(0, 4, size=120000)]
    ref[30000:30050] = ref[10000:10050]           # a repeat inside genome 0
    mid = synth.BASES[rng.integers(0, 4, size=40000)]
    mut = np.concatenate([ref[:60000], mid, ref[60000:]])
    mut[20000:26000] = synth.revcomp(mut[20000:26000])
    hit = rng.random(len(mut)) < 0.01
    mut[hit] = synth.BASES[(np.searchsorted(synth.BASES, mut[hit]) + 1) % 4]
    n = _compare(ctx, ref.tobytes(), mut.tobytes(), seed)
    assert n > 20


def _relatives(rng, ref, count, snp, inv):
    out = [ref]
    for _ in range(count - 1):
        m = ref.copy()
        hit = rng.random(len(m)) < snp
        m[hit] = synth.BASES[(np.searchsorted(synth.BASES, m[hit]) + rng.integers(1, 4, hit.sum())) % 4]
        at = int(rng.integers(0, len(m) - inv))
        m[at:at + inv] = synth.revcomp(m[at:at + inv])
        cut = int(rng.integers(0, len(m) - 50))
        m = np.concatenate([m[:cut], m[cut + int(rng.integers(1, 30)):]])
        out.append(m)
    return out


@pytest.mark.parametrize("ng,seed", [(3, po.get_seed(15, 3)), (3, po.get_seed(13, 0)), (4, po.get_seed(17, 3)),
                                     (4, po.get_seed(15, 0))])
def test_match_three_and_four_genomes(ctx, ng, seed):
    """PairwiseMatchFinder over three and four genomes (solid and spaced seeds,
    relatives with SNPs, an inversion and a deletion each): identical match
    lists (all k starts, lengths, order) and seed/collision counts."""
    rng = np.random.default_rng(ng * 1000 + (seed & 0xFF))
    seqs = [s.tobytes() for s in _relatives(rng, synth.BASES[rng.integers(0, 4, 60_000)], ng, 0.03, 6000)]
    smls = [mauvehip.SortedMerList(ctx, s, seed) for s in seqs]
    gs, gl, gst = mauvehip.find_matches(ctx, smls)
    os_, ol, ost = po.find_matches(seqs, seed)
    assert gst.seed_tuples == ost.hash_calls
    assert gst.collisions == ost.collisions
    assert len(gl) == len(ol) > 50
    assert np.array_equal(gs, os_)
    assert np.array_equal(gl, ol)
    assert (gs != 0).sum(axis=1).max() == 2      # pairwise matches: two genomes each


def _tandem(unit, copies):
    return np.frombuffer((unit * copies).encode(), dtype=np.uint8).copy()


@pytest.mark.parametrize("seed", [po.get_seed(15, 3), po.get_seed(17, 3), po.get_seed(11, 0)])
def test_match_repeat_restart(ctx, seed):
    """Runs above MER_REPEAT_LIMIT (poly-A, (AC)n, (ACG)n, (AGCT)n) in both
    genomes: the search restarts as the reference's SearchRange does
    (LM/MatchFinder.cpp:253-276), and the match lists and restart counts
    equal the oracle's."""
    rng = np.random.default_rng(seed & 0xFFF)
    r = lambda n: synth.BASES[rng.integers(0, 4, n)]
    a = np.concatenate([r(6000), np.full(1300, ord("A"), np.uint8), r(5000), _tandem("AC", 1500), r(3000),
                        _tandem("ACG", 1200), r(3000), _tandem("AGCT", 1100), r(8000)])
    b = _mutant(rng, a, snp=0.01, indel=0.0005, inv=3000)
    b = np.concatenate([r(2000), b[:9000], _tandem("AC", 900), b[9000:], np.full(800, ord("T"), np.uint8), r(500)])
    print(f"[repeat] seed {seed:#x} lengths {len(a)} {len(b)}", flush=True)
    _compare(ctx, a.tobytes(), b.tobytes(), seed)
    _, _, ost = po.find_matches([a.tobytes(), b.tobytes()], seed)
    assert _compare.last_stats.repeat_runs == ost.restarts >= 1


@pytest.mark.parametrize("trial", [32, 60, 2, 16])
def test_match_repeat_restart_at_buffer_ends(ctx, trial):
    """Poly-C runs whose sorted entries straddle a 10,000-entry buffer end: the
    restart fires (32, 60) or not (2, 16) on the refill check."""
    rng = np.random.default_rng(10000 + trial)
    seed = [po.get_seed(15, 3), po.get_seed(17, 3)][trial % 2]
    a = synth.BASES[rng.integers(0, 4, int(rng.integers(15000, 45000)))]
    b = synth.BASES[rng.integers(0, 4, int(rng.integers(15000, 45000)))]
    for s in (a, b):
        n = int(rng.integers(450, 700))
        at = int(rng.integers(0, len(s) - n))
        s[at:at + n] = ord("C")
    _compare(ctx, a.tobytes(), b.tobytes(), seed)
    assert (_compare.last_stats.repeat_runs > 0) == (trial in (32, 60))


@pytest.mark.parametrize("ng,seed", [(3, po.get_seed(15, 3)), (4, po.get_seed(17, 3)), (4, po.get_seed(11, 0))])
def test_match_repeat_restart_multi(ctx, ng, seed):
    """Repeat-limit restarts with three and four genomes (PairwiseMatchFinder):
    identical match lists and counts."""
    rng = np.random.default_rng(ng * 77 + (seed & 0xFF))
    seqs = _relatives(rng, synth.BASES[rng.integers(0, 4, 40_000)], ng, 0.02, 4000)
    seqs = [np.concatenate([s[:10000], np.full(600 + 400 * g, ord("A"), np.uint8), s[10000:20000],
                            _tandem("AC", 300 + 250 * (ng - g)), s[20000:]]) for g, s in enumerate(seqs)]
    seqs = [s.tobytes() for s in seqs]
    smls = [mauvehip.SortedMerList(ctx, s, seed) for s in seqs]
    gs, gl, gst = mauvehip.find_matches(ctx, smls)
    os_, ol, ost = po.find_matches(seqs, seed)
    assert gst.seed_tuples == ost.hash_calls
    assert gst.collisions == ost.collisions
    assert gst.repeat_runs == ost.restarts
    assert len(gl) == len(ol) > 50
    assert np.array_equal(gs, os_)
    assert np.array_equal(gl, ol)
