"""GPU parity of mh_find_matches_unique (UniqueMatchFinder, more than four
genomes in progressiveMauve) with the oracle: identical match lists (order
included), seed and collision counts, for families of 2 to 16 related
genomes with inversions, duplicated stretches (genomes left out of a run's
seed) and poly-A runs above the repeat limit (restarts), under solid seeds of
odd length (maximal exact matches), a solid seed of even length and spaced
seeds (the k-genome ExtendMatch walk)."""
import numpy as np
import pytest

import mauvehip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

COMP = str.maketrans("ACGT", "TGCA")


def _family(rng, n, ng, snp=0.02, polya=0):
    base = "".join("ACGT"[x] for x in rng.integers(0, 4, n))
    if polya:
        at = n // 3
        base = base[:at] + "A" * polya + "C" + base[at:]
    seqs = [base]
    for g in range(ng - 1):
        t = list(base)
        for i in np.nonzero(rng.random(len(t)) < snp)[0]:
            t[i] = "ACGT"[("ACGT".index(t[i]) + int(rng.integers(1, 4))) % 4]
        t = "".join(t)
        at = int(rng.integers(0, n // 2))
        t = t[:at] + t[at:at + n // 8][::-1].translate(COMP) + t[at + n // 8:]
        if g % 3 == 0:
            d0 = int(rng.integers(0, n - 800))
            t = t + t[d0:d0 + 700]
        seqs.append(t)
    return seqs


def _gpu(ctx, seqs, seed):
    smls = [mauvehip.SortedMerList(ctx, s, seed) for s in seqs]
    try:
        return mauvehip.find_matches_unique(ctx, smls)
    finally:
        for x in smls:
            x.free()


def _check(ctx, seqs, seed):
    gs, gl, gst = _gpu(ctx, seqs, seed)
    os_, ol, ost = po.find_matches_unique(seqs, seed)
    assert gs.shape == os_.shape
    assert np.array_equal(gs, os_)
    assert np.array_equal(gl, ol)
    assert gst.seed_tuples == ost.hash_calls
    assert gst.collisions == ost.collisions
    assert gst.matches == len(ol)
    assert gst.repeat_runs == ost.restarts
    return gst


@pytest.mark.parametrize("ng,k", [(2, 13), (3, 11), (5, 13), (8, 17), (8, 13)])
def test_unique_solid_odd(ctx, ng, k):
    rng = np.random.default_rng(40 + ng + k)
    seqs = _family(rng, 30000, ng)
    st = _check(ctx, seqs, (1 << k) - 1)
    assert st.matches > 100 and st.collisions > 0


@pytest.mark.parametrize("ng,seed", [(5, 0xFFF), (6, 0x7AC9AF), (8, 0x7B974EF), (3, 0x16DF6D)])
def test_unique_even_and_spaced_seeds(ctx, ng, seed):
    rng = np.random.default_rng(60 + ng)
    seqs = _family(rng, 20000, ng)
    st = _check(ctx, seqs, seed)
    assert st.matches > 50


def test_unique_two_genomes_equals_pairwise(ctx):
    rng = np.random.default_rng(7)
    seqs = _family(rng, 40000, 2)
    seed = (1 << 15) - 1
    gs, gl, _ = _gpu(ctx, seqs, seed)
    smls = [mauvehip.SortedMerList(ctx, s, seed) for s in seqs]
    try:
        pm, _ = mauvehip.find_matches_raw(ctx, smls)
    finally:
        for x in smls:
            x.free()
    assert np.array_equal(gs, pm["start"][:, :2])
    assert np.array_equal(gl, pm["length"])


@pytest.mark.parametrize("ng", [5, 7])
def test_unique_repeat_restarts(ctx, ng):
    rng = np.random.default_rng(80 + ng)
    seqs = _family(rng, 20000, ng, polya=700)
    st = _check(ctx, seqs, (1 << 13) - 1)
    assert st.repeat_runs >= 1


def test_unique_sixteen_genomes(ctx):
    rng = np.random.default_rng(99)
    seqs = _family(rng, 12000, 16, snp=0.01)
    st = _check(ctx, seqs, (1 << 15) - 1)
    assert st.matches > 100


def test_unique_eight_genome_tree_family(ctx):
    """The multi-genome workload's generator (eight genomes along a random
    binary tree: SNPs, indels, inversions, deletions) with a 300 kb ancestor
    and the default seed of weight 19 (rank 0, spaced)."""
    from mauvehip import synth
    fam = [q.tobytes() for q in synth.genome_family(length=300_000)]
    st = _check(ctx, fam, mauvehip.get_seed(19, 0))
    assert st.matches > 10000


@pytest.mark.parametrize("seed", [0xFFF, 0x7B974EF])
def test_unique_chain_walks_equal_every_seed_walk(ctx, monkeypatch, seed):
    """Spaced (and even solid) seeds walk once per chain of seeds one seed
    length apart on their diagonal: the lists equal those of a walk from every
    seed, with fewer walks."""
    rng = np.random.default_rng(66)
    seqs = _family(rng, 20000, 6)
    monkeypatch.setenv("MH_UQ_CHAINS", "0")
    s0, l0, st0 = _gpu(ctx, seqs, seed)
    monkeypatch.delenv("MH_UQ_CHAINS")
    s1, l1, st1 = _gpu(ctx, seqs, seed)
    assert np.array_equal(s0, s1) and np.array_equal(l0, l1)
    assert st0.collisions == st1.collisions
    assert st0.walks == st0.seed_tuples
    assert 0 < st1.walks < st0.walks
