"""GPU parity of the sorted-mer-list build (pack, keys, radix sort) against the oracle.

Bar: bit-exact packed words, keys and sorted order (ties in ascending position
on both sides).  The oracle is pinned to the reference SortedMerList build by
tests/test_oracle_keys.py.
"""
import numpy as np
import pytest

import mauvehip
from mauvehip import synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SEEDS = [po.get_seed(15, 3), po.get_seed(17, 3), po.get_seed(19, 3), po.get_seed(11, 0),
         po.get_seed(9, 3), po.get_seed(21, 1), po.get_seed(31, 0)]


def _check_sml(ctx, seq, seed):
    sml = mauvehip.SortedMerList(ctx, seq, seed)
    words, keys, pos = po.sml_build(seq, seed)
    hd = sml.header()
    assert hd.sml_length == len(keys)
    assert hd.seed_length == po.seed_length(seed) and hd.seed_weight == po.seed_weight(seed)
    gw = sml.sequence_words()
    assert np.array_equal(gw, words)
    gp, gk = sml.read()
    assert np.array_equal(gk, keys)
    assert np.array_equal(gp, pos)
    sml.free()


@pytest.mark.parametrize("seed", SEEDS)
def test_sml_random(ctx, seed):
    rng = np.random.default_rng(seed & 0xFFFF)
    for L in (po.seed_length(seed), 40, 100, 1000, 4097, 20000):
        seq = synth.BASES[rng.integers(0, 4, size=L)].tobytes()
        _check_sml(ctx, seq, seed)


def test_sml_iupac_and_case(ctx):
    rng = np.random.default_rng(3)
    alph = np.frombuffer(b"ACGTacgtNRYKMSWBDHVn", dtype=np.uint8)
    seq = alph[rng.integers(0, len(alph), size=5000)].tobytes()
    for seed in SEEDS[:3]:
        _check_sml(ctx, seq, seed)


def test_sml_low_complexity_ties(ctx):
    seq = (b"A" * 3000) + (b"ACGT" * 500) + (b"AT" * 700)
    for seed in SEEDS[:3]:
        _check_sml(ctx, seq, seed)


def test_sml_gap_is_an_error(ctx):
    with pytest.raises(mauvehip.MhError) as e:
        mauvehip.SortedMerList(ctx, b"ACGTACGTAC-GTACGTACGTACGTACGTA", SEEDS[0])
    assert e.value.code == -4


def test_sml_seed_mers_and_find_mer(ctx):
    rng = np.random.default_rng(4)
    seq = synth.BASES[rng.integers(0, 4, size=30000)].tobytes()
    seed = SEEDS[0]
    sml = mauvehip.SortedMerList(ctx, seq, seed)
    kp = po.keys_by_position(seq, seed)
    q = rng.integers(0, len(kp), size=500).astype(np.uint32)
    assert np.array_equal(sml.get_seed_mers(q), kp[q])
    _, keys, _ = po.sml_build(seq, seed)
    last = len(seq) - po.seed_length(seed)

    def bsearch(qm):   # LM/SortedMerList.cpp:380-394
        s, e = 0, last
        while True:
            m = (s + e) // 2
            if keys[m] == qm:
                return m
            if keys[m] < qm and m < e:
                s = m + 1
                continue
            if keys[m] > qm and s < m:
                e = m - 1
                continue
            return m

    for qm in list(keys[rng.integers(0, len(keys), 50)]) + [int(x) + 2 for x in keys[rng.integers(0, len(keys), 50)]]:
        f, idx = sml.find_mer(int(qm))
        assert idx == bsearch(int(qm))
        assert f == (int(keys[idx]) == int(qm))


@pytest.mark.parametrize("which", ["iid_200kb", "recoded_4.6Mb"])
def test_sml_full_size(ctx, which):
    if which == "iid_200kb":
        ref, mut = synth.pair_iid()
        seed = po.get_seed(15, 3)
    else:
        ref, mut, _ = synth.pair_recoded()
        seed = po.get_seed(17, 3)
    for g in (ref, mut):
        _check_sml(ctx, g.tobytes(), seed)
