"""Oracle keys / packing / sort against the reference's own SortedMerList
(golden vectors from oracle/_ref) and against first-principles brute force."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "keys_golden.npz")


def _cases():
    z = np.load(GOLD, allow_pickle=False)
    for ci in range(int(z["n_cases"][0])):
        yield (z[f"c{ci}_seq"].tobytes(), int(z[f"c{ci}_seed"][0]), z[f"c{ci}_words"], z[f"c{ci}_fill"],
               z[f"c{ci}_seedmer"], z[f"c{ci}_pos"])


@pytest.mark.parametrize("ci", range(7))
def test_oracle_matches_reference_sml(ci):
    seq, seed, words, fill, seedmer, pos = list(_cases())[ci]
    ow = po.packed(seq)
    assert np.array_equal(ow[:len(words)], words)
    kp = po.keys_by_position(seq, seed)
    assert np.array_equal(kp, seedmer)
    if po.seed_length(seed) == po.seed_weight(seed):
        assert np.array_equal(po.fill_dna_sml(seq, seed), fill)
    else:
        assert np.array_equal(kp, fill)
    _, okeys, opos = po.sml_build(seq, seed)
    # same key sequence; equal keys may come in another order (std::sort is not stable)
    assert np.array_equal(okeys, seedmer[pos])
    for k in np.unique(okeys):
        assert set(opos[okeys == k]) == set(pos[seedmer[pos] == k])


CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def brute_key(seq, p, seed):
    """Canonical seed key from the definition: the characters under the
    pattern's set bits (most significant bit = first character), two bits each,
    left-aligned; the reverse complement gets the LSB set; the smaller wins."""
    sl, sw = po.seed_length(seed), po.seed_weight(seed)
    low = (seed & -seed).bit_length() - 1
    pat = [(seed >> (low + sl - 1 - q)) & 1 for q in range(sl)]
    chars = [seq[p + q] for q in range(sl) if pat[q]]
    codes = [CODE.get(c.upper(), 0) if c.upper() in "ACGT" else {"B": 1, "Y": 1, "S": 2, "K": 2}.get(c.upper(), 0)
             for c in chars]
    f = 0
    for c in codes:
        f = (f << 2) | c
    r = 0
    for c in reversed(codes):
        r = (r << 2) | (3 - c)
    f64 = f << (64 - 2 * sw)
    r64 = (r << (64 - 2 * sw)) | 1
    return min(f64, r64)


@pytest.mark.parametrize("seed", [0x16df6d, 0x1FFFF, 0x7FFFF, 0x6dbedb, 0x1ab5, 0x7fffffff])
def test_keys_brute_force(seed):
    rng = np.random.default_rng(seed & 0xFFFF)
    seq = "".join("ACGTacgtNBYSK"[x] for x in rng.integers(0, 13, 300))
    kp = po.keys_by_position(seq, seed)
    for p in range(len(kp)):
        assert int(kp[p]) == brute_key(seq, p, seed)


def test_palindromic_seed_reverse_complement_symmetry():
    """For a palindromic pattern the masked key of a window equals that of the
    mirrored window in the reverse complement (the strand flag may flip)."""
    seed = 0x16df6d
    rng = np.random.default_rng(1)
    seq = "".join("ACGT"[x] for x in rng.integers(0, 4, 2000))
    rc = seq[::-1].translate(str.maketrans("ACGT", "TGCA"))
    k1 = po.keys_by_position(seq, seed)
    k2 = po.keys_by_position(rc, seed)
    sl = po.seed_length(seed)
    n = len(seq) - sl + 1
    assert np.array_equal(k1 >> np.uint64(1), k2[::-1] >> np.uint64(1))
    assert n == len(k1)


def test_sorted_invariants():
    rng = np.random.default_rng(2)
    seq = "".join("ACGT"[x] for x in rng.integers(0, 4, 20000))
    for seed in (0x16df6d, 0x1FFFF):
        _, keys, pos = po.sml_build(seq, seed)
        assert np.all(keys[1:] >= keys[:-1])
        assert np.array_equal(np.sort(pos), np.arange(len(pos)))
        kp = po.keys_by_position(seq, seed)
        assert np.array_equal(kp[pos], keys)
