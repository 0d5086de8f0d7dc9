""".sslist files (format version 5): mh_sml_save / mh_sml_load on the GPU.

The file is parsed with a ctypes restatement of SMLHeader
(LM/SortedMerList.h:48-63), so the header offsets are the C ABI's, not the
writer's own numbers.  Bar: the header fields FileSML::Create leaves, the
packed words and sorted positions equal to the oracle's; a loaded list equal
to the built one in every read-out and in the match list; every LoadFile2
status code (LM/FileSML.cpp:120-192) on a file damaged at the matching place.
"""
import ctypes as C
import os

import numpy as np
import pytest

import mauvehip
from mauvehip import synth
from oracle import pyoracle as po
from test_capi import SmlHeaderFile, SML_HEADER_BYTES

pytestmark = pytest.mark.gpu


def _header(path):
    with open(path, "rb") as f:
        return SmlHeaderFile.from_buffer_copy(f.read(SML_HEADER_BYTES))


def _n_words(length):
    return (2 * length + 31) // 32 + 2


def _seq(rng, L):
    return synth.BASES[rng.integers(0, 4, size=L)].tobytes()


CASES = [(po.get_seed(15, 3), L) for L in (0, 5, 21, 1000, 4097, 20001)] + \
        [(po.get_seed(11, 0), 777), (po.get_seed(19, 1), 3000), (po.get_seed(31, 0), 100)]


@pytest.mark.parametrize("seed,L", CASES)
def test_save_layout(ctx, tmp_path, seed, L):
    rng = np.random.default_rng(L + 17)
    seq = _seq(rng, L)
    sml = mauvehip.SortedMerList(ctx, seq, seed)
    path = str(tmp_path / "g.sslist")
    sml.save(path)
    words, keys, pos = po.sml_build(seq, seed)
    n = len(keys)
    assert os.path.getsize(path) == SML_HEADER_BYTES + 4 * _n_words(L) + 4 * n
    h = _header(path)
    assert (h.version, h.alphabet_bits, h.seed) == (5, 2, seed)
    assert (h.seed_length, h.seed_weight, h.length) == (po.seed_length(seed), po.seed_weight(seed), L)
    assert h.unique_mers == 0xFFFFFFFF and h.id == 0 and h.circular == 0
    with open(path, "rb") as f:
        assert f.read(SmlHeaderFile.description.offset + 2048)[SmlHeaderFile.description.offset:] == bytes(2048)
    raw = np.fromfile(path, dtype=np.uint32, offset=SML_HEADER_BYTES)
    assert np.array_equal(raw[:_n_words(L)], words)
    assert np.array_equal(raw[_n_words(L):], pos)
    sml.free()


def test_translation_table_is_the_packing(ctx, tmp_path):
    """Each byte of the stored translation table is the 2-bit code the packing
    gives that character (every byte value but NUL and the gap '-')."""
    sml = mauvehip.SortedMerList(ctx, b"ACGTACGTACGTACGTACGT", po.get_seed(11, 0))
    path = str(tmp_path / "t.sslist")
    sml.save(path)
    table = bytes(_header(path).translation_table)
    assert len(table) == 255
    for c in range(1, 255):
        if c == ord("-"):
            continue
        assert table[c] == int(po.packed(bytes([c]))[0]) >> 30, chr(c)
    sml.free()


def _same_list(ctx, a, b, seed, rng):
    ha, hb = a.header(), b.header()
    for f in ("version", "alphabet_bits", "seed", "seed_length", "seed_weight", "length", "sml_length",
              "circular"):
        assert getattr(ha, f) == getattr(hb, f), f
    pa, ka = a.read()
    pb, kb = b.read()
    assert np.array_equal(pa, pb) and np.array_equal(ka, kb)
    assert np.array_equal(a.sequence_words(), b.sequence_words())
    n = ha.sml_length
    if n:
        q = rng.integers(0, n, size=min(n, 500)).astype(np.uint32)
        assert np.array_equal(a.get_seed_mers(q), b.get_seed_mers(q))
        for mer in ka[rng.integers(0, n, size=5)]:
            assert a.find_mer(int(mer)) == b.find_mer(int(mer))


@pytest.mark.parametrize("seed,L", CASES)
def test_round_trip(ctx, tmp_path, seed, L):
    rng = np.random.default_rng(L + 29)
    seq = _seq(rng, L)
    built = mauvehip.SortedMerList(ctx, seq, seed)
    path = str(tmp_path / "r.sslist")
    built.save(path)
    loaded = mauvehip.SortedMerList.load(ctx, path)
    _same_list(ctx, built, loaded, seed, rng)
    _, keys, pos = po.sml_build(seq, seed)
    gp, gk = loaded.read()
    assert np.array_equal(gk, keys) and np.array_equal(gp, pos)
    built.free()
    loaded.free()


def test_loaded_lists_find_the_same_matches(ctx, tmp_path):
    rng = np.random.default_rng(5)
    seed = po.get_seed(15, 3)
    a = synth.BASES[rng.integers(0, 4, size=50000)]
    b = a.copy()
    hit = rng.random(len(b)) < 0.02
    b[hit] = synth.BASES[rng.integers(0, 4, size=int(hit.sum()))]
    b[20000:26000] = synth.revcomp(b[20000:26000])
    built = [mauvehip.SortedMerList(ctx, s.tobytes(), seed) for s in (a, b)]
    paths = [str(tmp_path / f"{k}.sslist") for k in "ab"]
    for s, p in zip(built, paths):
        s.save(p)
    loaded = [mauvehip.SortedMerList.load(ctx, p) for p in paths]
    s1, l1, st1 = mauvehip.find_matches(ctx, built)
    s2, l2, st2 = mauvehip.find_matches(ctx, loaded)
    assert len(l1) > 10
    assert np.array_equal(s1, s2) and np.array_equal(l1, l2)
    assert (st1.seed_tuples, st1.collisions) == (st2.seed_tuples, st2.collisions)
    # the --mums file of the match list reads back to the same list
    mums = str(tmp_path / "ab.mums")
    mauvehip.write_match_list(mums, s1, l1, paths, [len(a), len(b)])
    s3, l3, f3, n3 = mauvehip.read_match_list(mums)
    assert np.array_equal(s3, s1) and np.array_equal(l3, l1) and f3 == paths and n3 == [len(a), len(b)]
    for s in built + loaded:
        s.free()


def test_unwritten_padding_is_ignored(ctx, tmp_path):
    """The reference leaves the sequence's padding words (and the bits after
    the last character) unwritten: whatever they hold, the loaded list is the
    same."""
    rng = np.random.default_rng(9)
    seed = po.get_seed(13, 3)
    L = 1001                                   # 2002 bits: the last word is partly sequence
    seq = _seq(rng, L)
    built = mauvehip.SortedMerList(ctx, seq, seed)
    path = str(tmp_path / "p.sslist")
    built.save(path)
    raw = bytearray(open(path, "rb").read())
    nw = _n_words(L)
    w = np.frombuffer(raw, dtype=np.uint32, count=nw, offset=SML_HEADER_BYTES).copy()
    w[2002 // 32] |= (1 << (32 - 2002 % 32)) - 1
    w[2002 // 32 + 1:] = 0xFFFFFFFF
    raw[SML_HEADER_BYTES:SML_HEADER_BYTES + 4 * nw] = w.tobytes()
    raw[40:41] = b"\x00"                       # little_endian, word_size: ignored
    raw[36:40] = np.uint32(0xDEADBEEF).tobytes()
    open(path, "wb").write(bytes(raw))
    loaded = mauvehip.SortedMerList.load(ctx, path)
    _same_list(ctx, built, loaded, seed, rng)
    built.free()
    loaded.free()


def _damaged(tmp_path, raw, name):
    p = str(tmp_path / name)
    with open(p, "wb") as f:
        f.write(bytes(raw))
    return p


def test_load_status_codes(ctx, tmp_path):
    rng = np.random.default_rng(11)
    seed = po.get_seed(15, 3)
    L = 3000
    seq = _seq(rng, L)
    sml = mauvehip.SortedMerList(ctx, seq, seed)
    path = str(tmp_path / "s.sslist")
    sml.save(path)
    sml.free()
    raw = bytearray(open(path, "rb").read())
    nw = _n_words(L)
    pos0 = SML_HEADER_BYTES + 4 * nw
    _, keys, pos = po.sml_build(seq, seed)

    def status(p):
        with pytest.raises(mauvehip.SmlLoadError) as e:
            mauvehip.SortedMerList.load(ctx, p)
        return e.value.status, e.value.code

    assert status(str(tmp_path / "missing.sslist")) == (1, -7)
    assert status(_damaged(tmp_path, raw[:SML_HEADER_BYTES - 1], "h.sslist")) == (2, -7)
    v4 = raw.copy()
    v4[0:4] = np.uint32(4).tobytes()
    assert status(_damaged(tmp_path, v4, "v.sslist")) == (3, -7)
    assert status(_damaged(tmp_path, raw[:pos0 - 1], "w.sslist")) == (4, -7)
    assert status(_damaged(tmp_path, raw[:len(raw) - 4], "p.sslist")) == (5, -7)
    circ = raw.copy()
    circ[44] = 1
    assert status(_damaged(tmp_path, circ, "c.sslist")) == (6, -6)
    ab = raw.copy()
    ab[4:8] = np.uint32(5).tobytes()
    assert status(_damaged(tmp_path, ab, "a.sslist")) == (6, -6)
    sw = raw.copy()
    sw[20:24] = np.uint32(po.seed_weight(seed) + 1).tobytes()
    assert status(_damaged(tmp_path, sw, "sw.sslist")) == (6, -6)
    dup = raw.copy()
    dup[pos0 + 4:pos0 + 8] = dup[pos0:pos0 + 4]
    assert status(_damaged(tmp_path, dup, "d.sslist")) == (7, -1)
    big = raw.copy()
    big[pos0:pos0 + 4] = np.uint32(len(pos)).tobytes()
    assert status(_damaged(tmp_path, big, "b.sslist")) == (7, -1)
    # two neighbours with different keys swapped: out of order
    k = int(np.nonzero(keys[1:] != keys[:-1])[0][0])
    sw2 = raw.copy()
    p = np.frombuffer(sw2, dtype=np.uint32, count=len(pos), offset=pos0).copy()
    p[k], p[k + 1] = p[k + 1], p[k]
    sw2[pos0:] = p.tobytes()
    assert status(_damaged(tmp_path, sw2, "o.sslist")) == (7, -1)
    # neighbours with equal keys swapped: still a sorted list (the
    # reference's std::sort leaves ties in any order); the order is kept
    eq = np.nonzero(keys[1:] == keys[:-1])[0]
    if len(eq):
        k = int(eq[0])
        tie = raw.copy()
        p = np.frombuffer(tie, dtype=np.uint32, count=len(pos), offset=pos0).copy()
        p[k], p[k + 1] = p[k + 1], p[k]
        tie[pos0:] = p.tobytes()
        s = mauvehip.SortedMerList.load(ctx, _damaged(tmp_path, tie, "t.sslist"))
        gp, gk = s.read()
        assert np.array_equal(gp, p) and np.array_equal(gk, keys)
        s.free()
    # extra bytes after the list are not read
    s = mauvehip.SortedMerList.load(ctx, _damaged(tmp_path, raw + b"\x01" * 12, "x.sslist"))
    assert np.array_equal(s.read()[0], pos)
    s.free()


def test_load_smls_creates_then_loads(ctx, tmp_path):
    """GenericMatchList::LoadSMLs: missing files are created, present ones
    loaded, a seed mismatch recreates."""
    rng = np.random.default_rng(13)
    seqs = [_seq(rng, 5000), _seq(rng, 4000)]
    files = [str(tmp_path / "a.fa"), str(tmp_path / "b.fa")]
    seed = po.get_seed(15, 3)
    names = mauvehip.default_sml_file_names(files, 15, 3)
    assert names[0] == files[0] + "." + format(seed, "b") + ".sslist"
    msgs = []
    first = mauvehip.load_smls(ctx, seqs, names, seed, log=msgs.append)
    assert msgs == ["Creating sorted mer list"] * 2 and all(os.path.exists(n) for n in names)
    msgs.clear()
    again = mauvehip.load_smls(ctx, seqs, names, seed, log=msgs.append)
    assert msgs == ["Sorted mer list loaded successfully"] * 2
    for x, y in zip(first, again):
        _same_list(ctx, x, y, seed, rng)
    msgs.clear()
    seed2 = po.get_seed(13, 3)
    other = mauvehip.load_smls(ctx, seqs, names, seed2, log=msgs.append)
    assert msgs == ["Default seed mismatch.  A new sorted mer list will be created.", "Creating sorted mer list"] * 2
    assert all(s.seed() == seed2 for s in other) and _header(names[0]).seed == seed2
    for s in first + again + other:
        s.free()
