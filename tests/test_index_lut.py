"""The library's buildIndex LUT (mh_index_*, mauvehip.index) against the
oracle, on the CPU: bit-exact LUTs from XMFA text (the reference's fixture,
generated alignments, grammar corner cases) and after each pass on random
tables with holes, shifts and stray entries."""
import os

import numpy as np
import pytest

from mauvehip import index as mi
from oracle import index_oracle as io

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _both_walks(text, n):
    got = mi.lut_from_xmfa(text, n)
    want = io.lut_from_groups(io.parse_xmfa(text), n)
    assert got.tolist() == want
    return got


def test_example_fixture():
    text = open(os.path.join(GOLDEN, "example.xmfa")).read()
    lut = _both_walks(text, 1100)
    assert (lut != -1).sum() > 300


def _random_xmfa(rng, n_groups, glen, rlen, wrap=60):
    """Groups of random aligned rows (gaps, mismatches, lowercase) with random
    entry order, extra genomes and groups that lack genome 2."""
    alpha = np.array(list("ACGTacgtN-"))
    out = ["#FormatVersion Mauve1\n#Sequence1File\tg.fa\n"]
    for _ in range(n_groups):
        cols = int(rng.integers(1, 400))
        entries = []
        for num in rng.permutation([1, 2, 3])[:int(rng.integers(1, 4))]:
            row = "".join(alpha[rng.choice(10, cols, p=[.2, .2, .2, .2, .02, .02, .02, .02, .02, .1])])
            start = int(rng.integers(0, glen if num == 1 else rlen))
            strand = "+-|"[int(rng.integers(0, 3))]
            body = "\n".join(row[k:k + wrap] for k in range(0, len(row), wrap))
            entries.append(f"> {num}:{start}-{start + cols} {strand} /x/genome {num}.fa\n{body}\n")
        if rng.random() < 0.3 and len(entries) > 1:
            entries[0], entries[-1] = entries[-1], entries[0]
        out.append("".join(entries) + "=" + (" score=%d\n" % rng.integers(0, 999)))
    return "".join(out)


def test_random_alignments():
    rng = np.random.default_rng(7)
    for trial in range(20):
        glen, rlen = int(rng.integers(50, 3000)), int(rng.integers(50, 3000))
        text = _random_xmfa(rng, int(rng.integers(1, 12)), glen, rlen)
        _both_walks(text, glen)


@pytest.mark.parametrize("text", [
    "> 1:3-9 +\tnote\tx\nACGTACG\n> 2:1-7 + r\nACGTACG\n=\n",        # tab inside notes: entry 1 skipped
    "> 1:3-9 + \nACGTACG\nTTT\n> 2:1-7 + r\nACGTACG\nTTT\n=\n",      # notes taken from the next line
    ">1:3-9 | g\r\nACGTACG\r\n>2:1-7 - r\r\nACGTTCG\r\n=\r\n",       # '|' strand, CRLF lines
    "> 1:3-9 + g\nAC>GTACG\n> 2:1-7 + r\nACGTACG\n=\n",              # '>' inside a row ends it
    "> 1:3-9 + g\nACGT=ACG\n> 2:1-7 + r\nACGTACG\n=\n",              # '=' inside a row ends the group
    ">=\n> 1:1-4 + g\nACGT\n> 2:1-4 + r\nACGT\n= s\n",               # '>' then '=': no group there
    "> 1:1-4 + g\n\n> 2:1-4 + r\nACGT\n=\n",                        # empty row
    "> 1:0-4 + g\n-ACG\n> 2:5-9 + r\n-ACG\n=\n",                      # start 0, gap column: index -1 wraps
    "x > 1:1-4 + g\nACGT\n> 2:1-4 + r\nACGT\n=\n",                   # not at a line start
    "> 1:1-4 + g\nACGT\n> 2:1-4 + r\nACGT\n",                        # no '=' at all: no group
    "",
])
def test_grammar_corners(text):
    _both_walks(text, 12)


def test_short_reference_row_is_an_error():
    with pytest.raises(Exception):
        mi.lut_from_xmfa("> 1:1-4 + g\nACGT\n> 2:1-3 + r\nACG\n=\n", 10)


def _random_lut(rng, n):
    """Runs of consecutive mapping with holes, shifts and stray values."""
    lut = np.full(n, -1, dtype=np.int32)
    k, v = 0, int(rng.integers(0, 50))
    while k < n:
        run = int(rng.integers(1, 80))
        lut[k:k + run] = np.arange(v, v + min(run, n - k))
        k += run
        v += run
        u = rng.random()
        if u < 0.3:
            k += int(rng.integers(1, 40))                     # hole
            v += int(rng.integers(0, 45))
        elif u < 0.5:
            v += int(rng.integers(-5, 30))                    # shift
        elif u < 0.7 and k < n:
            lut[k] = int(rng.integers(0, 5 * n))              # stray entry
            k += 1
            v += 1
    return lut


@pytest.mark.parametrize("seed", range(12))
def test_passes_match_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(20, 4000))
    lut = _random_lut(rng, n)
    width, radius = int(rng.choice([1, 5, 40, 300])), int(rng.choice([1, 3, 20]))
    a = lut.copy()
    mi.fillGaps(a, max_gap_width=width)
    b = lut.tolist()
    io.fill_gaps(b, width)
    assert a.tolist() == b
    mi.smoothEdges(a, smoothing_radius=radius)
    io.smooth_edges(b, radius)
    assert a.tolist() == b


def test_fix_zero_matches_oracle():
    rng = np.random.default_rng(5)
    for trial in range(200):
        n = int(rng.integers(11, 40))
        lut = np.full(n, -1, dtype=np.int32)
        for k in rng.choice(n, int(rng.integers(1, n)), replace=False):
            lut[k] = int(rng.integers(0, 30))
        g = "".join(rng.choice(list("ACGT"), n))
        r = "".join(rng.choice(list("ACGT"), 80))
        a, b = lut.copy(), lut.tolist()
        try:
            io.fix_zero_idx(b, g, r)
            err = None
        except IndexError:
            err = IndexError
        if err:
            with pytest.raises(IndexError):
                mi.fixZeroIdx(a, g, r)
        else:
            mi.fixZeroIdx(a, g, r)
            assert a.tolist() == b


def test_build_index_from_xmfa(tmp_path):
    rng = np.random.default_rng(17)
    G = "".join(rng.choice(list("ACGT"), 1500))
    R = "".join(rng.choice(list("ACGT"), 1600))
    (tmp_path / "g.fa").write_text(">g genome\n" + "\n".join(G[k:k + 70] for k in range(0, len(G), 70)) + "\n")
    (tmp_path / "r.fa").write_text(">r\n" + R + "\n")
    text = _random_xmfa(rng, 8, len(G), len(R))
    (tmp_path / "a.xmfa").write_text(text)
    lut = mi.build_index_from_xmfa(str(tmp_path / "a.xmfa"), str(tmp_path / "g.fa"), str(tmp_path / "r.fa"),
                                   max_gap_width=50, smoothing_radius=7) \
        if _has_mapped_past_10(text, len(G)) else None
    if lut is None:
        pytest.skip("generated alignment maps nothing at or after index 10")
    want = io.lut_from_groups(io.parse_xmfa(text), len(G))
    io.fix_zero_idx(want, G, R)
    io.fill_gaps(want, 50)
    io.smooth_edges(want, 7)
    assert lut.tolist() == want
    assert lut.dtype == np.int32


def _has_mapped_past_10(text, n):
    lut = io.lut_from_groups(io.parse_xmfa(text), n)
    return lut[0] != -1 or any(v != -1 for v in lut[10:])


def test_helpers():
    lut = np.array([-1, 5, 6, 7, 20, 21, -1, 9], dtype=np.int32)
    assert mi.findEdges(lut).tolist() == [1, 1, 0, 0, 1, 0, 1, 1]
    assert mi.findMismatches(lut, "AAAAAAAA", "CCCCCAACCCCCCCCCCCCCCC").tolist() == [0, 0, 0, 1, 1, 1, 0, 1]
    assert mi.findDuplicateMappings(np.array([3, 3, 4, -1, 0, 5, 5], dtype=np.int32)) == 2


def test_long_rows_walked_in_chunks():
    """Rows longer than the library's 2^18-column walk chunks, with all-gap
    columns placed at the chunk borders (their writes must keep column order)."""
    rng = np.random.default_rng(23)
    n = 600_000
    alpha = np.array(list("ACGT-"))
    g = alpha[rng.choice(5, n, p=[.24, .24, .24, .24, .04])]
    r = g.copy()
    flip = rng.random(n) < 0.05
    r[flip] = alpha[rng.integers(0, 5, int(flip.sum()))]
    for b in (262144, 524288):
        g[b - 3:b + 3] = "-"
        r[b - 3:b + 3] = "-"
    gs, rs = "".join(g), "".join(r)
    body = lambda s: "\n".join(s[k:k + 80] for k in range(0, len(s), 80))
    text = f"> 1:5-{n} + g\n{body(gs)}\n> 2:9-{n} + r\n{body(rs)}\n=\n"
    _both_walks(text, n + 10)


@pytest.mark.parametrize("seed", range(2))
def test_passes_across_chunks(seed):
    """fillGaps and smoothEdges over more than 2^18 entries (the library finds
    fills in parallel chunks and keeps them in order): gaps and stray entries
    placed at and around the chunk borders, where a fill found in one chunk
    covers the first triggers of the next."""
    rng = np.random.default_rng(300 + seed)
    n = 700_000
    lut = _random_lut(rng, n)
    for b in (262144, 524288):
        lut[b - 5:b + 7] = -1
        lut[b + 9] = int(rng.integers(0, 5 * n))
    a = lut.copy()
    mi.fillGaps(a)
    b = lut.tolist()
    io.fill_gaps(b, 300)
    assert a.tolist() == b
    mi.smoothEdges(a)
    io.smooth_edges(b, 20)
    assert a.tolist() == b


def test_lut_from_file_matches_text(tmp_path):
    """The library's own read of the file gives the walk of the text, for a
    file longer than one 4 MiB read piece; a missing file is an I/O error."""
    rng = np.random.default_rng(31)
    n = 3_000_000
    alpha = np.array(list("ACGT-"))
    g = alpha[rng.choice(5, n, p=[.24, .24, .24, .24, .04])]
    r = g.copy()
    flip = rng.random(n) < 0.05
    r[flip] = alpha[rng.integers(0, 5, int(flip.sum()))]
    gs, rs = "".join(g), "".join(r)
    body = lambda s: "\n".join(s[k:k + 80] for k in range(0, len(s), 80))
    text = f"#FormatVersion Mauve1\n> 1:1-{n} + g\n{body(gs)}\n> 2:1-{n} + r\n{body(rs)}\n=\n"
    path = tmp_path / "a.xmfa"
    path.write_text(text)
    assert path.stat().st_size > (4 << 20)
    assert np.array_equal(mi.lut_from_xmfa_file(str(path), n + 5), mi.lut_from_xmfa(text, n + 5))
    with pytest.raises(mi.MhError):
        mi.lut_from_xmfa_file(str(tmp_path / "missing.xmfa"), 10)


def test_rows_with_scattered_whitespace():
    """Sequence text with every whitespace character the grammar removes
    (space, tab, CR, VT, FF, LF) at random places and in runs, over more than
    one squeeze chunk: the library removes them eight bytes at a time where a
    group holds none."""
    rng = np.random.default_rng(77)
    n = 400_000
    alpha = np.array(list("ACGT-"))
    g = alpha[rng.choice(5, n, p=[.24, .24, .24, .24, .04])]
    r = g.copy()
    flip = rng.random(n) < 0.05
    r[flip] = alpha[rng.integers(0, 5, int(flip.sum()))]
    ws = np.array([" ", "\t", "\r", "\v", "\f", "\n", "  \t", "\r\n"])

    def body(seq):
        out = list(seq)
        for at in sorted(rng.choice(len(out), 30_000, replace=False), reverse=True):
            out.insert(int(at), ws[rng.integers(0, len(ws))])
        return "".join(out)

    text = f"> 1:3-{n + 2} + g\n{body(g)}\n> 2:7-{n + 6} + r\n{body(r)}\n=\n"
    _both_walks(text, n + 10)
