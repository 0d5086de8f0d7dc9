"""Match-list files, format version 3 (`--mums` output and `--match-input`),
through the C ABI on the CPU (mh_match_list_write / mh_match_list_read).

The expected text below is spelled out from the stream operations of
WriteList (LM/MatchList.h:617-662) and UngappedLocalAlignment's operator<<
(LM/UngappedLocalAlignment.h:200-206): length, then each start, tab
separated; then the id, the subset count 0 and the superset count 0.  The
reader is checked on files in the reference's own form (hexadecimal match
addresses as ids) and on each malformed input ReadList rejects
(LM/MatchList.h:524-613).
"""
import numpy as np
import pytest

import mauvehip

STARTS = np.array([[1, 5], [-120, 77], [1000000, -4000000000], [7, 0]], dtype=np.int64)
LENGTHS = np.array([21, 15, 300, 12], dtype=np.uint64)


def test_known_text(tmp_path):
    p = tmp_path / "m.mums"
    mauvehip.write_match_list(p, STARTS, LENGTHS, ["a.fa", None], [5000, 4000000100])
    want = ("FormatVersion\t3\n"
            "SequenceCount\t2\n"
            "Sequence0File\ta.fa\n"
            "Sequence0Length\t5000\n"
            "Sequence1File\tnull\n"
            "Sequence1Length\t4000000100\n"
            "MatchCount\t4\n"
            "21\t1\t5\t0x1\t0\t0\n"
            "15\t-120\t77\t0x2\t0\t0\n"
            "300\t1000000\t-4000000000\t0x3\t0\t0\n"
            "12\t7\t0\t0x4\t0\t0\n")
    assert p.read_text() == want


def test_unknown_names_and_lengths(tmp_path):
    p = tmp_path / "m.mums"
    mauvehip.write_match_list(p, STARTS[:1], LENGTHS[:1])
    assert p.read_text().splitlines()[2:6] == ["Sequence0File\tnull", "Sequence0Length\t0",
                                               "Sequence1File\tnull", "Sequence1Length\t0"]


def test_empty_list_writes_nothing(tmp_path):
    p = tmp_path / "m.mums"
    mauvehip.write_match_list(p, np.zeros((0, 2), dtype=np.int64), np.zeros(0, dtype=np.uint64))
    assert p.exists() and p.read_bytes() == b""


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    for k in (2, 3, 4):
        n = 500
        st = rng.integers(-10 ** 9, 10 ** 9, size=(n, k))
        st[rng.random((n, k)) < 0.1] = 0
        ln = rng.integers(1, 10 ** 6, size=n).astype(np.uint64)
        files = [f"/data/genome {i}.fa" for i in range(k)]
        lens = list(rng.integers(1, 10 ** 10, size=k))
        p = tmp_path / f"r{k}.mums"
        mauvehip.write_match_list(p, st, ln, files, lens)
        s2, l2, f2, n2 = mauvehip.read_match_list(p)
        assert np.array_equal(s2, st) and np.array_equal(l2, ln)
        assert f2 == files and n2 == lens


def test_reads_the_reference_form(tmp_path):
    """Ids as the reference prints them (object addresses, or 0)."""
    txt = ("FormatVersion\t3\nSequenceCount\t2\nSequence0File\t/tmp/x y.fa\nSequence0Length\t10\n"
           "Sequence1File\tnull\nSequence1Length\t0\nMatchCount\t2\n"
           "9\t1\t-2\t0x55d3c0a8e2f0\t0\t0\n12\t3\t4\t0\t0\t0\n")
    p = tmp_path / "ref.mums"
    p.write_text(txt)
    st, ln, files, lens = mauvehip.read_match_list(p)
    assert st.tolist() == [[1, -2], [3, 4]] and ln.tolist() == [9, 12]
    assert files == ["/tmp/x y.fa", "null"] and lens == [10, 0]


@pytest.mark.parametrize("text,code", [
    ("FormatVersion\t2\nSequenceCount\t2\n", -7),
    ("Format\t3\n", -7),
    ("FormatVersion\t3\nSequenceCount\t1\nSequence0File\ta\nSequence0Length\t1\nMatchCount\t0\n", -7),
    ("FormatVersion\t3\nSequenceCount\t2\nSequence0File\ta\nSequence0Length\t1\nSequence1File\tb\n"
     "Sequence1Length\t1\nMatchCount\t2\n5\t1\t1\t0x1\t0\t0\n", -7),              # count differs
    ("FormatVersion\t3\nSequenceCount\t2\nSequence0File\ta\nSequence0Length\t1\nSequence1File\tb\n"
     "Sequence1Length\t1\nMatchCount\t1\n5\t1\t1\t0x1\t1\t0\n", -7),              # subset data
    ("FormatVersion\t3\nSequenceCount\t5\n", -6),
])
def test_rejects(tmp_path, text, code):
    p = tmp_path / "bad.mums"
    p.write_text(text)
    with pytest.raises(mauvehip.MhError) as e:
        mauvehip.read_match_list(p)
    assert e.value.code == code


def test_missing_file(tmp_path):
    with pytest.raises(mauvehip.MhError) as e:
        mauvehip.read_match_list(tmp_path / "none.mums")
    assert e.value.code == -7
