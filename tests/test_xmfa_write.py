"""XMFA writer (mh_xmfa_write) on the CPU.

The expected text is spelled out from the stream operations of
WriteStandardAlignment (LM/IntervalList.h:353-442) and the extension table of
libGenome's gnSourceFactory (gnSourceFactory.cpp:34-46, 88-95).  A written
file is read back with the XMFA grammar of the index oracle and with the
library's own LUT walk.
"""
import numpy as np

import mauvehip
from mauvehip import index as mi
from oracle import index_oracle as io


def test_known_text(tmp_path):
    p = tmp_path / "a.xmfa"
    row_a = "ACGT-ACGTA"
    row_b = "ACGTTAC-TA"
    mauvehip.write_xmfa(p, [[(1, 9, row_a), (-101, 9, row_b)]], ["/d/g1.fasta", "/d/g2.gbk"])
    want = ("#FormatVersion Mauve1\n"
            "#Sequence1File\t/d/g1.fasta\n"
            "#Sequence1Format\tFastA\n"
            "#Sequence2File\t/d/g2.gbk\n"
            "#Sequence2Format\tGenBank\n"
            "#Annotation2File\t/d/g2.gbk\n"
            "#Annotation2Format\tGenBank\n"
            "> 1:1-9 + /d/g1.fasta\n"
            "ACGT-ACGTA\n"
            "> 2:101-109 - /d/g2.gbk\n"
            "ACGTTAC-TA\n"
            "=\n")
    assert p.read_text() == want


def test_formats_and_single_input(tmp_path):
    p = tmp_path / "b.xmfa"
    mauvehip.write_xmfa(p, [[(5, 2, "AC"), (7, 2, "AC")]], ["x.v1/genome", "x.v1/genome"], backbone_file="b.bb")
    lines = p.read_text().splitlines()
    # the same file for both genomes: Entry lines, and every entry is named after file 1
    assert lines[:8] == ["#FormatVersion Mauve1", "#Sequence1File\tx.v1/genome", "#Sequence1Entry\t1",
                         "#Sequence1Format\tFastA", "#Sequence2File\tx.v1/genome", "#Sequence2Entry\t2",
                         "#Sequence2Format\tFastA", "#BackboneFile\tb.bb"]
    p2 = tmp_path / "c.xmfa"
    mauvehip.write_xmfa(p2, [[(1, 1, "A"), (1, 1, "A"), (1, 1, "A")]], ["a.RAW", "b.seq", "c.GB"])
    t = p2.read_text()
    assert "#Sequence1Format\traw\n" in t and "#Annotation1" not in t
    assert "#Sequence2Format\tDNAstar\n#Annotation2File\tb.seq\n#Annotation2Format\tDNAstar\n" in t
    assert "#Sequence3Format\tGenBank\n#Annotation3File\tc.GB\n" in t


def test_absent_genomes_wrapping_and_empty_intervals(tmp_path):
    p = tmp_path / "d.xmfa"
    long_row = "A" * 161
    ivs = [
        [(0, 0, "-" * 3), (10, 3, "CCC")],        # absent genome 1 in the first interval: 0-0 +
        [(1, 0, ""), (1, 0, "")],                 # alignment length 0: skipped
        [(0, 0, "---"), (20, 3, "GGG")],          # absent in a later interval: left out
        [(50, 161, long_row), (-300, 161, long_row)],
    ]
    mauvehip.write_xmfa(p, ivs, ["g1.fa", "g2.fa"])
    body = p.read_text().split("#Sequence2Format\tFastA\n")[1]
    want = ("> 1:0-0 + g1.fa\n---\n> 2:10-12 + g2.fa\nCCC\n=\n"
            "> 2:20-22 + g2.fa\nGGG\n=\n"
            "> 1:50-210 + g1.fa\n" + "A" * 80 + "\n" + "A" * 80 + "\nA\n"
            "> 2:300-460 - g2.fa\n" + "A" * 80 + "\n" + "A" * 80 + "\nA\n=\n")
    assert body == want


def test_empty_alignment_writes_nothing(tmp_path):
    p = tmp_path / "e.xmfa"
    mauvehip.write_xmfa(p, [], ["a.fa", "b.fa"])
    assert p.read_bytes() == b""


def test_reads_back(tmp_path):
    rng = np.random.default_rng(8)
    ivs = []
    pos = [1, 1]
    for _ in range(6):
        n = int(rng.integers(1, 500))
        rows = ["".join(rng.choice(list("ACGT-"), n)) for _ in range(2)]
        iv = []
        for s in range(2):
            bases = n - rows[s].count("-")
            iv.append((pos[s], bases, rows[s]))
            pos[s] += bases + 5
        ivs.append(iv)
    p = tmp_path / "f.xmfa"
    mauvehip.write_xmfa(p, ivs, ["g1.fa", "g2.fa"])
    text = p.read_text()
    groups = io.parse_xmfa(text)
    assert len(groups) == len(ivs)
    for g, iv in zip(groups, ivs):
        assert [(e[0], e[1], e[3], e[5]) for e in g[0]] == [(s + 1, iv[s][0], "+", iv[s][2]) for s in range(2)]
        assert [e[2] for e in g[0]] == [iv[s][0] + iv[s][1] - 1 for s in range(2)]
    n = pos[0] + 10
    assert mi.lut_from_xmfa(text, n).tolist() == io.lut_from_groups(groups, n)
