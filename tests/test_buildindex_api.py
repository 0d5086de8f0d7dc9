"""buildIndex / runMauve front end (mauvehip.buildindex): the reference's
signatures (PY/buildindex.py:56-138) and the argument checks that run before
any GPU work."""
import inspect

import pytest

from mauvehip import buildindex as bi


def test_signatures_match_the_reference():
    # PY/buildindex.py:90-92 and 56
    p = inspect.signature(bi.buildIndex).parameters
    assert list(p) == ["genome_fp", "ref_genome_fp", "genome_seq", "ref_genome_seq", "fill_gaps", "max_gap_width",
                       "smooth_edges", "smoothing_radius"]
    assert [p[k].default for k in list(p)[2:]] == [None, None, True, 300, True, 20]
    q = inspect.signature(bi.runMauve).parameters
    assert list(q) == ["fasta_files", "flags"] and q["flags"].default == {}
    assert list(inspect.signature(bi.getGenomeLength).parameters) == ["genome_fp"]


def test_genome_length(tmp_path):
    fp = tmp_path / "g.fa"
    fp.write_text(">g\nACGT\nAC\n")
    assert bi.getGenomeLength(str(fp)) == 6


def test_run_mauve_argument_checks(tmp_path):
    fa = []
    for n in ("a", "b", "c"):
        fp = tmp_path / (n + ".fa")
        fp.write_text(">x\nACGTACGT\n")
        fa.append(str(fp))
    with pytest.raises(ValueError, match="two genomes"):
        bi.runMauve(fa, {"--output": str(tmp_path / "o.xmfa")})
    with pytest.raises(ValueError, match="not supported"):
        bi.runMauve(fa[:2], {"--output": str(tmp_path / "o.xmfa"), "--seed-family": ""})
    with pytest.raises(ValueError, match="--output"):
        bi.runMauve(fa[:2], {})
    empty = tmp_path / "e.fa"
    empty.write_text(">e\n")
    with pytest.raises(ValueError, match="no sequence"):
        bi.runMauve([fa[0], str(empty)], {"--output": str(tmp_path / "o.xmfa")})
