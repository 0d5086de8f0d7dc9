"""buildIndex run end to end on the GPU from two FASTA files: the same LUT as
the pairwise path with progressiveMauve's default seed followed by the LUT
passes on its XMFA file (each stage of which tests/test_gpu_pipeline.py checks
against the oracle)."""
import numpy as np
import pytest

import mauvehip
from mauvehip import buildindex as bi
from mauvehip import index as mindex
from mauvehip import pipeline, synth

pytestmark = pytest.mark.gpu


def _fasta(path, name, seq):
    with open(path, "wb") as f:
        f.write(b">" + name + b"\n")
        for i in range(0, len(seq), 80):
            f.write(seq[i:i + 80] + b"\n")


def test_build_index_from_fasta(ctx, tmp_path):
    ref, mut = synth.pair_iid()
    g, r = np.asarray(mut).tobytes(), np.asarray(ref).tobytes()
    gfp, rfp = str(tmp_path / "mut.fa"), str(tmp_path / "ref.fa")
    _fasta(gfp, b"mut", g)
    _fasta(rfp, b"ref", r)
    lut = bi.buildIndex(gfp, rfp)
    assert lut.dtype == mindex.IDX_ARRAY_DTYPE and len(lut) == len(g)
    # the same composition by hand
    seed = mauvehip.get_seed(mauvehip.get_default_seed_weight((len(g) + len(r)) // 2), mauvehip.CODING_SEED)
    d = [ctx.alloc(len(g) + 16), ctx.alloc(len(r) + 16)]
    try:
        ctx.h2d(d[0], np.frombuffer(g, dtype=np.uint8))
        ctx.h2d(d[1], np.frombuffer(r, dtype=np.uint8))
        xf = str(tmp_path / "by_hand.xmfa")
        pipeline.align_pair(ctx, d, (g, r), seed, xf, [gfp, rfp], lut=False)
    finally:
        for p in d:
            ctx.free(p)
    want = mindex.build_index_from_xmfa(xf, gfp, rfp)
    assert np.array_equal(lut, want)
    # a related pair (1% SNPs, 0.1% indels): nearly every position maps, and
    # mapped positions mostly carry the reference's base
    m = lut >= 0
    assert m.mean() > 0.95
    gb, rb = np.frombuffer(g, dtype=np.uint8), np.frombuffer(r, dtype=np.uint8)
    assert (gb[m] == rb[lut[m]]).mean() > 0.97
    # no-pass variant: the XMFA column walk alone maps only columns of equal
    # bases; its indices count residues from the start index
    # (PY/buildindex.py:116-129: genome_idx = start_idx - 1, then += 1 on a
    # residue before the mapping), so entry p maps base p - 1 to base
    # raw[p] - 1
    raw = bi.buildIndex(gfp, rfp, fill_gaps=False, smooth_edges=False)
    p = np.nonzero(raw >= 1)[0]
    p = p[p >= 1]
    assert len(p) > 0.95 * len(g)
    assert np.array_equal(gb[p - 1], rb[raw[p] - 1])
