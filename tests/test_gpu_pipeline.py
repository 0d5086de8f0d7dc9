"""The pairwise path end to end on the GPU (mauvehip.pipeline.align_pair) on
the 200 kb i.i.d. pair (1% SNPs, 0.1%/bp indels, one 5 kb inversion, seed
weight 15): every stage checked against the oracle on its own input.

* the rows keep every residue of both genomes in order;
* the refined rows equal the refinement oracle run on the unrefined rows;
* the XMFA file parses (oracle grammar) to those rows;
* the LUT equals the index oracle's walk and passes over that file.
"""
import os

import numpy as np
import pytest

import mauvehip
from mauvehip import pipeline, synth
from oracle import index_oracle as io
from oracle import pyoracle as po
from oracle import refine_oracle as ro

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(ctx):
    ref, mut = synth.pair_iid()
    g, r = np.asarray(mut).tobytes(), np.asarray(ref).tobytes()
    d = [ctx.alloc(len(g) + 16), ctx.alloc(len(r) + 16)]
    ctx.h2d(d[0], np.frombuffer(g, dtype=np.uint8))
    ctx.h2d(d[1], np.frombuffer(r, dtype=np.uint8))
    yield d, (g, r)
    for p in d:
        ctx.free(p)


def test_pipeline_stages(ctx, pair, tmp_path):
    d, (g, r) = pair
    seed = mauvehip.get_seed(15, 3)
    names = ["mut.fa", "ref.fa"]
    ua, ub, _, info0 = pipeline.align_pair(ctx, d, (g, r), seed, str(tmp_path / "u.xmfa"), names, refine=False,
                                           lut=False)
    ua, ub = ua.tobytes(), ub.tobytes()
    assert ua.replace(b"-", b"") == g and ub.replace(b"-", b"") == r
    assert info0["chain_anchors"] > 100
    tm = {}
    ra, rb, lut, info = pipeline.align_pair(ctx, d, (g, r), seed, str(tmp_path / "a.xmfa"), names, timings=tm)
    ra, rb = ra.tobytes(), rb.tobytes()
    assert {k for k in tm if "." not in k} == {"sml", "match", "chain", "recurse", "refine", "xmfa", "lut"}
    assert {k for k in tm if "." in k} <= {"chain.dp", "chain.rows", "recurse.gaps", "recurse.search",
                                           "recurse.chain", "lut.walk", "lut.passes"}
    assert ra.replace(b"-", b"") == g and rb.replace(b"-", b"") == r
    want = ro.gapped_alignment(ua.decode(), ub.decode())
    assert (ra.decode(), rb.decode()) == want
    text = open(tmp_path / "a.xmfa").read()
    groups = io.parse_xmfa(text)
    assert len(groups) == 1
    assert [e[5] for e in groups[0][0]] == [ra.decode(), rb.decode()]
    wl = io.lut_from_groups(groups, len(g))
    io.fix_zero_idx(wl, g.decode(), r.decode())
    io.fill_gaps(wl, 300)
    io.smooth_edges(wl, 20)
    assert lut.tolist() == wl
    assert (lut >= 0).mean() > 0.9


def test_pipeline_recursive_anchors_are_the_oracles(ctx, pair):
    d, (g, r) = pair
    seed = mauvehip.get_seed(15, 3)
    sa = mauvehip.SortedMerList.from_device(ctx, d[0], len(g), seed)
    sb = mauvehip.SortedMerList.from_device(ctx, d[1], len(r), seed)
    mt, _ = mauvehip.find_matches_raw(ctx, [sa, sb])
    sa.free()
    sb.free()
    chain, _ = mauvehip.chain_gap_jobs_raw(mt, len(g), len(r), 0)
    ra, rb = mauvehip.chain_rows(mt, chain, g, r)
    gaps = mauvehip.recursion_gaps(ra, rb, 1, 1, 200)
    want = po.recursion_gaps(ra.decode(), rb.decode(), 1, 1, 200)
    assert [(tuple(int(x) for x in q["lend"]), tuple(int(x) for x in q["rend"])) for q in gaps] == want
    st, ln, first = mauvehip.anchor_gaps_device(ctx, d[0], len(g), d[1], len(r), gaps)
    for i, (lend, rend) in enumerate(want):
        ws, wl = po.anchor_gap(g, r, lend, rend)
        assert np.array_equal(st[first[i]:first[i + 1]], ws)
        assert np.array_equal(ln[first[i]:first[i + 1]], wl)
