"""The pairwise path end to end: two genomes in, XMFA and index LUT out.

Stages, in the order progressiveMauve runs them for two genomes (each through
the library; the anchoring and the DP on the GPU):

  sml        sorted mer lists of both genomes          mh_sml_create_device
  match      pairwise seed matching                    mh_find_matches_pairwise_list
  chain      anchors of one interval                   mh_chain_gap_jobs (stand-in for
                                                       the LCB pass) + mh_chain_rows
  recurse    anchor search in the gaps of more than    mh_recursion_gaps + mh_anchor_gaps,
             200 residues on both sides, new anchors   then the chain again
             chained with the old ones
  refine     the profile pass and the final            mh_refine_pairwise (every DP of a
             refinement pass                           pass in one GPU batch)
  xmfa       the alignment file                        mh_xmfa_write
  lut        buildIndex's LUT from that file           mh_index_* (fixZeroIdx, fillGaps,
                                                       smoothEdges)

The LCB pass of the reference (greedy breakpoint elimination) is not part of
this library: the chain of mh_chain_gap_jobs stands in for it, so the
alignment has one interval over both whole genomes and anchors on the forward
strand only.  Every later stage is the reference's, given those anchors.
"""
import time

import numpy as np

from . import (MATCH_DTYPE, SortedMerList, anchor_gaps_device, chain_gap_jobs_raw, chain_rows, find_matches_raw,
               recursion_gaps, refine_pairwise, write_xmfa)
from . import index as mindex

MIN_RECURSIVE_GAP = 200          # default_min_r_gap_size, LM/Aligner.cpp:179


def align_pair(ctx, d_seqs, seqs, seed, xmfa_path, names, recursive=True, refine=True, lut=True, timings=None,
               refine_stats=None):
    """Align genome 1 (seqs[0]) with genome 2 (seqs[1]); d_seqs are the same
    genomes in device memory (ASCII).  Writes xmfa_path and returns
    (row_1, row_2, lut or None, info), the rows as uint8 arrays; timings
    (dict) accumulates seconds per stage, refine_stats (RefineStats) receives
    the DP cells of the passes."""
    t = time.perf_counter
    tm = timings if timings is not None else {}

    def mark(name, t0):
        tm[name] = tm.get(name, 0.0) + (t() - t0)
        return t()

    (d0, d1), (s0, s1) = d_seqs, seqs
    L0, L1 = len(s0), len(s1)
    t0 = t()
    sa = SortedMerList.from_device(ctx, d0, L0, seed)
    sb = SortedMerList.from_device(ctx, d1, L1, seed)
    t0 = mark("sml", t0)
    mt, stats = find_matches_raw(ctx, [sa, sb])
    sa.free()
    sb.free()
    t0 = mark("match", t0)
    t1 = t()
    chain, _ = chain_gap_jobs_raw(mt, L0, L1, 0)
    t1 = mark("chain.dp", t1)
    ra, rb = chain_rows(mt, chain, s0, s1, as_array=True)
    mark("chain.rows", t1)
    t0 = mark("chain", t0)
    n_rec = 0
    if recursive:
        t1 = t()
        gaps = recursion_gaps(ra, rb, 1, 1, MIN_RECURSIVE_GAP)
        t1 = mark("recurse.gaps", t1)
        if len(gaps):
            gs, gl, _ = anchor_gaps_device(ctx, d0, L0, d1, L1, gaps)
            t1 = mark("recurse.search", t1)
            n_rec = len(gl)
            if n_rec:
                extra = np.zeros(n_rec, dtype=MATCH_DTYPE)
                extra["start"][:, :2] = gs
                extra["length"] = gl
                mt = np.concatenate([mt, extra])
                chain, _ = chain_gap_jobs_raw(mt, L0, L1, 0)
                ra, rb = chain_rows(mt, chain, s0, s1, as_array=True)
                t1 = mark("recurse.chain", t1)
        t0 = mark("recurse", t0)
    if refine:
        (ra, rb), = refine_pairwise(ctx, [(ra, rb)], stats=refine_stats, as_array=True)
        t0 = mark("refine", t0)
    write_xmfa(xmfa_path, [[(1, L0, ra), (1, L1, rb)]], names)
    t0 = mark("xmfa", t0)
    table = None
    if lut:
        t1 = t()
        table = mindex.lut_from_xmfa_file(xmfa_path, L0)
        t1 = mark("lut.walk", t1)
        mindex.fixZeroIdx(table, s0, s1)
        mindex.fillGaps(table)
        mindex.smoothEdges(table)
        mark("lut.passes", t1)
        t0 = mark("lut", t0)
    info = {"matches": int(len(mt)) - n_rec, "seed_tuples": int(stats.seed_tuples), "chain_anchors": int(len(chain)),
            "recursive_anchors": n_rec, "columns": len(ra)}
    return ra, rb, table, info
