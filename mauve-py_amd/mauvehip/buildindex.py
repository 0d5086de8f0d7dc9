"""buildIndex with the alignment run by this library, a drop-in for
mauve/buildindex.py (PY/buildindex.py:56-138): the same functions, arguments
and defaults.

runMauve (PY/buildindex.py:56-80) runs progressiveMauveStatic on FASTA files;
here it runs the pairwise path of mauvehip.pipeline.align_pair on the GPU
(sorted mer lists, matching, anchoring, recursive anchor search, the profile
and refinement passes) and writes the XMFA file that flags['--output'] names.
The seed is progressiveMauve's default for two genomes
(MA/progressiveMauve.cpp:433-455: LoadSMLs with the coding seed; the weight
from GetDefaultMerSize, LM/MatchList.h:380-385, unless '--seed-weight' gives
it).  The LCB pass of the reference is not part of this library (the chain of
mh_chain_gap_jobs stands in for it), so the XMFA is this library's alignment,
not progressiveMauve's.

buildIndex (PY/buildindex.py:90-138) then reads the LUT from that file and
runs fixZeroIdx, fillGaps and smoothEdges as the reference does
(mauvehip.index).
"""
import os
import tempfile

import numpy as np

from . import CODING_SEED, Context, get_default_seed_weight, get_seed
from . import index as mindex
from . import pipeline

# flags runMauve accepts: the output path, the seed weight, and the scratch
# paths the reference's buildIndex passes (used there by progressiveMauve for
# its sorted mer list files; this library keeps those in device memory)
_FLAGS = {"--output", "--seed-weight", "--scratch-path-1", "--scratch-path-2"}


def getGenomeLength(genome_fp):
    """The genome length of a FASTA file holding one record
    (PY/buildindex.py:83-87)."""
    return mindex.genome_length(genome_fp)


def runMauve(fasta_files, flags={}):
    """Align the genomes of two FASTA files (one record each) and write the
    XMFA alignment to flags['--output'] (PY/buildindex.py:56-80).  The first
    file is sequence 1 of the XMFA file."""
    if len(fasta_files) != 2:
        raise ValueError("runMauve aligns two genomes in this build")
    unknown = set(flags) - _FLAGS
    if unknown:
        raise ValueError("flags not supported by this build: " + ", ".join(sorted(unknown)))
    if "--output" not in flags:
        raise ValueError("runMauve needs flags['--output'], the XMFA file to write")
    paths = [os.path.abspath(fp) for fp in fasta_files]
    seqs = [mindex.seq_from_file(fp).encode("latin-1") for fp in paths]
    if not all(seqs):
        raise ValueError("a FASTA file holds no sequence")
    weight = int(flags.get("--seed-weight", 0)) or get_default_seed_weight(sum(len(s) for s in seqs) // 2)
    seed = get_seed(weight, CODING_SEED)
    ctx = Context(0)
    d = []
    try:
        for s in seqs:
            p = ctx.alloc(len(s) + 16)
            d.append(p)
            ctx.h2d(p, np.frombuffer(s, dtype=np.uint8))
        pipeline.align_pair(ctx, d, seqs, seed, str(flags["--output"]), paths, lut=False)
    finally:
        for p in d:
            ctx.free(p)
        ctx.close()


def buildIndex(genome_fp, ref_genome_fp, genome_seq=None, ref_genome_seq=None,
               fill_gaps=True, max_gap_width=300, smooth_edges=True,
               smoothing_radius=20):
    """The index mapping positions of the genome in genome_fp to the genome
    in ref_genome_fp (PY/buildindex.py:90-138): -1 where unmapped; gaps up to
    max_gap_width filled when fill_gaps, edges smoothed when smooth_edges."""
    genome_fp = os.path.abspath(genome_fp)
    ref_genome_fp = os.path.abspath(ref_genome_fp)
    with tempfile.TemporaryDirectory() as temp_dir:
        output_fp = os.path.join(temp_dir, "mauveout.xmfa")
        runMauve([genome_fp, ref_genome_fp], flags={"--output": output_fp})
        return mindex.build_index_from_xmfa(output_fp, genome_fp, ref_genome_fp, genome_seq=genome_seq,
                                            ref_genome_seq=ref_genome_seq, fill_gaps=fill_gaps,
                                            max_gap_width=max_gap_width, smooth_edges=smooth_edges,
                                            smoothing_radius=smoothing_radius)
