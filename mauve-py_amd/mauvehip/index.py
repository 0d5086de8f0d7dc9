"""Genome index LUT: the post-processing of mauve.buildIndex
(PY/buildindex.py:90-138, PY/indexutils.pyx:27-108) over an XMFA file.

``build_index_from_xmfa`` takes the XMFA alignment of two genomes (sequence 1
the engineered genome, sequence 2 the reference) and returns the int32 LUT
mapping every position of genome 1 to genome 2 (-1 unmapped), with the same
options and passes as buildIndex.  The column walk and the passes run in the
library (mh_index_*), in place on a contiguous int32 array.

Sequences are read as the first record of a FASTA file, whitespace removed
(what the reference's libnano getSeqFromFile returns for the single-record
files buildIndex is documented for).
"""
import ctypes as C
import os
import re

import numpy as np

from . import MhError, _check, load_library

IDX_ARRAY_DTYPE = np.int32
MASK_ARRAY_DTYPE = np.uint8

_FASTA_RECORD = re.compile(rb">(.*)[\n\r]([^>]+)")


def _bytes(s):
    if isinstance(s, str):
        return s.encode("latin-1")
    return bytes(s)


def _lut_arg(idx_lut):
    if not (isinstance(idx_lut, np.ndarray) and idx_lut.dtype == IDX_ARRAY_DTYPE and idx_lut.ndim == 1
            and idx_lut.flags.c_contiguous and idx_lut.flags.writeable):
        raise ValueError("idx_lut must be a writeable contiguous 1-d int32 array")
    return idx_lut.ctypes.data, len(idx_lut)


def seq_from_file(path):
    """The first FASTA record's sequence of ``path``, whitespace removed."""
    with open(path, "rb") as f:
        m = _FASTA_RECORD.search(f.read())
    return b"".join(m.group(2).split()).decode("latin-1") if m else ""


def genome_length(path):
    return len(seq_from_file(path))


def lut_from_xmfa(xmfa, genome_length):
    """The LUT of the XMFA column walk (PY/buildindex.py:110-129), before any
    pass.  ``xmfa`` is the file's text (str/bytes)."""
    L = load_library()
    text = _bytes(xmfa)
    lut = np.empty(genome_length, dtype=IDX_ARRAY_DTYPE)
    ng = C.c_uint64()
    _check(L.mh_index_from_xmfa(text, len(text), genome_length, lut.ctypes.data, C.byref(ng)))
    return lut


def lut_from_xmfa_file(path, genome_length):
    """lut_from_xmfa of the XMFA file at path (read by the library)."""
    L = load_library()
    lut = np.empty(genome_length, dtype=IDX_ARRAY_DTYPE)
    ng = C.c_uint64()
    _check(L.mh_index_from_xmfa_file(os.fsencode(path), genome_length, lut.ctypes.data, C.byref(ng)))
    return lut


def fixZeroIdx(idx_lut, genome, ref_genome):
    """PY/indexutils.pyx:27-43, in place.  IndexError as in the reference
    when no mapped entry exists at or after index 10."""
    L = load_library()
    p, n = _lut_arg(idx_lut)
    g, r = _bytes(genome), _bytes(ref_genome)
    if n == 0:
        raise IndexError("Out of bounds on buffer access (axis 0)")
    try:
        _check(L.mh_index_fix_zero(p, n, g, len(g), r, len(r)))
    except MhError as e:
        raise IndexError(str(e)) from None


def fillGaps(idx_lut, max_gap_width=300):
    """PY/indexutils.pyx:46-74, in place."""
    L = load_library()
    p, n = _lut_arg(idx_lut)
    _check(L.mh_index_fill_gaps(p, n, int(max_gap_width)))


def smoothEdges(idx_lut, smoothing_radius=20):
    """PY/indexutils.pyx:77-108, in place."""
    L = load_library()
    p, n = _lut_arg(idx_lut)
    _check(L.mh_index_smooth_edges(p, n, int(smoothing_radius)))


fix_zero_idx, fill_gaps, smooth_edges = fixZeroIdx, fillGaps, smoothEdges


def build_index_from_xmfa(xmfa_fp, genome_fp, ref_genome_fp, genome_seq=None, ref_genome_seq=None, fill_gaps=True,
                          max_gap_width=300, smooth_edges=True, smoothing_radius=20):
    """buildIndex (PY/buildindex.py:90-138) from the XMFA alignment of
    genome_fp (sequence 1) against ref_genome_fp (sequence 2)."""
    lut = lut_from_xmfa_file(xmfa_fp, genome_length(genome_fp))
    if fill_gaps:
        genome_seq = genome_seq or seq_from_file(genome_fp)
        ref_genome_seq = ref_genome_seq or seq_from_file(ref_genome_fp)
        fixZeroIdx(lut, genome_seq, ref_genome_seq)
        fillGaps(lut, max_gap_width=max_gap_width)
    if smooth_edges:
        smoothEdges(lut, smoothing_radius=smoothing_radius)
    return lut


def findEdges(idx_lut):
    """PY/indexutils.pyx:142-156 as a uint8 mask (the reference packs it into
    a bitarray): 1 at index 0 and wherever the mapping is not consecutive."""
    edge = np.zeros(len(idx_lut), dtype=MASK_ARRAY_DTYPE)
    if len(idx_lut):
        edge[1:][(idx_lut[1:] - idx_lut[:-1]) != 1] = 1
        edge[0] = 1
    return edge


def findMismatches(idx_lut, genome, ref_genome):
    """PY/indexutils.pyx:159-175 as a uint8 mask: mapped positions of genome
    whose base differs from the mapped reference base."""
    g = np.frombuffer(_bytes(genome), dtype=np.uint8)
    r = np.frombuffer(_bytes(ref_genome), dtype=np.uint8)
    out = np.zeros(len(idx_lut), dtype=MASK_ARRAY_DTYPE)
    idx = idx_lut[:len(g)]
    m = idx != -1
    pos = np.nonzero(m)[0]
    out[pos] = g[pos] != r[idx[pos]]
    return out


def findDuplicateMappings(idx_lut):
    """PY/indexutils.pyx:178-186: reference positions mapped more than once,
    less one (the clipped -1 entries share position 0)."""
    counts = np.bincount(np.clip(idx_lut, 0, 2 ** 31 - 1))
    return int(np.sum(counts > 1)) - 1
