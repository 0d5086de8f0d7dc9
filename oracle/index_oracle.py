"""CPU oracle of buildIndex's LUT: XMFA parse, column walk and the three
post-passes.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
mh_index_* entry points; nothing in the product calls it.

Restates, with plain Python loops (small inputs only):
  parse_xmfa        the grammar of PY/xmfa.py:15-68 (the reference calls
                    libnano.fileio.xmfa.parseXMFA, an un-vendored third-party
                    module with no pinned version; the local xmfa.py is the
                    readable stand-in): a group is a line-start '>' up to the
                    first '=', with the rest of that line as its notes; in a
                    group, an entry is "> seq:start-end strand notes" (strand
                    one of + | -) and its sequence is everything up to the next
                    '>' with whitespace removed
  lut_from_groups   PY/buildindex.py:110-129: LUT of -1; for each group with
                    sequences 1 and 2, walk the columns, advance each index on
                    a non-gap character and map where the characters are equal
                    (strand ignored; numpy's wrap-around for index -1 kept).
                    Both indices start at start-1 and are advanced before
                    their first use, so the column of 0-based positions
                    (g, r) sets lut[g + 1] = r + 1, as in the reference
  fix_zero_idx      PY/indexutils.pyx:27-43 (IndexError when no mapped entry at
                    or after index 10 exists, as the bounds-checked buffer raises)
  fill_gaps         PY/indexutils.pyx:46-74
  smooth_edges      PY/indexutils.pyx:77-108
  seq_from_fasta    the first record of a FASTA file, whitespace removed
                    (libnano.fileio.getSeqFromFile; PY/fasta.py:31-52 is the
                    local stand-in)

Parity pins (tests/test_index_oracle.py): the reference's own fixture
T/example.xmfa for the grammar (groups, entry fields, notes and rows read
off the file); hand-worked known answers for the three passes; alignments
built from a known edit script for the column walk.  Parity against libnano
itself is unpinned (the module is absent).
"""
import re

_GROUP = re.compile(r"^>[^=]+=.*$", re.M)
_ENTRY = re.compile(r"^>\s*(\d+):(\d+)-(\d+)\s*([+|-])\s*([ \S]+)\s*?$([^=>]+)", re.M)
_WS = re.compile(r"\s+")


def parse_xmfa(text):
    """[(entries, notes)], entries = [(seq_num, start, end, strand, notes, seq)]."""
    groups = []
    for g in _GROUP.finditer(text):
        body, _, notes = g.group(0).partition("=")
        entries = []
        for e in _ENTRY.finditer(body):
            entries.append((int(e.group(1)), int(e.group(2)), int(e.group(3)), e.group(4), e.group(5),
                            _WS.sub("", e.group(6))))
        groups.append((entries, notes))
    return groups


def _first(entries, num):
    for e in entries:
        if e[0] == num:
            return e
    return None


def lut_from_groups(groups, genome_length):
    lut = [-1] * genome_length
    for entries, _ in groups:
        g, r = _first(entries, 1), _first(entries, 2)
        if g is None or r is None:
            continue
        gi, ri = g[1] - 1, r[1] - 1
        gs, rs = g[5], r[5]
        for col, base in enumerate(gs):
            rb = rs[col]                       # IndexError if the reference row is shorter
            if base != "-":
                gi += 1
            if rb != "-":
                ri += 1
            if base == rb and gi < genome_length:
                lut[gi] = ri                   # gi == -1 writes the last entry, as numpy does
    return lut


def fix_zero_idx(lut, genome, ref):
    if lut[0] != -1:
        return
    nxt = 1
    while lut[nxt] == -1 or nxt < 10:          # IndexError past the end
        nxt += 1
    r = lut[nxt] - nxt
    if r > -1 and genome[0] == ref[r]:
        lut[0] = r


def _fill_from(lut, lower, idx, width):
    lm = lut[lower]
    for up in range(idx + 1, len(lut)):
        if up - idx > width:
            return
        if lut[up] - lm == up - lower:
            for k in range(lower, up + 1):
                lut[k] = lm + (k - lower)
            return


def fill_gaps(lut, max_gap_width=300):
    for idx in range(1, len(lut)):
        if lut[idx] == -1 and lut[idx - 1] != -1:
            _fill_from(lut, idx - 1, idx, max_gap_width)


def smooth_edges(lut, smoothing_radius=20):
    for idx in range(1, len(lut)):
        if lut[idx] != lut[idx - 1] + 1 and lut[idx - 1] != -1:
            _fill_from(lut, idx - 1, idx, smoothing_radius)


def seq_from_fasta(text):
    m = re.search(r">(.*)[\n\r]([^>]+)", text)
    return "".join(m.group(2).split()) if m else ""
