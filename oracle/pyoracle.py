"""ctypes view of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product package never
imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OR_MAX_SEQ = 16
CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}


class OrMatch(C.Structure):
    _fields_ = [("start", C.c_int64 * OR_MAX_SEQ), ("length", C.c_uint64)]


class OrStats(C.Structure):
    _fields_ = [("hash_calls", C.c_int64), ("collisions", C.c_int64),
                ("restarts", C.c_int64), ("matches", C.c_int64)]


def build():
    """Compile liboracle.so in place (gcc, a few seconds)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_get_seed.restype = C.c_int64
        L.or_get_seed.argtypes = [C.c_int, C.c_int]
        L.or_get_seed_length.argtypes = [C.c_int64]
        L.or_get_seed_weight.argtypes = [C.c_int64]
        L.or_default_seed_weight.restype = C.c_uint32
        L.or_default_seed_weight.argtypes = [C.c_uint64]
        L.or_packed_words.restype = C.c_uint64
        L.or_packed_words.argtypes = [C.c_uint64]
        L.or_translate32.argtypes = [C.c_char_p, C.c_uint64, C.c_void_p]
        L.or_get_dna_seed_mer.restype = C.c_uint64
        L.or_get_dna_seed_mer.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_int]
        L.or_revcomp_mer.restype = C.c_uint64
        L.or_revcomp_mer.argtypes = [C.c_uint64, C.c_int]
        L.or_fill_dna_sml.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p]
        L.or_fill_dna_seed_sml.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_void_p]
        L.or_sml_build.restype = C.c_int64
        L.or_sml_build.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_score6.argtypes = [C.c_void_p]
        L.or_find_mer_keys.restype = C.c_int
        L.or_find_mer_keys.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_uint64, C.POINTER(C.c_uint64)]
        L.or_find_matches_pairwise.restype = C.c_int64
        L.or_find_matches_pairwise.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64),
                                               C.c_uint64, C.POINTER(OrMatch), C.c_int64, C.POINTER(OrStats)]
        L.or_find_matches_unique.restype = C.c_int64
        L.or_find_matches_unique.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64),
                                             C.c_uint64, C.POINTER(OrMatch), C.c_int64, C.POINTER(OrStats)]
        L.or_seed_tuples_unique.restype = C.c_int64
        L.or_seed_tuples_unique.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64),
                                            C.c_uint64, C.POINTER(OrMatch), C.c_void_p, C.c_int64]
        L.or_seed_tuples_pairwise.restype = C.c_int64
        L.or_seed_tuples_pairwise.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64),
                                              C.c_uint64, C.POINTER(OrMatch), C.c_void_p, C.c_int64]
        L.or_extend_seed.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64),
                                     C.c_uint64, C.POINTER(OrMatch)]
        L.or_eliminate_overlaps_pairwise.restype = C.c_int64
        L.or_eliminate_overlaps_pairwise.argtypes = [C.POINTER(OrMatch), C.c_int64, C.c_uint64]
        L.or_nwsmall.restype = C.c_int64
        L.or_nwsmall.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_char_p,
                                 C.POINTER(C.c_float)]
        L.or_nwsmall_region.restype = C.c_int64
        L.or_nwsmall_region.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_char_p, C.c_void_p, C.c_void_p]
        L.or_nwsmall_window.restype = C.c_int64
        L.or_nwsmall_window.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_char_p, C.c_void_p, C.c_void_p]
        L.or_window_base.restype = C.c_int64
        L.or_window_base.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_path_dev.restype = C.c_uint32
        L.or_path_dev.argtypes = [C.c_uint32, C.c_uint32, C.c_char_p, C.c_int64]
        L.or_path_score.restype = C.c_float
        L.or_path_score.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_char_p, C.c_int64]
        L.or_refine_gapped.restype = C.c_int64
        L.or_refine_gapped.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                       C.POINTER(C.c_uint64)]
        L.or_refine_pass.restype = C.c_int64
        L.or_refine_pass.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_int64]
        L.or_column_scores.restype = C.c_int64
        L.or_column_scores.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_void_p]
        L.or_anchor_cols.restype = C.c_int64
        L.or_anchor_cols.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_void_p, C.c_int64]
        L.or_xmfa_lut.restype = C.c_int64
        L.or_xmfa_lut.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int64,
                                  C.c_int64, C.c_int64, C.c_char_p, C.c_char_p, C.c_void_p, C.c_int64]
        L.or_recursion_gaps.restype = C.c_int64
        L.or_recursion_gaps.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                        C.c_void_p, C.c_int64]
        L.or_nwsmall_batch.restype = C.c_int64
        L.or_nwsmall_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int64, C.c_int64, C.c_void_p]
        _LIB = L
    return _LIB


def _b(s):
    return s if isinstance(s, bytes) else s.encode()


def get_seed(weight, rank=0):
    return lib().or_get_seed(weight, rank)


def seed_length(seed):
    return lib().or_get_seed_length(seed)


def seed_weight(seed):
    return lib().or_get_seed_weight(seed)


def default_seed_weight(avg_len):
    return lib().or_default_seed_weight(avg_len)


def revcomp_mer(mer, length):
    return lib().or_revcomp_mer(mer, length)


def packed(seq):
    seq = _b(seq)
    nw = lib().or_packed_words(len(seq))
    out = np.zeros(nw, dtype=np.uint32)
    rc = lib().or_translate32(seq, len(seq), out.ctypes.data)
    if rc != 0:
        raise ValueError("gap character in genome sequence")
    return out


def keys_by_position(seq, seed):
    """GetDnaSeedMer at every SML position, in position order."""
    seq = _b(seq)
    words = packed(seq)
    sl, sw = seed_length(seed), seed_weight(seed)
    n = len(seq) - sl + 1
    out = np.zeros(max(n, 0), dtype=np.uint64)
    if n > 0:
        lib().or_fill_dna_seed_sml(words.ctypes.data, n, seed, sl, sw, out.ctypes.data)
    return out


def fill_dna_sml(seq, seed):
    """The incremental solid-seed fill (FillDnaSML) in position order."""
    seq = _b(seq)
    words = packed(seq)
    sl = seed_length(seed)
    n = len(seq) - sl + 1
    out = np.zeros(max(n, 0), dtype=np.uint64)
    if n > 0:
        lib().or_fill_dna_sml(words.ctypes.data, len(seq), sl, out.ctypes.data)
    return out


def sml_build(seq, seed):
    """Returns (packed words, sorted keys, sorted positions)."""
    seq = _b(seq)
    L = len(seq)
    words = np.zeros(lib().or_packed_words(L), dtype=np.uint32)
    n = max(L - seed_length(seed) + 1, 0)
    keys = np.zeros(max(n, 1), dtype=np.uint64)
    pos = np.zeros(max(n, 1), dtype=np.uint32)
    r = lib().or_sml_build(seq, L, seed, words.ctypes.data, keys.ctypes.data, pos.ctypes.data)
    if r < 0:
        raise ValueError("SML build failed")
    return words, keys[:r], pos[:r]


def find_mer(keys, seq_len, seed, q):
    """SortedMerList::FindMer on the sorted keys of a sequence of seq_len
    bases: (1 found / 0 not found / -1 too short, result or None if unset)."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    r = C.c_uint64(0xFFFFFFFFFFFFFFFF)
    f = lib().or_find_mer_keys(k.ctypes.data if len(k) else None, seq_len, seed_length(seed), int(q), C.byref(r))
    return f, (None if f < 0 else r.value)


def _seq_args(seqs):
    bs = [_b(s) for s in seqs]
    arr = (C.c_char_p * len(bs))(*bs)
    lens = (C.c_uint64 * len(bs))(*[len(s) for s in bs])
    return bs, arr, lens


def find_matches(seqs, seed, max_out=None):
    """Pairwise match list in reference order: (starts[k] int64 array, lengths)."""
    k = len(seqs)
    bs, arr, lens = _seq_args(seqs)
    cap = max_out or (max(len(s) for s in bs) + 16)
    out = (OrMatch * cap)()
    st = OrStats()
    n = lib().or_find_matches_pairwise(k, arr, lens, seed, out, cap, C.byref(st))
    if n < 0:
        raise ValueError("match finding failed")
    starts = np.array([[out[i].start[s] for s in range(k)] for i in range(n)], dtype=np.int64).reshape(n, k)
    lengths = np.array([out[i].length for i in range(n)], dtype=np.uint64)
    return starts, lengths, st


def find_matches_unique(seqs, seed, max_out=None):
    """UniqueMatchFinder's match list in reference order: (starts (n, k) int64,
    lengths, stats)."""
    k = len(seqs)
    bs, arr, lens = _seq_args(seqs)
    cap = max_out or (max(len(s) for s in bs) + 16)
    out = (OrMatch * cap)()
    st = OrStats()
    n = lib().or_find_matches_unique(k, arr, lens, seed, out, cap, C.byref(st))
    if n < 0:
        raise ValueError("match finding failed")
    starts = np.array([[out[i].start[s] for s in range(k)] for i in range(n)], dtype=np.int64).reshape(n, k)
    lengths = np.array([out[i].length for i in range(n)], dtype=np.uint64)
    return starts, lengths, st


def seed_tuples_unique(seqs, seed):
    """UniqueMatchFinder's HashMatch calls in order: (starts (n, k), offsets)."""
    k = len(seqs)
    bs, arr, lens = _seq_args(seqs)
    n = lib().or_seed_tuples_unique(k, arr, lens, seed, None, None, 0)
    out = (OrMatch * max(n, 1))()
    off = np.zeros(max(n, 1), dtype=np.int64)
    n2 = lib().or_seed_tuples_unique(k, arr, lens, seed, out, off.ctypes.data, n)
    starts = np.array([[out[i].start[s] for s in range(k)] for i in range(n2)], dtype=np.int64).reshape(n2, k)
    return starts, off[:n2]


def eliminate_overlaps(starts, lengths, min_length=0):
    """EliminateOverlaps_v2 then LengthFilter for two-genome matches:
    (starts (n, 2), lengths) in the resulting order."""
    n = len(lengths)
    arr = (OrMatch * max(n, 1))()
    for i in range(n):
        arr[i].start[0] = int(starts[i][0])
        arr[i].start[1] = int(starts[i][1])
        arr[i].length = int(lengths[i])
    m = lib().or_eliminate_overlaps_pairwise(arr, n, min_length)
    st = np.array([[arr[i].start[0], arr[i].start[1]] for i in range(m)], dtype=np.int64).reshape(m, 2)
    return st, np.array([arr[i].length for i in range(m)], dtype=np.uint64)


MIN_ANCHOR_LENGTH = 9                                 # LM/Aligner.h:266
MIN_DNA_SEED_WEIGHT = 5                               # LM/SeedMasks.h:375


def anchor_gap(seq_a, seq_b, lend, rend):
    """ProgressiveAligner::pairwiseAnchorSearch (LM/ProgressiveAligner.cpp:
    590-679) for one gap: lend/rend are getInterveningCoordinates' gap_lend /
    gap_rend in each genome (1-based); the searched sequence starts at lend and
    is rend - lend - 1 long (when positive).  Returns (starts (n, 2) in genome
    coordinates, lengths)."""
    seqs = []
    for g, s in enumerate((_b(seq_a), _b(seq_b))):
        diff = rend[g] - lend[g]
        diff = diff - 1 if diff > 0 else 0
        if diff > 0 and lend[g] + diff - 1 <= len(s):
            seqs.append(s[lend[g] - 1:lend[g] - 1 + diff])
        else:
            seqs.append(b"")
    avg = (len(seqs[0]) + len(seqs[1])) // 2
    w = default_seed_weight(avg)
    empty = (np.zeros((0, 2), dtype=np.int64), np.zeros(0, dtype=np.uint64))
    if w < MIN_DNA_SEED_WEIGHT:
        return empty
    seed = get_seed(w, 0)
    if min(len(s) for s in seqs) < seed_length(seed):
        return empty                                  # a list without seeds pairs with nothing
    st, ln, _ = find_matches(seqs, seed)
    if len(ln) == 0:
        return empty
    st, ln = eliminate_overlaps(st, ln, MIN_ANCHOR_LENGTH + 3)
    for g in range(2):
        # SetLeftEnd(LeftEnd + gap_lend - 1), keeping the strand
        sign = np.where(st[:, g] < 0, -1, 1)
        st[:, g] = sign * (np.abs(st[:, g]) + lend[g] - 1)
    return st, ln


def recursion_gaps(row_a, row_b, left_a=1, left_b=1, min_gap=200):
    """The gaps recurseOnPairs searches (LM/ProgressiveAligner.cpp:838-905)
    for an interval forward in both genomes: [((lend_a, lend_b), (rend_a,
    rend_b))] as getInterveningCoordinates gives them for the zero-length
    anchors placed at the aligned columns around each gap."""
    right_a = left_a + sum(c != "-" for c in row_a) - 1
    right_b = left_b + sum(c != "-" for c in row_b) - 1
    out = []
    ci = cj = pi = pj = 0
    in_gap = False
    n = len(row_a)
    for col in range(n + 1):
        if col == n or (row_a[col] != "-" and row_b[col] != "-"):
            if in_gap and ci - pi > min_gap and cj - pj > min_gap:
                lend = (left_a + pi, left_b + pj)        # End()+1 of a zero-length match = its left end
                if cj != right_b and ci != right_a:
                    rend = (left_a + ci, left_b + cj)
                else:
                    rend = (right_a + 1, right_b + 1)
                out.append((lend, rend))
            pi, pj = ci, cj
            in_gap = False
        else:
            in_gap = True
        if col < n:
            ci += row_a[col] != "-"
            cj += row_b[col] != "-"
    return out


def seed_tuples(seqs, seed):
    k = len(seqs)
    bs, arr, lens = _seq_args(seqs)
    n = lib().or_seed_tuples_pairwise(k, arr, lens, seed, None, None, 0)
    out = (OrMatch * max(n, 1))()
    off = np.zeros(max(n, 1), dtype=np.int64)
    n2 = lib().or_seed_tuples_pairwise(k, arr, lens, seed, out, off.ctypes.data, n)
    starts = np.array([[out[i].start[s] for s in range(k)] for i in range(n2)], dtype=np.int64).reshape(n2, k)
    return starts, off[:n2]


def extend_seed(seqs, seed, starts, length):
    k = len(seqs)
    bs, arr, lens = _seq_args(seqs)
    m = OrMatch()
    for s in range(k):
        m.start[s] = int(starts[s])
    m.length = length
    lib().or_extend_seed(k, arr, lens, seed, C.byref(m))
    return [m.start[s] for s in range(k)], m.length


_CLASS = np.full(256, 5, dtype=np.uint8)
for _i, _c in enumerate(b"ACGTX"):
    _CLASS[_c] = _i
    _CLASS[ord(chr(_c).lower())] = _i
_CLASS[ord("U")] = _CLASS[ord("u")] = 3


def codes(seq):
    """DP operand classes: A,C,G,T (U as T) -> 0..3, X -> 4, any other letter
    -> 5 (libMUSCLE's wildcard counts, see or_score6; U's class is the one
    under which libMUSCLE's paths in tests/golden/dp_wild_golden.txt come out)."""
    return _CLASS[np.frombuffer(_b(seq), dtype=np.uint8)]


def score6():
    out = np.zeros(36, dtype=np.float32)
    lib().or_score6(out.ctypes.data)
    return out.reshape(6, 6)


def nwsmall(a, b):
    """Path string ('M'/'D'/'I') and terminal score for one job (A, B strings or code arrays)."""
    A = codes(a) if isinstance(a, (str, bytes)) else np.ascontiguousarray(a, dtype=np.uint8)
    B = codes(b) if isinstance(b, (str, bytes)) else np.ascontiguousarray(b, dtype=np.uint8)
    buf = C.create_string_buffer(len(A) + len(B) + 1)
    sc = C.c_float()
    n = lib().or_nwsmall(A.ctypes.data, len(A), B.ctypes.data, len(B), buf, C.byref(sc))
    if n < 0:
        raise ValueError("NWSmall: length=0")
    return buf.raw[:n].decode(), sc.value


def nwsmall_region(a, b, band, stripe_w):
    """Banded NWSmall over the stripe region (band 0: whole matrix):
    (path, int score, deviation)."""
    A = codes(a) if isinstance(a, (str, bytes)) else np.ascontiguousarray(a, dtype=np.uint8)
    B = codes(b) if isinstance(b, (str, bytes)) else np.ascontiguousarray(b, dtype=np.uint8)
    buf = C.create_string_buffer(len(A) + len(B) + 1)
    sc = C.c_int32()
    dv = C.c_uint32()
    n = lib().or_nwsmall_region(A.ctypes.data, len(A), B.ctypes.data, len(B), int(band), int(stripe_w), buf,
                                C.byref(sc), C.byref(dv))
    if n == -1:
        raise ValueError("NWSmall: length=0")
    if n < 0:
        return None, None, None
    return buf.raw[:n].decode(), int(sc.value), int(dv.value)


def nwsmall_window(a, b, band, width):
    """Banded NWSmall over the sliding window of `width` columns:
    (path, int score, deviation)."""
    A = codes(a) if isinstance(a, (str, bytes)) else np.ascontiguousarray(a, dtype=np.uint8)
    B = codes(b) if isinstance(b, (str, bytes)) else np.ascontiguousarray(b, dtype=np.uint8)
    buf = C.create_string_buffer(len(A) + len(B) + 1)
    sc = C.c_int32()
    dv = C.c_uint32()
    n = lib().or_nwsmall_window(A.ctypes.data, len(A), B.ctypes.data, len(B), int(band), int(width), buf,
                                C.byref(sc), C.byref(dv))
    if n == -1:
        raise ValueError("NWSmall: length=0")
    if n < 0:
        return None, None, None
    return buf.raw[:n].decode(), int(sc.value), int(dv.value)


def window_base(la, lb, band, width, i):
    return int(lib().or_window_base(int(la), int(lb), int(band), int(width), int(i)))


def path_dev(la, lb, path):
    """ceil(max |i*LB - j*LA| / LA) over the path's positions."""
    return int(lib().or_path_dev(int(la), int(lb), _b(path), len(path)))


def path_score(a, b, path):
    A = codes(a) if isinstance(a, (str, bytes)) else np.ascontiguousarray(a, dtype=np.uint8)
    B = codes(b) if isinstance(b, (str, bytes)) else np.ascontiguousarray(b, dtype=np.uint8)
    return lib().or_path_score(A.ctypes.data, len(A), B.ctypes.data, len(B), _b(path), len(path))


def rows_from_path(a, b, path):
    """Aligned rows (as AlignTwoMSAsGivenPath lays them out) from a path."""
    a = a.decode() if isinstance(a, bytes) else a
    b = b.decode() if isinstance(b, bytes) else b
    ra, rb, i, j = [], [], 0, 0
    for c in path:
        if c == "M":
            ra.append(a[i]); rb.append(b[j]); i += 1; j += 1
        elif c == "D":
            ra.append(a[i]); rb.append("-"); i += 1
        else:
            ra.append("-"); rb.append(b[j]); j += 1
    return "".join(ra), "".join(rb)


# ---- compiled host stages of the whole path (oracle/cpu_path.cpp) ----

def c_refine_gapped(r1, r2):
    """Profile pass then refinement pass of two aligned rows, compiled
    (refine_oracle.gapped_alignment restated in C++); returns (row1, row2,
    DP cells)."""
    r1, r2 = _b(r1), _b(r2)
    cap = 2 * len(r1) + 64
    o1, o2 = C.create_string_buffer(cap), C.create_string_buffer(cap)
    cells = C.c_uint64()
    n = lib().or_refine_gapped(r1, r2, len(r1), o1, o2, cap, C.byref(cells))
    if n < 0:
        raise ValueError(f"or_refine_gapped: {n}")
    return o1.raw[:n].decode(), o2.raw[:n].decode(), cells.value


def c_refine_pass(r1, r2, profile):
    r1, r2 = _b(r1), _b(r2)
    cap = 2 * len(r1) + 64
    o1, o2 = C.create_string_buffer(cap), C.create_string_buffer(cap)
    n = lib().or_refine_pass(r1, r2, len(r1), 1 if profile else 0, o1, o2, cap)
    if n < 0:
        raise ValueError(f"or_refine_pass: {n}")
    return o1.raw[:n].decode(), o2.raw[:n].decode()


def c_column_scores(r1, r2):
    out = np.zeros(len(r1), dtype=np.float32)
    lib().or_column_scores(_b(r1), _b(r2), len(r1), out.ctypes.data)
    return out


def c_anchor_cols(r1, r2):
    out = np.zeros(len(r1) + 1, dtype=np.int64)
    n = lib().or_anchor_cols(_b(r1), _b(r2), len(r1), out.ctypes.data, len(out))
    return [int(x) for x in out[:n]]


def c_xmfa_lut(path, names, r1, r2, len1, len2, seq1, seq2):
    """XMFA of one interval written to path, read back, LUT over genome 1
    with the three post-passes; returns (lut int32, xmfa bytes)."""
    lut = np.zeros(len1, dtype=np.int32)
    r1, r2 = _b(r1), _b(r2)
    n = lib().or_xmfa_lut(_b(path), _b(names[0]), _b(names[1]), r1, r2, len(r1), len1, len2, _b(seq1), _b(seq2),
                          lut.ctypes.data, len1)
    if n < 0:
        raise ValueError(f"or_xmfa_lut: {n}")
    return lut, n


def c_nwsmall_batch(A, B, a_off, a_len, b_off, b_len, lo, hi, scores=None):
    """NWSmall over jobs [lo, hi) of a batch on one thread; returns the cells."""
    n = lib().or_nwsmall_batch(A.ctypes.data, B.ctypes.data, a_off.ctypes.data, a_len.ctypes.data,
                               b_off.ctypes.data, b_len.ctypes.data, lo, hi,
                               scores.ctypes.data if scores is not None else None)
    if n < 0:
        raise ValueError("NWSmall: length=0")
    return n


def c_recursion_gaps(row_a, row_b, left_a=1, left_b=1, min_gap=200):
    """recursion_gaps compiled: the same list of ((lend_a, lend_b), (rend_a, rend_b))."""
    ra, rb = _b(row_a), _b(row_b)
    n = lib().or_recursion_gaps(ra, rb, len(ra), left_a, left_b, min_gap, None, 0)
    out = np.zeros((max(n, 1), 4), dtype=np.int64)
    lib().or_recursion_gaps(ra, rb, len(ra), left_a, left_b, min_gap, out.ctypes.data, n)
    return [((int(o[0]), int(o[1])), (int(o[2]), int(o[3]))) for o in out[:n]]
