"""Synthetic genomes and DP job sets for the benchmark and parity workloads.

All generators are deterministic functions of a numpy PCG64 seed.

* ``pair_iid``        -- 200 kb i.i.d. reference + mutant (1% SNPs, 0.1%/bp
                         indels of 1-20 bp, one 5 kb inversion), seeds 101/102.
* ``pair_recoded``    -- 4.6 Mb, 50% GC reference with ORFs on both strands
                         covering 88%; derivative with 5% of codons replaced by
                         a different synonymous codon, 0.2% SNPs, 20 indels of
                         1-200 bp and 3 inversions of 10-50 kb, seeds 201/202.
* ``dp_jobs``         -- inter-anchor DP pairs: len(A) = floor(10*2000^u),
                         B derived at divergence U(0.02, 0.30) with
                         substitutions:indels 9:1 and Geom(0.3) indel lengths.
"""
import numpy as np

BASES = np.frombuffer(b"ACGT", dtype=np.uint8)
_COMP = np.zeros(256, dtype=np.uint8)
for _a, _b in zip(b"ACGTN", b"TGCAN"):
    _COMP[_a] = _b

_CODON_AA = {}
_AA = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG"
for _i, _c1 in enumerate("TCAG"):
    for _j, _c2 in enumerate("TCAG"):
        for _k, _c3 in enumerate("TCAG"):
            _CODON_AA[_c1 + _c2 + _c3] = _AA[16 * _i + 4 * _j + _k]
_SYN = {}
for _cod, _aa in _CODON_AA.items():
    _SYN.setdefault(_aa, []).append(_cod)
SENSE = [c for c, a in _CODON_AA.items() if a != "*"]
STOPS = ["TAA", "TAG", "TGA"]


def revcomp(seq):
    """Reverse complement of an ASCII uint8 array."""
    return _COMP[seq[::-1]]


def random_bases(rng, n, gc=0.5):
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return BASES[rng.choice(4, size=n, p=p)]


def _mutate_snps(rng, seq, rate):
    n = len(seq)
    hit = np.nonzero(rng.random(n) < rate)[0]
    if len(hit):
        cur = np.searchsorted(BASES, seq[hit])
        cur = np.where(cur < 4, cur, 0)
        shift = rng.integers(1, 4, size=len(hit))
        seq = seq.copy()
        seq[hit] = BASES[(cur + shift) % 4]
    return seq


def _apply_indels(rng, seq, starts, lens, ins_flags):
    """Apply indels at sorted positions (left to right, in original coordinates)."""
    out, prev = [], 0
    for p, ln, ins in zip(starts, lens, ins_flags):
        if p < prev:
            continue
        if ins:
            out.append(seq[prev:p + 1])
            out.append(BASES[rng.integers(0, 4, size=ln)])
            prev = p + 1
        else:
            out.append(seq[prev:p])
            prev = min(p + ln, len(seq))
    out.append(seq[prev:])
    return np.concatenate(out)


def _invert(seq, start, length):
    seq = seq.copy()
    seq[start:start + length] = revcomp(seq[start:start + length])
    return seq


def pair_iid(length=200_000, seed_ref=101, seed_mut=102):
    """Reference + mutant for the 200 kb i.i.d. case (ASCII uint8 arrays)."""
    r = np.random.Generator(np.random.PCG64(seed_ref))
    ref = BASES[r.integers(0, 4, size=length)]
    m = np.random.Generator(np.random.PCG64(seed_mut))
    mut = _mutate_snps(m, ref, 0.01)
    ev = np.nonzero(m.random(len(mut)) < 0.001)[0]
    lens = m.integers(1, 21, size=len(ev))
    ins = m.random(len(ev)) < 0.5
    mut = _apply_indels(m, mut, ev, lens, ins)
    st = int(m.integers(0, len(mut) - 5000))
    mut = _invert(mut, st, 5000)
    return ref, mut


def _branch(rng, seq):
    """One branch of the multi-genome tree: 1-5% SNPs, indels at 0.2% per base
    (lengths 1-20, half insertions), two inversions of 10-50 kb and one
    deletion of 2-20 kb."""
    out = _mutate_snps(rng, seq, rng.uniform(0.01, 0.05))
    ev = np.nonzero(rng.random(len(out)) < 0.002)[0]
    out = _apply_indels(rng, out, ev, rng.integers(1, 21, size=len(ev)), rng.random(len(ev)) < 0.5)
    for _ in range(2):
        ln = int(rng.integers(10_000, 50_001))
        out = _invert(out, int(rng.integers(0, max(1, len(out) - ln))), ln)
    ln = int(rng.integers(2_000, 20_001))
    at = int(rng.integers(0, max(1, len(out) - ln)))
    return np.concatenate([out[:at], out[at + ln:]])


def genome_family(length=5_000_000, n=8, seed=501, gc=0.45):
    """n genomes derived from one random ancestor (length bases, GC fraction
    gc) along a random binary tree: a random leaf is split into two children
    until there are n leaves, each child one branch of mutations (_branch).
    Returns the leaves (ASCII uint8 arrays) in the order they were made."""
    rng = np.random.Generator(np.random.PCG64(seed))
    leaves = [random_bases(rng, length, gc)]
    while len(leaves) < n:
        parent = leaves.pop(int(rng.integers(0, len(leaves))))
        leaves += [_branch(rng, parent), _branch(rng, parent)]
    return leaves


def _orf(rng, n_codons):
    body = rng.integers(0, len(SENSE), size=n_codons - 2)
    cods = ["ATG"] + [SENSE[i] for i in body] + [STOPS[rng.integers(0, 3)]]
    return np.frombuffer("".join(cods).encode(), dtype=np.uint8)


def pair_recoded(length=4_600_000, seed_ref=201, seed_mut=202, coverage=0.88):
    """Recoded-genome pair (ASCII uint8 arrays) plus the ORF table of the reference.

    ORFs are laid out left to right with random spacers so that they cover
    ``coverage`` of the sequence; each is on the forward or reverse strand.
    """
    r = np.random.Generator(np.random.PCG64(seed_ref))
    ref = random_bases(r, length, 0.5)
    orfs = []
    pos = 0
    mean_orf = (300 + 3000) / 2.0
    mean_gap = mean_orf * (1 - coverage) / coverage
    while True:
        gap = int(r.exponential(mean_gap))
        n_cod = int(r.integers(100, 1001))
        L = 3 * n_cod
        if pos + gap + L > length:
            break
        s = pos + gap
        o = _orf(r, n_cod)
        strand = int(r.integers(0, 2))
        ref[s:s + L] = o if strand == 0 else revcomp(o)
        orfs.append((s, L, strand))
        pos = s + L
    m = np.random.Generator(np.random.PCG64(seed_mut))
    mut = ref.copy()
    # synonymous recoding of 5% of codons (in ORF frame, either strand)
    for s, L, strand in orfs:
        n_cod = L // 3
        pick = np.nonzero(m.random(n_cod - 2) < 0.05)[0] + 1
        if len(pick) == 0:
            continue
        seg = mut[s:s + L] if strand == 0 else revcomp(mut[s:s + L])
        seg = seg.copy()
        for ci in pick:
            cod = bytes(seg[3 * ci:3 * ci + 3]).decode()
            alts = [c for c in _SYN.get(_CODON_AA.get(cod, "*"), []) if c != cod]
            if not alts:
                continue
            seg[3 * ci:3 * ci + 3] = np.frombuffer(alts[int(m.integers(0, len(alts)))].encode(), dtype=np.uint8)
        mut[s:s + L] = seg if strand == 0 else revcomp(seg)
    mut = _mutate_snps(m, mut, 0.002)
    ev = np.sort(m.choice(len(mut) - 1000, size=20, replace=False))
    lens = m.integers(1, 201, size=20)
    ins = m.random(20) < 0.5
    mut = _apply_indels(m, mut, ev, lens, ins)
    for _ in range(3):
        ln = int(m.integers(10_000, 50_001))
        st = int(m.integers(0, len(mut) - ln))
        mut = _invert(mut, st, ln)
    return ref, mut, orfs


def _derive_chunk(rng, A, la, d):
    """B for a run of pairs whose A sequences are concatenated in A."""
    n = len(la)
    starts = np.zeros(n, dtype=np.int64)
    starts[1:] = np.cumsum(la)[:-1]
    ends = starts + la
    pair_id = np.repeat(np.arange(n, dtype=np.int32), la)
    u = rng.random(len(A), dtype=np.float32)
    ev = u < d[pair_id].astype(np.float32)
    kind = rng.random(len(A), dtype=np.float32)
    sub = ev & (kind < 0.9)
    ins = ev & (kind >= 0.9) & (kind < 0.95)
    dele = ev & (kind >= 0.95)
    B = A.copy()
    B[sub] = (B[sub] + rng.integers(1, 4, size=int(sub.sum()), dtype=np.uint8)) % 4
    dpos = np.nonzero(dele)[0]
    keep = np.ones(len(A), dtype=bool)
    if len(dpos):
        glen = rng.geometric(0.3, size=len(dpos))
        stop = np.minimum(dpos + glen, ends[pair_id[dpos]])
        cover = np.bincount(dpos, minlength=len(A) + 1) - np.bincount(stop, minlength=len(A) + 1)
        keep = np.cumsum(cover[:-1]) == 0
    icount = np.zeros(len(A), dtype=np.int64)
    ipos = np.nonzero(ins)[0]
    if len(ipos):
        icount[ipos] = rng.geometric(0.3, size=len(ipos))
    reps = keep.astype(np.int64) + icount
    src = np.repeat(B, reps)
    is_orig = np.zeros(len(src), dtype=bool)
    first = np.cumsum(reps) - reps
    is_orig[first[keep]] = True
    newb = rng.integers(0, 4, size=len(src), dtype=np.uint8)
    Bout = np.where(is_orig, src, newb).astype(np.uint8)
    blen = np.bincount(pair_id, weights=reps, minlength=n).astype(np.int64)
    return Bout, blen


def _dp_chunk(args):
    ss, la, d = args
    rng = np.random.Generator(np.random.PCG64(ss))
    A = rng.integers(0, 4, size=int(la.sum()), dtype=np.uint8)
    B, blen = _derive_chunk(rng, A, la, d)
    return A, B, blen


def dp_jobs(n=100_000, seed=301, min_len=10, max_ratio=2000, div=(0.02, 0.30), chunk=2000, workers=1):
    """Batched DP job set as letter codes (0..3).

    len(A) = floor(min_len * max_ratio**u), u ~ U(0,1); B derived from A at
    divergence d ~ U(div): per base an event with probability d, 90%
    substitutions and 10% indels (half insertions, half deletions) of
    Geom(0.3) length.  An empty B gets one random base (the DP rejects empty
    operands, as the reference does).  Lengths and divergences come from
    PCG64(seed); the sequences of each run of `chunk` pairs come from its own
    spawned PCG64 stream, so chunks can be generated by `workers` parallel
    processes (spawned: the calling script needs an `if __name__ == "__main__"`
    guard; call this before opening a GPU context).

    Returns (a_codes, a_off, a_len, b_codes, b_off, b_len); all A sequences
    are concatenated in a_codes (uint8) and likewise for B.
    """
    ss = np.random.SeedSequence(seed)
    rng = np.random.Generator(np.random.PCG64(ss))
    u = rng.random(n)
    la = np.maximum(np.floor(min_len * np.power(float(max_ratio), u)).astype(np.int64), 1)
    d = rng.uniform(div[0], div[1], size=n)
    bounds = list(range(0, n, chunk)) + [n]
    kids = ss.spawn(len(bounds) - 1)
    tasks = [(kids[i], la[bounds[i]:bounds[i + 1]], d[bounds[i]:bounds[i + 1]]) for i in range(len(bounds) - 1)]
    if workers > 1 and n > 4 * chunk:
        import concurrent.futures as cf
        import multiprocessing as mp
        with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
            parts = list(ex.map(_dp_chunk, tasks))
    else:
        parts = [_dp_chunk(t) for t in tasks]
    A = np.concatenate([p[0] for p in parts])
    B = np.concatenate([p[1] for p in parts])
    blen = np.concatenate([p[2] for p in parts])
    b_off = np.zeros(n, dtype=np.int64)
    b_off[1:] = np.cumsum(blen)[:-1]
    empty = np.nonzero(blen == 0)[0]
    if len(empty):
        fill = rng.integers(0, 4, size=len(empty), dtype=np.uint8)
        pieces, prev = [], 0
        for k, e in enumerate(empty):
            pieces.append(B[prev:b_off[e]])
            pieces.append(fill[k:k + 1])
            prev = b_off[e]
        pieces.append(B[prev:])
        B = np.concatenate(pieces)
        blen[empty] = 1
        b_off[1:] = np.cumsum(blen)[:-1]
    a_off = np.zeros(n, dtype=np.int64)
    a_off[1:] = np.cumsum(la)[:-1]
    return A, a_off, la, B, b_off, blen


def ascii_to_codes(seq):
    """ACGT (any case) -> 0..3; anything else raises."""
    s = np.frombuffer(bytes(seq), dtype=np.uint8) if not isinstance(seq, np.ndarray) else seq
    lut = np.full(256, 255, dtype=np.uint8)
    for i, c in enumerate(b"ACGT"):
        lut[c] = i
        lut[c + 32] = i
    out = lut[s]
    if (out == 255).any():
        raise ValueError("non-ACGT residue")
    return out


def codes_to_ascii(codes):
    return BASES[np.asarray(codes, dtype=np.uint8)]
