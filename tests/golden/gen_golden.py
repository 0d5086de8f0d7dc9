"""Write the golden vectors under tests/golden/ from the reference's own code.

Runs oracle/_ref/ref_harness, built by `make -C oracle ref` from the reference
sources (libMUSCLE, libGenome, SortedMerList.cpp, MemorySML.cpp,
DNAMemorySML.cpp) in this container, and stores its outputs as data:

* dp_golden.txt   -- "A B rowA rowB" per line: ProfileProfile (MU/profile.cpp:68-93,
                     driving GlobalAlign -> NWSmall -> BitTraceBack) on two
                     one-sequence profiles set up as ProfileAlignFast does.
* keys_golden.npz -- per case: the sequence, seed, packed words, the fill keys
                     in position order, GetDnaSeedMer per position and the sorted
                     positions exactly as the reference std::sort left them.
* findmer_golden.npz -- per case: the sequence, seed, query keys and
                     SortedMerList::FindMer's answer to each (found flag, and
                     the index its bsearch left, LM/SortedMerList.cpp:170-179,
                     380-394), as the repeat-limit restart calls it.

Usage: python tests/golden/gen_golden.py   (needs the reference tree)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def _mutate(rng, a, d):
    out = []
    for ch in a:
        r = rng.random()
        if r < d * 0.9:
            out.append("ACGT"[rng.integers(0, 4)])
        elif r < d * 0.95:
            continue
        elif r < d:
            out.append(ch)
            out.append("ACGT"[rng.integers(0, 4)])
        else:
            out.append(ch)
    return "".join(out) or "A"


def dp_cases():
    rng = np.random.default_rng(20240601)
    pairs = []
    # every small shape, including the one-row / one-column boundary cases
    for la in (1, 2, 3, 4):
        for lb in (1, 2, 3, 4, 9):
            for _ in range(3):
                pairs.append(("".join("ACGT"[x] for x in rng.integers(0, 4, la)),
                              "".join("ACGT"[x] for x in rng.integers(0, 4, lb))))
    for _ in range(200):
        la = int(rng.choice([5, 8, 16, 33, 64, 100, 150]))
        a = "".join("ACGT"[x] for x in rng.integers(0, 4, la))
        if rng.random() < 0.7:
            b = _mutate(rng, a, float(rng.choice([0.02, 0.1, 0.25, 0.5])))
        else:
            b = "".join("ACGT"[x] for x in rng.integers(0, 4, int(rng.integers(1, 150))))
        pairs.append((a, b))
    return pairs


def write_dp(path):
    pairs = dp_cases()
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.txt"), os.path.join(td, "out.txt")
        with open(fin, "w") as f:
            for a, b in pairs:
                f.write(f"{a} {b}\n")
        subprocess.check_call([HARNESS, "nw", fin, fout])
        rows = [l.split() for l in open(fout)]
    assert len(rows) == len(pairs)
    with open(path, "w") as f:
        f.write("# A B rowA rowB  (reference ProfileProfile/NWSmall output)\n")
        for (a, b), (ra, rb) in zip(pairs, rows):
            f.write(f"{a} {b} {ra} {rb}\n")
    return len(pairs)


def read_keys_bin(path):
    b = open(path, "rb").read()
    o = 0

    def take(dt, n):
        nonlocal o
        v = np.frombuffer(b, dt, n, o).copy()
        o += v.nbytes
        return v

    nw = int(take(np.uint64, 1)[0])
    words = take(np.uint32, nw)
    n = int(take(np.uint64, 1)[0])
    fill = take(np.uint64, n)
    seedmer = take(np.uint64, n)
    npos = int(take(np.uint64, 1)[0])
    pos = take(np.uint32, npos)
    return words, fill, seedmer, pos


KEY_CASES = [
    # (length, alphabet, seed)
    (3000, "ACGT", 0x16df6d),           # weight 15, rank 3 (spaced, the CLI's w15 seed)
    (3000, "ACGT", 0x1FFFF),            # weight 17 -> solid
    (2500, "ACGT", 0x7FFFF),            # weight 19 -> solid
    (2000, "ACGTNRYKMacgt", 0x7ac9af),  # weight 15, rank 0, IUPAC + lower case
    (1500, "AAAAAAAACGT", 0x6dbedb),    # weight 17 rank 0, low complexity (many ties)
    (800, "ACGT", 0x7954f),             # weight 12 rank 0 (the table's weight-11 pattern)
    (700, "ACGT", 0x7fffffff),          # solid 31
]


def write_keys(path):
    rng = np.random.default_rng(777)
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for ci, (L, alph, seed) in enumerate(KEY_CASES):
            seq = "".join(alph[x] for x in rng.integers(0, len(alph), L))
            fs, fb = os.path.join(td, "s.seq"), os.path.join(td, "k.bin")
            open(fs, "w").write(seq)
            subprocess.check_call([HARNESS, "keys", hex(seed), fs, fb])
            words, fill, seedmer, pos = read_keys_bin(fb)
            out[f"c{ci}_seq"] = np.frombuffer(seq.encode(), np.uint8)
            out[f"c{ci}_seed"] = np.array([seed], np.uint64)
            out[f"c{ci}_words"] = words
            out[f"c{ci}_fill"] = fill
            out[f"c{ci}_seedmer"] = seedmer
            out[f"c{ci}_pos"] = pos
    np.savez_compressed(path, n_cases=np.array([len(KEY_CASES)]), **out)
    return len(KEY_CASES)


FINDMER_CASES = [
    # (sequence builder, seed)
    ("random", 3000, 0x16df6d),
    ("polya", 3000, 0x1FFFF),
    ("tandem", 4000, 0x16df6d),
    ("random", 1200, 0x7954f),
    ("lowcomp", 2000, 0x1FFFF),
    ("short", 10, 0x1FFFF),       # shorter than the seed: FindMer leaves result unset
    ("exact", 17, 0x1FFFF),       # one mer
]


def _findmer_seq(rng, kind, L):
    r = lambda n: "".join("ACGT"[x] for x in rng.integers(0, 4, n))
    if kind == "polya":
        return r(L // 3) + "A" * (L // 3) + r(L - 2 * (L // 3)) + "T" * 40
    if kind == "tandem":
        return r(L // 4) + "AC" * (L // 8) + r(L // 4) + "ACG" * (L // 12) + r(L - L // 2 - 2 * (L // 8) - 3 * (L // 12))
    if kind == "lowcomp":
        return "".join("AAAAAACG"[x] for x in rng.integers(0, 8, L))
    return r(L)


def write_findmer(path):
    import ctypes
    rng = np.random.default_rng(4242)
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for ci, (kind, L, seed) in enumerate(FINDMER_CASES):
            seq = _findmer_seq(rng, kind, L)
            fs, fb = os.path.join(td, "s.seq"), os.path.join(td, "k.bin")
            open(fs, "w").write(seq)
            subprocess.check_call([HARNESS, "keys", hex(seed), fs, fb])
            _, _, seedmer, pos = read_keys_bin(fb)
            keys = np.sort(seedmer[pos]) if len(pos) else np.zeros(0, np.uint64)
            w = bin(seed).count("1")
            unit = np.uint64(1) << np.uint64(64 - 2 * w)
            masked = keys & ~(unit - np.uint64(1)) if len(keys) else keys
            q = [keys, keys + np.uint64(1), masked + unit, masked | np.uint64(1), masked - unit,
                 np.array([0, 1, unit, 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFFFFFFFFFF - unit + 1], dtype=np.uint64),
                 (rng.integers(0, 2 ** 63, 200, dtype=np.int64).astype(np.uint64) << np.uint64(1))
                 & ~(unit - np.uint64(1))]
            qs = np.unique(np.concatenate(q).astype(np.uint64))
            if len(qs) > 3000:
                qs = np.sort(rng.choice(qs, 3000, replace=False))
            fq, fo = os.path.join(td, "q.bin"), os.path.join(td, "o.bin")
            qs.tofile(fq)
            subprocess.check_call([HARNESS, "findmer", hex(seed), fs, fq, fo])
            res = np.fromfile(fo, dtype=np.uint64).reshape(-1, 2)
            out[f"c{ci}_seq"] = np.frombuffer(seq.encode(), np.uint8)
            out[f"c{ci}_seed"] = np.array([seed], np.uint64)
            out[f"c{ci}_query"] = qs
            out[f"c{ci}_found"] = res[:, 0].astype(np.uint8)
            out[f"c{ci}_result"] = res[:, 1]
    np.savez_compressed(path, n_cases=np.array([len(FINDMER_CASES)]), **out)
    return len(FINDMER_CASES)


def main():
    if not os.path.exists(HARNESS):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    n1 = write_dp(os.path.join(HERE, "dp_golden.txt"))
    n2 = write_keys(os.path.join(HERE, "keys_golden.npz"))
    n3 = write_findmer(os.path.join(HERE, "findmer_golden.npz"))
    print(f"wrote {n1} DP cases, {n2} key cases and {n3} FindMer cases")


if __name__ == "__main__":
    sys.exit(main())
