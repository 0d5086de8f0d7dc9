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


def main():
    if not os.path.exists(HARNESS):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    n1 = write_dp(os.path.join(HERE, "dp_golden.txt"))
    n2 = write_keys(os.path.join(HERE, "keys_golden.npz"))
    print(f"wrote {n1} DP cases and {n2} key cases")


if __name__ == "__main__":
    sys.exit(main())
