"""Write the golden vectors under tests/golden/ from the reference's own code.

Runs oracle/_ref/ref_harness, built by `make -C oracle ref` from the reference
sources (libMUSCLE, libGenome, SortedMerList.cpp, MemorySML.cpp,
DNAMemorySML.cpp) in this container, and stores its outputs as data:

* dp_golden.txt   -- "A B rowA rowB" per line: ProfileProfile (MU/profile.cpp:68-93,
                     driving GlobalAlign -> NWSmall -> BitTraceBack) on two
                     one-sequence profiles set up as ProfileAlignFast does.
* dp_wild_golden.txt -- the same for operands with wildcard and other
                     letters (X, N, R, Y, ..., lower case, U, invalid letters),
                     which libMUSCLE scores with fractional profile counts in
                     float (MU/msa2.cpp:19-90).
* keys_golden.npz -- per case: the sequence, seed, packed words, the fill keys
                     in position order, GetDnaSeedMer per position and the sorted
                     positions exactly as the reference std::sort left them.
* refine_golden.txt -- libMUSCLE's own profile pass and refinement windows on
                     two-row gapped alignments (gaps of many lengths, so the
                     column scores carry fractional gap shares; terminal gaps;
                     lower case; wildcard letters in the score-only cases):
                     "APP row1 row2" followed by LetterObjScoreXP's column
                     scores (float bits), FindAnchorColsPP's anchor columns
                     and AnchoredProfileProfile's rows (MU/anchoredpp.cpp),
                     and "RW window row1 row2" followed by RefineW's rows
                     (MU/refinew.cpp) under RefineFast's globals.
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


def dp_wild_cases():
    rng = np.random.default_rng(20261203)
    alph = "ACGT"
    wild = "XNRYKMSWBDHVnxuUZ"
    pairs = []

    def seq(n, pw):
        return "".join(wild[rng.integers(0, len(wild))] if rng.random() < pw else alph[rng.integers(0, 4)]
                       for _ in range(n))
    for la in (1, 2, 3, 5):
        for lb in (1, 2, 4, 9):
            pairs.append((seq(la, 0.5), seq(lb, 0.5)))
    for _ in range(120):
        la = int(rng.choice([4, 12, 40, 90, 160]))
        a = seq(la, float(rng.choice([0.02, 0.1, 0.4])))
        if rng.random() < 0.7:
            b = _mutate(rng, a, float(rng.choice([0.02, 0.1, 0.25])))
            b = "".join(wild[rng.integers(0, len(wild))] if rng.random() < 0.05 else c for c in b)
        else:
            b = seq(int(rng.integers(1, 150)), 0.2)
        pairs.append((a, b))
    pairs.append(("N" * 30, "N" * 25))
    pairs.append(("X" * 20 + "ACGT", "AGAG" * 6))
    return pairs


def write_dp_wild(path):
    pairs = dp_wild_cases()
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.txt"), os.path.join(td, "out.txt")
        with open(fin, "w") as f:
            for a, b in pairs:
                f.write(f"{a} {b}\n")
        subprocess.check_call([HARNESS, "nw", fin, fout], stderr=subprocess.DEVNULL)   # MUSCLE warns of invalid letters
        rows = [l.split() for l in open(fout)]
    assert len(rows) == len(pairs)
    with open(path, "w") as f:
        f.write("# A B rowA rowB  (reference ProfileProfile/NWSmall output; wildcard letters)\n")
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


def _app_safe(r1, r2):
    """True when every range between the oracle's anchor columns holds
    residues of both rows (otherwise libMUSCLE's NWSmall quits the process)."""
    sys.path.insert(0, ROOT)
    from oracle import refine_oracle as ro
    for lo, hi in ro.cols_to_ranges(ro.anchor_cols(r1, r2), len(r1)):
        if hi == lo:
            continue
        if not ro._residues(r1[lo:hi]) or not ro._residues(r2[lo:hi]):
            return False
    return True


def _refine_inputs():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refine_cases import alignment
    rng = np.random.default_rng(99)
    app, rw, score_only = [], [], []
    while len(app) < 50:
        n = int(rng.choice([40, 120, 300, 600]))
        r1, r2 = alignment(rng, n, div=float(rng.choice([0.02, 0.1, 0.25, 0.4])),
                           marble=float(rng.choice([0.0, 0.1, 0.5])), lower=bool(rng.random() < 0.2))
        if len(r1) < 2 or not _app_safe(r1, r2):
            continue
        app.append((r1, r2))
    # hand-placed gaps of 1..9 columns (shares of -400 by 1..9), terminal gaps
    for k in range(1, 10):
        core = "".join("ACGT"[x] for x in rng.integers(0, 4, 80))
        r1 = core[:30] + "-" * k + core[30:]
        r2 = core[:30] + "".join("ACGT"[x] for x in rng.integers(0, 4, k)) + core[30:]
        r1 = "-" * (k % 3) + r1 + "-" * (k % 2)
        r2 = "".join("ACGT"[x] for x in rng.integers(0, 4, k % 3)) + r2 + "".join("ACGT"[x] for x in rng.integers(0, 4, k % 2))
        if _app_safe(r1, r2):
            app.append((r1, r2))
    # wildcards: column scores and anchor columns only
    for _ in range(12):
        r1, r2 = alignment(rng, int(rng.choice([60, 200, 400])), div=0.1, marble=0.1)
        b1, b2 = bytearray(r1.encode()), bytearray(r2.encode())
        for b in (b1, b2):
            for i in rng.integers(0, len(b), max(1, len(b) // 30)):
                if b[i] != ord("-"):
                    b[i] = ord(rng.choice(list("NRYKMSWBDHVXUu")))
        if _app_safe(b1.decode(), b2.decode()):
            score_only.append((b1.decode(), b2.decode()))
    while len(rw) < 24:
        n = int(rng.choice([150, 600, 1200]))
        r1, r2 = alignment(rng, n, div=float(rng.choice([0.02, 0.1, 0.3])), marble=float(rng.choice([0.0, 0.2])),
                           lower=bool(rng.random() < 0.2))
        w = int(rng.choice([200, 500]))
        # RefineW needs residues in every window
        ok = all(any(c != "-" for c in r1[f:f + w] + r2[f:f + w]) for f in range(0, len(r1), w))
        if ok:
            rw.append((w, r1, r2))
    return app, score_only, rw


def write_refine(path):
    app, wild, rw = _refine_inputs()
    lines = ["# libMUSCLE profile pass (APP, WILD with wildcard letters) and refinement windows (RW); see gen_golden.py"]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.txt"), os.path.join(td, "out.txt")
        for tag, cases in (("APP", app), ("WILD", wild)):
            with open(fin, "w") as f:
                for r1, r2 in cases:
                    f.write(f"{r1} {r2}\n")
            subprocess.check_call([HARNESS, "app", fin, fout])
            res = open(fout).read().splitlines()
            assert len(res) == 3 * len(cases)
            for (r1, r2), k in zip(cases, range(0, len(res), 3)):
                lines.append(f"{tag} {r1} {r2}")
                lines.extend(res[k:k + 3])
        for w, r1, r2 in rw:
            with open(fin, "w") as f:
                f.write(f"{r1} {r2}\n")
            subprocess.check_call([HARNESS, "refinew", str(w), fin, fout])
            res = open(fout).read().split()
            lines.append(f"RW {w} {r1} {r2}")
            lines.append(f"O {res[0]} {res[1]}")
    open(path, "w").write("\n".join(lines) + "\n")
    return len(app) + len(wild), len(rw)


def main():
    if not os.path.exists(HARNESS):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    n1 = write_dp(os.path.join(HERE, "dp_golden.txt"))
    n1 += write_dp_wild(os.path.join(HERE, "dp_wild_golden.txt"))
    n2 = write_keys(os.path.join(HERE, "keys_golden.npz"))
    n3 = write_findmer(os.path.join(HERE, "findmer_golden.npz"))
    n4, n5 = write_refine(os.path.join(HERE, "refine_golden.txt"))
    print(f"wrote {n1} DP cases, {n2} key cases, {n3} FindMer cases, {n4} profile-pass and {n5} window cases")


if __name__ == "__main__":
    sys.exit(main())
