"""Synthetic two-row alignments for the refinement tests: an edit script of
substitutions and indels, with some stretches laid out unaligned (genome 1's
residues against gaps, then gaps against genome 2's), optionally one long
one-sided stretch (over 5000 columns) and lowercase residues."""
import numpy as np

BASES = np.array(list("ACGT"))


def alignment(rng, n, div=0.08, marble=0.05, big_gap=0, lower=False):
    a = BASES[rng.integers(0, 4, n)]
    r1, r2 = [], []
    i = 0
    while i < n:
        u = rng.random()
        if u < div * 0.1:                                   # insertion in genome 2
            k = int(rng.geometric(0.3))
            r1.extend("-" * k)
            r2.extend(BASES[rng.integers(0, 4, k)])
        elif u < div * 0.2:                                 # deletion from genome 2
            k = min(int(rng.geometric(0.3)), n - i)
            r1.extend(a[i:i + k])
            r2.extend("-" * k)
            i += k
        elif u < div:                                       # substitution
            r1.append(a[i])
            r2.append(BASES[(np.searchsorted(BASES, a[i]) + rng.integers(1, 4)) % 4])
            i += 1
        elif u < div + marble * 0.02:                       # an unaligned stretch
            k = min(int(rng.integers(5, 200)), n - i)
            kb = int(rng.integers(0, 200))
            r1.extend(list(a[i:i + k]) + ["-"] * kb)
            r2.extend(["-"] * k + list(BASES[rng.integers(0, 4, kb)]))
            i += k
        else:
            r1.append(a[i])
            r2.append(a[i])
            i += 1
        if big_gap and i > n // 2:
            r1.extend(BASES[rng.integers(0, 4, big_gap)])
            r2.extend("-" * big_gap)
            big_gap = 0
    s1, s2 = "".join(r1), "".join(r2)
    if lower:
        s1, s2 = s1.lower(), s2.lower()
    return s1, s2
