"""Calibration of the oracle's NWSmall against libMUSCLE's own, compiled from
the reference sources (oracle/_ref/ref_harness, `make -C oracle ref`).

TEST INFRASTRUCTURE ONLY: runs in the build container, where the reference
tree and the harness exist; nothing in the product or on the GPU runs it.

Takes the batched-DP workload (100,000 pairs, synth.dp_jobs seed 301), the
jobs of a seeded random order up to about 1.2e9 cells (the order the bench's
CPU baseline samples), and times on one thread
  * ref_harness nw: MU/profile.cpp ProfileProfile on two one-sequence MSAs
    (NWSmall, MU/nwsmall.cpp:500-670, and BitTraceBack), whole process;
  * or_nwsmall (oracle/mh_oracle.c) on the same jobs;
then checks the two aligned rows of every job are identical.

    python oracle/calibrate_nwsmall.py [--cells 1.2e9]
"""
import argparse
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "mauve-py_amd"))

from oracle import pyoracle as po  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=float, default=1.2e9)
    args = ap.parse_args()
    harness = os.path.join(HERE, "_ref", "ref_harness")
    if not os.path.exists(harness):
        print("oracle/_ref/ref_harness is not built (make -C oracle ref)", file=sys.stderr)
        return 2
    from mauvehip import synth
    A, ao, al, B, bo, bl = synth.dp_jobs(n=100_000, seed=301, workers=min(8, os.cpu_count() or 1))
    order = np.random.default_rng(5).permutation(len(al))
    cells = np.cumsum(al[order].astype(np.int64) * bl[order])
    n = int(np.searchsorted(cells, args.cells) + 1)
    jobs = order[:n]
    letters = np.frombuffer(b"ACGT", dtype=np.uint8)
    seqs = [(letters[A[ao[k]:ao[k] + al[k]]].tobytes().decode(), letters[B[bo[k]:bo[k] + bl[k]]].tobytes().decode())
            for k in jobs]
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.txt"), os.path.join(d, "out.txt")
        with open(fin, "w") as f:
            for a, b in seqs:
                f.write(f"{a} {b}\n")
        t0 = time.perf_counter()
        subprocess.run([harness, "nw", fin, fout], check=True, stderr=subprocess.DEVNULL)
        t_ref = time.perf_counter() - t0
        ref_rows = [tuple(line.split()) for line in open(fout)]
    t0 = time.perf_counter()
    ours = []
    for a, b in seqs:
        path, _ = po.nwsmall(a, b)
        ours.append(po.rows_from_path(a, b, path))
    t_or = time.perf_counter() - t0
    total = int(cells[n - 1])
    same = ref_rows == ours
    print(f"jobs {n}, cells {total}: libMUSCLE ProfileProfile {t_ref:.2f} s ({total / t_ref / 1e9:.4f} GCUPS), "
          f"oracle or_nwsmall {t_or:.2f} s ({total / t_or / 1e9:.4f} GCUPS), ratio {t_ref / t_or:.3f}; "
          f"rows identical: {same}")
    return 0 if same else 1


if __name__ == "__main__":
    sys.exit(main())
