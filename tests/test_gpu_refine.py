"""GPU parity of mh_refine_pairwise (the pairwise gapped-alignment passes) with
the refinement oracle: identical rows for every interval of a batch, for the
profile pass alone, the refinement pass alone and both."""
import numpy as np
import pytest

import mauvehip
from oracle import refine_oracle as ro
from refine_cases import alignment

pytestmark = pytest.mark.gpu


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.choice([60, 400, 2500, 9000, 24000]))
        out.append(alignment(rng, n, div=float(rng.choice([0.02, 0.1, 0.25])),
                             marble=float(rng.choice([0.0, 0.2, 0.6])),
                             big_gap=int(rng.choice([0, 0, 5600])) if n > 5000 else 0, lower=bool(k % 4 == 3)))
    return out


@pytest.mark.parametrize("passes", [mauvehip.REFINE_PROFILE, mauvehip.REFINE_WINDOWS,
                                    mauvehip.REFINE_PROFILE | mauvehip.REFINE_WINDOWS])
def test_refine_batch_matches_oracle(ctx, passes):
    ivs = _cases(7 + passes, 8)
    got = [(a.decode(), b.decode()) for a, b in mauvehip.refine_pairwise(ctx, ivs, passes)]
    for (r1, r2), (g1, g2) in zip(ivs, got):
        if passes == mauvehip.REFINE_PROFILE:
            want = ro.refine_alignment(r1, r2, True)
        elif passes == mauvehip.REFINE_WINDOWS:
            want = ro.refine_alignment(r1, r2, False)
        else:
            want = ro.gapped_alignment(r1, r2)
        assert (g1, g2) == want
        assert len(g1) == len(g2)
        assert g1.replace("-", "") == r1.replace("-", "") and g2.replace("-", "") == r2.replace("-", "")


def test_refine_single_interval_and_corners(ctx):
    cases = [("ACGTACGTAC", "ACGTACGTAC"), ("A" * 30, "-" * 30), ("ACGT" * 30, "AC-T" * 30)]
    got = [(a.decode(), b.decode()) for a, b in mauvehip.refine_pairwise(ctx, cases)]
    for (r1, r2), g in zip(cases, got):
        assert g == ro.gapped_alignment(r1, r2)
    assert got[1] == cases[1]                       # one genome only: left as it is


def _with_wildcards(rng, r, rate=0.01):
    b = bytearray(r.encode())
    for i in np.nonzero(rng.random(len(b)) < rate)[0]:
        if b[i] != ord("-"):
            b[i] = ord(rng.choice(list("NNNNXRYUunZ")))
    return b.decode()


@pytest.mark.parametrize("passes", [mauvehip.REFINE_PROFILE, mauvehip.REFINE_PROFILE | mauvehip.REFINE_WINDOWS])
def test_refine_wildcards_match_oracle(ctx, passes):
    """Rows with N, X, R, Y, U, lower case and letters outside the alphabet:
    the anchored ranges and windows that hold them run NWSmall in float with
    libMUSCLE's wildcard scores; rows equal the oracle's (itself pinned to
    libMUSCLE on wildcard pairs and profile passes)."""
    rng = np.random.default_rng(90 + passes)
    ivs = []
    for r1, r2 in _cases(91 + passes, 8):
        ivs.append((_with_wildcards(rng, r1), _with_wildcards(rng, r2)))
    ivs.append(("ACGTN" * 60, "ACGTA" * 60))
    ivs.append(("NNNNNNNNNNACGTACGTACNNNN" * 10, "ACGTACGTACGTACGTACGTACGT" * 10))
    got = [(a.decode(), b.decode()) for a, b in mauvehip.refine_pairwise(ctx, ivs, passes)]
    for (r1, r2), g in zip(ivs, got):
        want = ro.refine_alignment(r1, r2, True) if passes == mauvehip.REFINE_PROFILE else ro.gapped_alignment(r1, r2)
        assert g == want


def test_refine_after_dp_batches_and_growing_staging():
    """A fresh context that first runs a batched DP (its job table in pinned
    memory), then refinement batches whose pinned staging grows from one call
    to the next: every result equals the oracle's.  (Growing the staging once
    released the job table's pinned memory, which the next DP call then wrote
    into.)"""
    from mauvehip import synth
    c = mauvehip.Context(0)
    try:
        A, ao, al, B, bo, bl = synth.dp_jobs(n=3000, seed=301, max_ratio=300)
        mauvehip.global_align(c, A, B, mauvehip.make_jobs(ao, al, bo, bl))
        for seed, count in ((61, 2), (62, 6), (63, 12)):
            ivs = _cases(seed, count)
            got = [(a.decode(), b.decode()) for a, b in mauvehip.refine_pairwise(c, ivs)]
            for (r1, r2), g in zip(ivs, got):
                assert g == ro.gapped_alignment(r1, r2)
            mauvehip.global_align(c, A, B, mauvehip.make_jobs(ao, al, bo, bl))
    finally:
        c.close()
