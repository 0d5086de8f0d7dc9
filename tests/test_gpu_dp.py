"""GPU parity of the batched NWSmall DP (mh_dp_align) against the oracle.

Bar: bit-exact paths and terminal scores (integer DP).  The oracle itself is
pinned to the reference libMUSCLE build by tests/test_oracle_dp.py.
"""
import numpy as np
import pytest

import mauvehip
from mauvehip import synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _mutate(rng, a, d):
    out = []
    for x in a:
        r = rng.random()
        if r < d * 0.9:
            out.append(int(rng.integers(0, 4)))
        elif r < d * 0.95:
            continue
        elif r < d:
            out.append(int(x))
            out.append(int(rng.integers(0, 4)))
        else:
            out.append(int(x))
    return np.array(out or [0], dtype=np.uint8)


def _jobs_from_pairs(pairs):
    a = np.concatenate([p[0] for p in pairs]).astype(np.uint8)
    b = np.concatenate([p[1] for p in pairs]).astype(np.uint8)
    la = np.array([len(p[0]) for p in pairs], dtype=np.uint32)
    lb = np.array([len(p[1]) for p in pairs], dtype=np.uint32)
    ao = np.concatenate([[0], np.cumsum(la)[:-1]]).astype(np.uint64)
    bo = np.concatenate([[0], np.cumsum(lb)[:-1]]).astype(np.uint64)
    return a, b, mauvehip.make_jobs(ao, la, bo, lb)


def _check_against_oracle(ctx, pairs):
    a, b, jobs = _jobs_from_pairs(pairs)
    paths, res = mauvehip.global_align(ctx, a, b, jobs)
    bad = []
    for j, (pa, pb) in enumerate(pairs):
        p = mauvehip.path_of(paths, res, j)
        op, osc = po.nwsmall(pa, pb)
        if p != op or int(res[j]["score"]) != int(osc):
            bad.append((j, len(pa), len(pb)))
    return bad


def test_dp_small_edge_cases(ctx):
    rng = np.random.default_rng(11)
    pairs = []
    for la in (1, 2, 3, 7, 8, 9, 63, 64, 65):
        for lb in (1, 2, 3, 7, 8, 9, 63, 64, 65, 511, 512, 513):
            pairs.append((rng.integers(0, 4, la, dtype=np.uint8), rng.integers(0, 4, lb, dtype=np.uint8)))
    assert _check_against_oracle(ctx, pairs) == []


def test_dp_random_related(ctx):
    rng = np.random.default_rng(12)
    pairs = []
    for _ in range(400):
        la = int(rng.choice([5, 20, 100, 300, 700, 1500]))
        a = rng.integers(0, 4, la, dtype=np.uint8)
        b = _mutate(rng, a, float(rng.choice([0.01, 0.05, 0.15, 0.3])))
        pairs.append((a, b))
    assert _check_against_oracle(ctx, pairs) == []


def test_dp_multi_stripe_long(ctx):
    rng = np.random.default_rng(13)
    pairs = []
    for la, lb in ((3000, 2000), (1025, 1537), (4000, 3999), (130, 5000), (5000, 130)):
        a = rng.integers(0, 4, la, dtype=np.uint8)
        b = _mutate(rng, a, 0.1)[:lb] if lb <= la else np.concatenate([_mutate(rng, a, 0.1), rng.integers(0, 4, lb, dtype=np.uint8)])[:lb]
        pairs.append((a, b))
    assert _check_against_oracle(ctx, pairs) == []


@pytest.mark.parametrize("extra", [0, 1, 2, 3])
def test_dp_row_packed_shapes(ctx, extra):
    """Jobs up to 256 columns wide with short trace-back walks run four to a
    wavefront (one per 16-lane row, stripes of 64 to 256 columns in steps of
    32, 4 to 16 columns a lane): the shape boundaries, walks either side of the
    wavefront trace-back limit
    (a + b = 1024), related and unrelated pairs, and job counts that leave
    the last wavefront part empty."""
    rng = np.random.default_rng(40 + extra)
    pairs = []
    for lb in (17, 18, 63, 64, 65, 95, 96, 97, 127, 128, 129, 159, 160, 161, 191, 192, 193, 223, 224, 225, 255, 256,
               257):
        for la in (1, 2, 16, 17, 100, 766, 767, 768, 1000):
            a = rng.integers(0, 4, la, dtype=np.uint8)
            if rng.random() < 0.5:
                b = _mutate(rng, a, 0.05)
                b = np.concatenate([b, rng.integers(0, 4, max(0, lb - len(b)), dtype=np.uint8)])[:lb]
            else:
                b = rng.integers(0, 4, lb, dtype=np.uint8)
            pairs.append((a, b))
    pairs = pairs[:len(pairs) - (len(pairs) % 4) + extra] if extra else pairs[:len(pairs) - (len(pairs) % 4)]
    assert _check_against_oracle(ctx, pairs) == []


def test_dp_zero_length_is_an_error(ctx):
    jobs = mauvehip.make_jobs([0], [0], [0], [5])
    with pytest.raises(mauvehip.MhError) as e:
        mauvehip.global_align(ctx, np.zeros(4, np.uint8), np.zeros(5, np.uint8), jobs)
    assert e.value.code == -5


def test_dp_batch_sample_parity(ctx):
    """A slice of the 100k-pair batched DP workload (seed 301) checked job by job."""
    A, ao, al, B, bo, bl = synth.dp_jobs(n=2000, seed=301, max_ratio=300)
    jobs = mauvehip.make_jobs(ao, al, bo, bl)
    paths, res = mauvehip.global_align(ctx, A, B, jobs)
    rng = np.random.default_rng(5)
    for j in rng.choice(len(al), 300, replace=False):
        a = A[ao[j]:ao[j] + al[j]]
        b = B[bo[j]:bo[j] + bl[j]]
        op, osc = po.nwsmall(a, b)
        assert mauvehip.path_of(paths, res, j) == op, j
        assert int(res[j]["score"]) == int(osc), j
    # every path consumes exactly its operands
    for j in range(len(al)):
        p = mauvehip.path_of(paths, res, j)
        assert p.count("M") + p.count("D") == al[j]
        assert p.count("M") + p.count("I") == bl[j]
