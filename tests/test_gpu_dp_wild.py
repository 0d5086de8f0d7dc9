"""GPU parity of the DP on operands with letters other than A, C, G, T:
such jobs run NWSmall in float with libMUSCLE's wildcard scores
(dp_float.hip).  Paths equal libMUSCLE's own (tests/golden/dp_wild_golden.txt)
and the oracle's; the score is the oracle's float score rounded; a batch that
mixes them with A/C/G/T jobs, the device-resident, banded and no-allocation
entry points give the same results."""
import os

import numpy as np
import pytest

import mauvehip
from mauvehip import synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    out = []
    for line in open(os.path.join(HERE, "golden", "dp_wild_golden.txt")):
        if line.startswith("#"):
            continue
        a, b, ra, rb = line.split()
        path = "".join("M" if x != "-" and y != "-" else ("D" if x != "-" else "I") for x, y in zip(ra, rb))
        out.append((a, b, path))
    return out


def _pack(pairs):
    a = np.concatenate([po.codes(p[0]) for p in pairs])
    b = np.concatenate([po.codes(p[1]) for p in pairs])
    la = np.array([len(p[0]) for p in pairs], dtype=np.uint32)
    lb = np.array([len(p[1]) for p in pairs], dtype=np.uint32)
    ao = np.concatenate([[0], np.cumsum(la)[:-1]]).astype(np.uint64)
    bo = np.concatenate([[0], np.cumsum(lb)[:-1]]).astype(np.uint64)
    return a, b, mauvehip.make_jobs(ao, la, bo, lb)


def _mixed(seed=5, n=400):
    """A/C/G/T pairs of the batched-DP distribution, a third of them with
    wildcard letters sprinkled in."""
    A, ao, al, B, bo, bl = synth.dp_jobs(n=n, seed=seed, max_ratio=100)
    rng = np.random.default_rng(seed)
    letters = np.frombuffer(b"ACGT", dtype=np.uint8)
    pairs = []
    for j in range(n):
        a = letters[A[ao[j]:ao[j] + al[j]]].tobytes().decode()
        b = letters[B[bo[j]:bo[j] + bl[j]]].tobytes().decode()
        if j % 3 == 0:
            a = "".join(rng.choice(list("NXRYUn")) if rng.random() < 0.03 else c for c in a)
            b = "".join(rng.choice(list("NXRYUn")) if rng.random() < 0.03 else c for c in b)
        pairs.append((a, b))
    return pairs


def test_wildcard_pairs_match_libmuscle(ctx):
    cases = _golden()
    a, b, jobs = _pack([(x, y) for x, y, _ in cases])
    paths, res = mauvehip.global_align(ctx, a, b, jobs)
    for j, (x, y, want) in enumerate(cases):
        got = mauvehip.path_of(paths, res, j)
        assert got == want, j
        op, osc = po.nwsmall(x, y)
        assert got == op and int(res[j]["score"]) == int(np.rint(np.float32(osc))), j


def test_mixed_batch_and_entry_points(ctx):
    pairs = _mixed()
    a, b, jobs = _pack(pairs)
    paths, res = mauvehip.global_align(ctx, a, b, jobs)
    want = [po.nwsmall(x, y) for x, y in pairs]
    for j, (op, osc) in enumerate(want):
        assert mauvehip.path_of(paths, res, j) == op, j
        assert int(res[j]["score"]) == int(np.rint(np.float32(osc))), j
    total = int((jobs["a_len"].astype(np.int64) + jobs["b_len"]).sum())
    dA, dB = ctx.alloc(len(a) + 16), ctx.alloc(len(b) + 16)
    dP, dR, dD = ctx.alloc(total + 16), ctx.alloc(16 * len(jobs) + 16), ctx.alloc(4 * len(jobs) + 16)
    try:
        ctx.h2d(dA, a)
        ctx.h2d(dB, b)
        for name in ("device", "noalloc", "banded"):
            if name == "device":
                mauvehip.global_align_device(ctx, dA, dB, jobs, dP, dR)
            elif name == "noalloc":
                mauvehip._check(ctx.L.mh_dp_align_device_noalloc(ctx.h, dA, dB, jobs.ctypes.data, len(jobs), dP, dR))
            else:
                bands = np.full(len(jobs), 16, dtype=np.uint32)
                mauvehip.global_align_banded_device(ctx, dA, dB, jobs, bands, dP, dR, dD)
            p2 = ctx.d2h(np.zeros(total + 16, np.uint8), dP)
            r2 = ctx.d2h(np.zeros(len(jobs), mauvehip.DP_RESULT_DTYPE), dR)
            for j, (op, osc) in enumerate(want):
                wild = j % 3 == 0
                if name != "banded" or wild:
                    assert mauvehip.path_of(p2, r2, j) == op, (name, j)
            if name == "banded":
                dev = ctx.d2h(np.zeros(len(jobs), np.uint32), dD)
                for j in range(0, len(jobs), 3):     # the float jobs run the full DP: dev is the label
                    assert dev[j] == po.path_dev(len(pairs[j][0]), len(pairs[j][1]), want[j][0]), j
    finally:
        for p in (dA, dB, dP, dR, dD):
            ctx.free(p)


def test_ascii_to_codes_classes(ctx):
    text = b"ACGTacgtUuXxNnRYKMSWBDHVZ*"
    d = ctx.alloc(len(text) + 16)
    c = ctx.alloc(len(text) + 16)
    try:
        ctx.h2d(d, np.frombuffer(text, dtype=np.uint8))
        mauvehip.ascii_to_codes_device(ctx, d, len(text), c)
        got = ctx.d2h(np.zeros(len(text), np.uint8), c)
        assert got.tolist() == po.codes(text).tolist()
        ctx.h2d(d, np.frombuffer(b"ACG-T" + bytes(len(text) - 5), dtype=np.uint8))
        with pytest.raises(mauvehip.MhError) as e:
            mauvehip.ascii_to_codes_device(ctx, d, 5, c)
        assert e.value.code == -4
    finally:
        ctx.free(d)
        ctx.free(c)
