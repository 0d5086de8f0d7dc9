"""mh_dp_align_device_noalloc: the DP with every scratch buffer in the
library's fixed per-device array, against mh_dp_align_device (bit-exact paths
and scores), on two calls in a row, the second with a smaller input; and the
refusal of a job larger than the array.  Calls go through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import mauvehip
from mauvehip import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jobset():
    # lengths 10-3000, divergence 2-30%: every kernel shape (one-lane,
    # row-packed, stripes) and many scratch-sized rounds
    return synth.dp_jobs(n=1500, seed=311, max_ratio=300)


def _call(ctx, name, dA, dB, jobs):
    total = int((jobs["a_len"].astype(np.int64) + jobs["b_len"]).sum())
    dP = ctx.alloc(total + 16)
    dR = ctx.alloc(16 * len(jobs) + 16)
    try:
        rc = getattr(ctx.L, name)(ctx.h, dA, dB, jobs.ctypes.data, len(jobs), dP, dR)
        mauvehip._check(rc)
        paths = ctx.d2h(np.zeros(total + 16, np.uint8), dP)
        res = ctx.d2h(np.zeros(len(jobs), mauvehip.DP_RESULT_DTYPE), dR)
    finally:
        ctx.free(dP)
        ctx.free(dR)
    return paths, res


def _same(p1, r1, p2, r2):
    assert np.array_equal(r1, r2)
    for j in range(len(r1)):
        s, n = int(r1[j]["path_start"]), int(r1[j]["path_len"])
        assert p1[s:s + n].tobytes() == p2[s:s + n].tobytes(), j


def test_noalloc_matches_dp_align_device(ctx, jobset):
    A, ao, al, B, bo, bl = jobset
    assert ctx.L.mh_dp_noalloc_scratch_bytes() == 64 << 20
    dA, dB = ctx.alloc(len(A) + 16), ctx.alloc(len(B) + 16)
    try:
        ctx.h2d(dA, A)
        ctx.h2d(dB, B)
        for n in (len(al), 300):           # two calls in a row, the second smaller
            jobs = mauvehip.make_jobs(ao[:n], al[:n], bo[:n], bl[:n])
            cells = int((jobs["a_len"].astype(np.int64) * jobs["b_len"]).sum())
            print(f"[noalloc] {n} jobs, {cells / 1e6:.0f} Mcells", flush=True)
            p_ref, r_ref = _call(ctx, "mh_dp_align_device", dA, dB, jobs)
            p_na, r_na = _call(ctx, "mh_dp_align_device_noalloc", dA, dB, jobs)
            _same(p_ref, r_ref, p_na, r_na)
    finally:
        ctx.free(dA)
        ctx.free(dB)


def test_noalloc_refuses_a_job_above_the_scratch(ctx, jobset):
    A, ao, al, B, bo, bl = jobset
    assert len(A) > 12000 and len(B) > 12000
    dA, dB = ctx.alloc(len(A) + 16), ctx.alloc(len(B) + 16)
    try:
        ctx.h2d(dA, A)
        ctx.h2d(dB, B)
        jobs = mauvehip.make_jobs(np.array([0, 0]), np.array([100, 12000]), np.array([0, 0]),
                                  np.array([90, 12000]))
        dP, dR = ctx.alloc(24100), ctx.alloc(64)
        try:
            rc = ctx.L.mh_dp_align_device_noalloc(ctx.h, dA, dB, jobs.ctypes.data, len(jobs), dP, dR)
        finally:
            ctx.free(dP)
            ctx.free(dR)
        assert rc == -1
        msg = ctx.L.mh_last_error().decode()
        assert "12000 x 12000" in msg and "64 MiB" in msg
    finally:
        ctx.free(dA)
        ctx.free(dB)
