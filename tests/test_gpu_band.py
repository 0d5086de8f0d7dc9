"""GPU parity of the banded DP and its edge fallback.

Bar: bit-exact paths, scores and path deviations.  Without fallback, every
job equals the oracle's banded region DP (or_nwsmall_region) at the stripe
width the scheduler reports (64-lane stripes of 4, 8, 16 columns a lane;
16-lane stripes of 4 or 8, four jobs a wavefront: MH_DP_COLS 36, 40); the region oracle is pinned by
tests/test_oracle_dp.py (whole matrix = NWSmall = libMUSCLE, brute force over
the region's paths, containment of the full path).  With fallback, every job
equals the oracle model of the fallback chain (region of band 2W accepted
within W, region 4W accepted within 2W, then full).
"""
import os

import numpy as np
import pytest

import mauvehip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _pairs(rng, n, la_range, div=0.1):
    A, B, jobs = [], [], []
    ao = bo = 0
    for _ in range(n):
        la = int(rng.integers(*la_range))
        a = rng.integers(0, 4, la, dtype=np.uint8)
        b = a.copy()
        hit = rng.random(la) < div
        b[hit] = rng.integers(0, 4, int(hit.sum()), dtype=np.uint8)
        if la > 4 and rng.random() < 0.7:   # an indel or two
            for _ in range(int(rng.integers(1, 3))):
                c = int(rng.integers(0, len(b)))
                if rng.random() < 0.5:
                    b = np.concatenate([b[:c], rng.integers(0, 4, int(rng.integers(1, 30)), dtype=np.uint8), b[c:]])
                else:
                    b = np.concatenate([b[:c], b[c + int(rng.integers(1, 30)):]])
        if len(b) == 0:
            b = rng.integers(0, 4, 1, dtype=np.uint8)
        A.append(a)
        B.append(b)
        jobs.append((ao, len(a), bo, len(b)))
        ao += len(a)
        bo += len(b)
    j = np.array(jobs, dtype=np.int64)
    return np.concatenate(A), np.concatenate(B), mauvehip.make_jobs(j[:, 0], j[:, 1], j[:, 2], j[:, 3])


def _oracle(a, b, band, rerun=False):
    """The oracle DP over the region the library gives this job and band (in
    the fallback's second pass with rerun)."""
    eff, kind, width = mauvehip.band_region(len(a), len(b), band, rerun)
    if kind == mauvehip.REGION_WINDOW:
        return po.nwsmall_window(a, b, eff, width), eff
    return po.nwsmall_region(a, b, eff, width if kind == mauvehip.REGION_STRIPES else 0), eff


def _ops(A, B, jobs, k):
    a = A[jobs[k]["a_off"]:jobs[k]["a_off"] + jobs[k]["a_len"]]
    b = B[jobs[k]["b_off"]:jobs[k]["b_off"] + jobs[k]["b_len"]]
    return a, b


@pytest.fixture
def cols_env():
    old = os.environ.get("MH_DP_COLS")
    yield
    if old is None:
        os.environ.pop("MH_DP_COLS", None)
    else:
        os.environ["MH_DP_COLS"] = old


@pytest.fixture
def kind_env():
    old = os.environ.get("MH_DP_BAND_KIND")
    yield
    if old is None:
        os.environ.pop("MH_DP_BAND_KIND", None)
    else:
        os.environ["MH_DP_BAND_KIND"] = old


@pytest.mark.parametrize("cols", ["4", "8", "16", "36", "40", "auto", "window"])
def test_banded_equals_region_oracle(ctx, cols, cols_env, kind_env):
    os.environ.pop("MH_DP_COLS", None)
    os.environ.pop("MH_DP_BAND_KIND", None)
    if cols == "window":
        os.environ["MH_DP_BAND_KIND"] = "window"
    elif cols != "auto":
        os.environ["MH_DP_COLS"] = cols
        os.environ["MH_DP_BAND_KIND"] = "stripes"
    rng = np.random.default_rng(int(cols) if cols.isdigit() else (99 if cols == "auto" else 98))
    A, B, jobs = _pairs(rng, 60, (1, 900))
    # skewed shapes: empty stripes on the long side, one-row and one-column jobs
    extra = [(1, 700), (700, 1), (40, 1500), (1500, 40), (1, 1), (2, 3000)]
    for la, lb in extra:
        A = np.concatenate([A, rng.integers(0, 4, la, dtype=np.uint8)])
        B = np.concatenate([B, rng.integers(0, 4, lb, dtype=np.uint8)])
    j2 = mauvehip.make_jobs(np.cumsum([0] + [e[0] for e in extra])[:-1] + int(jobs["a_off"][-1] + jobs["a_len"][-1]),
                            [e[0] for e in extra],
                            np.cumsum([0] + [e[1] for e in extra])[:-1] + int(jobs["b_off"][-1] + jobs["b_len"][-1]),
                            [e[1] for e in extra])
    jobs = np.concatenate([jobs, j2])
    bands = rng.choice([1, 2, 5, 17, 64, 300], size=len(jobs)).astype(np.uint32)
    paths, res, dev, st = mauvehip.global_align_banded(ctx, A, B, jobs, bands, flags=mauvehip.BAND_NO_FALLBACK)
    assert st.rerun_doubled == 0 and st.rerun_full == 0
    for k in range(len(jobs)):
        a, b = _ops(A, B, jobs, k)
        (p, sc, dv), _ = _oracle(a, b, 2 * int(bands[k]))   # band W evaluates the region of band 2W
        if p is None:     # the region holds no path: the library flags a failed trace-back
            assert res[k]["path_len"] == 0
            continue
        assert mauvehip.path_of(paths, res, k) == p, (k, len(a), len(b), int(bands[k]),
                                                     mauvehip.band_region(len(a), len(b), 2 * int(bands[k])))
        assert int(res[k]["score"]) == sc
        assert int(dev[k]) == dv


def test_full_mode_labels(ctx):
    rng = np.random.default_rng(5)
    A, B, jobs = _pairs(rng, 80, (1, 1200), div=0.2)
    paths, res, dev, st = mauvehip.global_align_banded(ctx, A, B, jobs, None)
    assert st.banded == 0
    gp, gr = mauvehip.global_align(ctx, A, B, jobs)
    for k in range(len(jobs)):
        p = mauvehip.path_of(paths, res, k)
        assert p == mauvehip.path_of(gp, gr, k) and res[k]["score"] == gr[k]["score"]
        assert int(dev[k]) == po.path_dev(jobs[k]["a_len"], jobs[k]["b_len"], p)


def test_fallback_chain(ctx):
    rng = np.random.default_rng(6)
    A, B, jobs = _pairs(rng, 70, (30, 1500), div=0.15)
    bands = rng.choice([1, 2, 4, 8, 30], size=len(jobs)).astype(np.uint32)
    paths, res, dev, st = mauvehip.global_align_banded(ctx, A, B, jobs, bands)
    n2 = 0
    for k in range(len(jobs)):
        a, b = _ops(A, B, jobs, k)
        la, lb, W = len(a), len(b), int(bands[k])
        # region 2W accepted within W, region 4W accepted within 2W, then full
        want, e1 = _oracle(a, b, 2 * W)
        if e1 and (want[0] is None or want[2] >= W):
            want, e2 = _oracle(a, b, 4 * W, rerun=True)
            n2 += 1
            if e2 and (want[0] is None or want[2] >= 2 * W):
                want = po.nwsmall_region(a, b, 0, 0)
        assert mauvehip.path_of(paths, res, k) == want[0], k
        assert int(res[k]["score"]) == want[1] and int(dev[k]) == want[2]
        full_p, full_s = po.nwsmall(a, b)
        assert int(res[k]["score"]) <= int(full_s)
    assert n2 > 0 and st.rerun_doubled + st.rerun_full > 0
