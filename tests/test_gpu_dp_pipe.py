"""GPU parity of pipelined stripe jobs: a 64-lane stripe job whose stripes P
wavefronts share, handing the boundary column over through device memory.

A round with fewer work items than the forward launch's wavefront slots pipes
its stripe jobs, P a job by its size (MH_DP_PIPE sets one P for the first
jobs; 0 or 1 keeps one wavefront a job).
Bar: bit-exact paths and scores against the oracle (full DP: NWSmall; banded:
the region DP at the stripe width used), and identical output for every P.
"""
import os

import numpy as np
import pytest

import mauvehip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture
def dp_env():
    keys = ("MH_DP_PIPE", "MH_DP_COLS", "MH_DP_BAND_KIND")
    old = {k: os.environ.get(k) for k in keys}
    yield
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def _related(rng, la, lb, div=0.1):
    a = rng.integers(0, 4, la, dtype=np.uint8)
    b = a.copy() if lb <= la else np.concatenate([a, rng.integers(0, 4, lb - la, dtype=np.uint8)])
    b = b[:lb].copy()
    hit = rng.random(lb) < div
    b[hit] = rng.integers(0, 4, int(hit.sum()), dtype=np.uint8)
    return a, b


def _batch(pairs):
    A = np.concatenate([p[0] for p in pairs])
    B = np.concatenate([p[1] for p in pairs])
    la = np.array([len(p[0]) for p in pairs], dtype=np.int64)
    lb = np.array([len(p[1]) for p in pairs], dtype=np.int64)
    ao = np.concatenate([[0], np.cumsum(la)[:-1]])
    bo = np.concatenate([[0], np.cumsum(lb)[:-1]])
    return A, B, mauvehip.make_jobs(ao, la, bo, lb)


# shapes: one row, a row per chunk edge, many stripes at every width
SHAPES = [(1, 3000), (2, 1100), (63, 700), (64, 2000), (65, 1300), (500, 2600), (1500, 3100),
          (3000, 900), (700, 5000), (129, 257)]


def _pipe_parts(ctx):
    ms, n = ctx.kernel_time("dp_pipe_jobs")
    return int(round(ms * n))


@pytest.mark.parametrize("cols", ["4", "8", "16"])
def test_pipelined_full_dp_matches_oracle(ctx, dp_env, cols):
    rng = np.random.default_rng(700 + int(cols))
    pairs = [_related(rng, la, lb) for la, lb in SHAPES]
    A, B, jobs = _batch(pairs)
    os.environ["MH_DP_COLS"] = cols
    want = [po.nwsmall(a, b) for a, b in pairs]
    outs = {}
    for P in ("1", "2", "3", "5", "auto"):
        if P == "auto":
            os.environ.pop("MH_DP_PIPE", None)
        else:
            os.environ["MH_DP_PIPE"] = P
        ctx.set_timing(True)
        ctx.reset_kernel_times()
        paths, res = mauvehip.global_align(ctx, A, B, jobs)
        npiped = _pipe_parts(ctx)
        ctx.set_timing(False)
        if P == "1":
            assert npiped == 0
        elif P == "auto":
            # ten items, far fewer than the slots: the large jobs get up to eight parts
            assert npiped > 0
        else:
            # every job with at least P stripes of 64*cols columns runs pipelined
            sw = 64 * int(cols)
            assert npiped == sum((lb + sw - 1) // sw >= int(P) for _, lb in SHAPES), P
        got = [(mauvehip.path_of(paths, res, k), int(res[k]["score"])) for k in range(len(pairs))]
        for k, (p, sc) in enumerate(got):
            assert p == want[k][0] and sc == int(want[k][1]), (P, k, SHAPES[k])
        outs[P] = got
    assert all(v == outs["1"] for v in outs.values())


@pytest.mark.parametrize("cols", ["4", "16"])
def test_pipelined_banded_matches_region_oracle(ctx, dp_env, cols):
    rng = np.random.default_rng(710 + int(cols))
    shapes = SHAPES + [(4000, 3000), (2500, 4100)]
    pairs = [_related(rng, la, lb, 0.05) for la, lb in shapes]
    A, B, jobs = _batch(pairs)
    bands = rng.choice([3, 17, 64, 300], size=len(pairs)).astype(np.uint32)
    os.environ["MH_DP_COLS"] = cols
    os.environ["MH_DP_BAND_KIND"] = "stripes"
    ref = None
    for P in ("1", "2", "4", "auto"):
        if P == "auto":
            os.environ.pop("MH_DP_PIPE", None)
        else:
            os.environ["MH_DP_PIPE"] = P
        paths, res, dev, st = mauvehip.global_align_banded(ctx, A, B, jobs, bands, flags=mauvehip.BAND_NO_FALLBACK)
        got = [(mauvehip.path_of(paths, res, k), int(res[k]["score"]), int(dev[k]), int(res[k]["path_len"]))
               for k in range(len(pairs))]
        if ref is None:
            ref = got
            for k, (a, b) in enumerate(pairs):
                eff, kind, width = mauvehip.band_region(len(a), len(b), 2 * int(bands[k]))
                p, sc, dv = po.nwsmall_region(a, b, eff, width if kind == mauvehip.REGION_STRIPES else 0)
                if p is None:
                    assert got[k][3] == 0
                    continue
                assert got[k][0] == p and got[k][1] == sc and got[k][2] == dv, (k, shapes[k], int(bands[k]))
        assert got == ref, P


def test_pipelined_prefix_with_many_small_jobs(ctx, dp_env):
    """A round whose work items nearly fill the slots pipes only a prefix of
    its stripe jobs; the output equals the unpipelined one."""
    rng = np.random.default_rng(720)
    pairs = [_related(rng, la, lb) for la, lb in [(2000, 3000)] * 6]
    pairs += [_related(rng, int(x), int(x)) for x in rng.integers(300, 600, 2900)]
    A, B, jobs = _batch(pairs)
    os.environ["MH_DP_COLS"] = "4"
    res_by = {}
    for P in ("1", "2"):
        os.environ["MH_DP_PIPE"] = P
        ctx.set_timing(True)
        ctx.reset_kernel_times()
        paths, res = mauvehip.global_align(ctx, A, B, jobs)
        npiped = _pipe_parts(ctx)
        ctx.set_timing(False)
        res_by[P] = (paths.copy(), res.copy(), npiped)
    assert res_by["1"][2] == 0 and res_by["2"][2] > 0
    assert np.array_equal(res_by["1"][1], res_by["2"][1])
    assert np.array_equal(res_by["1"][0], res_by["2"][0])
    for k in range(6):
        p, sc = po.nwsmall(*pairs[k])
        assert mauvehip.path_of(*res_by["2"][:2], k) == p and int(res_by["2"][1][k]["score"]) == int(sc)
