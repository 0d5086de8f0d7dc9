"""The DP round as one forward launch fed by the length-binned queue, and its
captured graphs: every kernel variant and the graph replays give the
oracle's paths and scores.

* a mixed batch (one-lane jobs, every row-packed shape, stripes of 8 and 16
  columns a lane over several stripes, long trace-back walks) against the
  oracle, launched directly (MH_DP_GRAPH=0);
* the same batch through the captured graphs, four calls in a row with other
  work on the device (another DP batch, a sorted mer list build) and host
  synchronisations between them: every call equals the direct launch;
* the mixed batch on each forward variant that holds every shape, and
  rounds of narrow shapes on every variant (MH_DP_VARIANT: one workgroup
  per item or persistent wavefronts, 2 to 6 wavefronts a SIMD);
* several rounds in one call (a small trace budget), graphs replayed per
  round.
"""
import os

import numpy as np
import pytest

import mauvehip
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _related(rng, n, d=0.1):
    a = rng.integers(0, 4, n, dtype=np.uint8)
    out = []
    for x in a:
        r = rng.random()
        if r < d * 0.9:
            out.append(int(rng.integers(0, 4)))
        elif r < d * 0.95:
            continue
        elif r < d:
            out += [int(x), int(rng.integers(0, 4))]
        else:
            out.append(int(x))
    return a, np.array(out or [1], dtype=np.uint8)


def _batch(seed, shapes):
    rng = np.random.default_rng(seed)
    pairs = []
    for la, lb in shapes:
        a, b = _related(rng, max(la, lb))
        pairs.append((a[:la] if len(a) >= la else a, b[:lb] if len(b) >= lb else b))
    a = np.concatenate([p[0] for p in pairs]).astype(np.uint8)
    b = np.concatenate([p[1] for p in pairs]).astype(np.uint8)
    la = np.array([len(p[0]) for p in pairs], dtype=np.uint32)
    lb = np.array([len(p[1]) for p in pairs], dtype=np.uint32)
    ao = np.concatenate([[0], np.cumsum(la)[:-1]]).astype(np.uint64)
    bo = np.concatenate([[0], np.cumsum(lb)[:-1]]).astype(np.uint64)
    return pairs, a, b, mauvehip.make_jobs(ao, la, bo, lb)


MIXED = ([(5, 9), (40, 12), (3, 16)] * 20                      # one lane a job
         + [(n, n) for n in range(20, 257, 12)] * 3            # every row-packed shape
         + [(600, 580), (700, 1500), (1200, 1300), (2500, 2600)]   # stripes, long walks
         + [(300, 520), (90, 700), (1100, 60)])


def _run(ctx, a, b, jobs, graph, variant=None):
    old = {k: os.environ.get(k) for k in ("MH_DP_GRAPH", "MH_DP_VARIANT")}
    os.environ["MH_DP_GRAPH"] = "1" if graph else "0"
    if variant is None:
        os.environ.pop("MH_DP_VARIANT", None)
    else:
        os.environ["MH_DP_VARIANT"] = str(variant)
    try:
        paths, res = mauvehip.global_align(ctx, a, b, jobs)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return [mauvehip.path_of(paths, res, j) for j in range(len(jobs))], res["score"].copy()


def _oracle(pairs):
    out = [po.nwsmall(pa, pb) for pa, pb in pairs]
    return [p for p, _ in out], np.array([int(s) for _, s in out])


def test_mixed_batch_direct_matches_oracle(ctx):
    pairs, a, b, jobs = _batch(1, MIXED)
    got_p, got_s = _run(ctx, a, b, jobs, graph=False)
    want_p, want_s = _oracle(pairs)
    assert got_p == want_p
    assert got_s.tolist() == want_s.tolist()


def test_graph_replays_match_direct_launch(ctx):
    pairs, a, b, jobs = _batch(2, MIXED)
    ref_p, ref_s = _run(ctx, a, b, jobs, graph=False)
    other_pairs, oa, ob, ojobs = _batch(3, [(150, 140), (800, 900), (7, 11)] * 5)
    seq = np.frombuffer(b"ACGT" * 5000, dtype=np.uint8)
    for k in range(4):
        got_p, got_s = _run(ctx, a, b, jobs, graph=True)
        assert got_p == ref_p, k
        assert got_s.tolist() == ref_s.tolist(), k
        # other device work between the replays: a direct DP batch and an SML build
        mauvehip.global_align(ctx, oa, ob, ojobs)
        sml = mauvehip.SortedMerList(ctx, seq, mauvehip.get_seed(11, 0))
        sml.free()
        ctx.synchronize()
    want_p, want_s = _oracle(pairs)
    assert ref_p == want_p and ref_s.tolist() == want_s.tolist()


@pytest.mark.parametrize("variant", [0, 2, 3, 4, 6])
def test_mixed_batch_on_every_wide_variant(ctx, variant):
    pairs, a, b, jobs = _batch(20 + variant, MIXED)
    want_p, want_s = _oracle(pairs)
    for graph in (False, True):
        got_p, got_s = _run(ctx, a, b, jobs, graph=graph, variant=variant)
        assert got_p == want_p
        assert got_s.tolist() == want_s.tolist()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6])
def test_narrow_rounds_on_every_variant(ctx, variant):
    shapes = [(n, n) for n in range(17, 129, 7)] * 4 + [(9, 14)] * 30 + [(200, 120), (400, 300)]
    pairs, a, b, jobs = _batch(4 + variant, shapes)
    for graph in (False, True):
        got_p, got_s = _run(ctx, a, b, jobs, graph=graph, variant=variant)
        want_p, want_s = _oracle(pairs)
        assert got_p == want_p
        assert got_s.tolist() == want_s.tolist()


def test_rounds_split_by_trace_budget(ctx):
    pairs, a, b, jobs = _batch(8, MIXED)
    os.environ["MH_DP_TB_BUDGET_GB"] = "0.002"
    try:
        got_p, got_s = _run(ctx, a, b, jobs, graph=True)
    finally:
        os.environ.pop("MH_DP_TB_BUDGET_GB", None)
    want_p, want_s = _oracle(pairs)
    assert got_p == want_p
    assert got_s.tolist() == want_s.tolist()
