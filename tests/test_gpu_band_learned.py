"""The learned band end to end on the GPU: train on one gap set, band a
held-out set, and compare with the method without the network -- the plain
full DP, the default -- and, for reference, with one constant band chosen on
the same training labels by the same cost, with the same edge fallback.

Bar: zero score and path mismatches against the full DP after the edge
fallback for both bandings, and at most half of the full DP's cells (these
gaps are at most 4,000 long; the 100k-gap bench set reaches 20,000).  The
learned band is not asserted to beat the constant one: on these short gaps
the constant band chosen by the same cost evaluates about as many cells, and
on the 100k-gap set the two come out within about 10% of each other in time
(stripe regions cost more per cell as they narrow).
"""
import numpy as np
import pytest

import mauvehip
from mauvehip import band, synth

pytestmark = pytest.mark.gpu


def _set(ctx, n, seed):
    A, ao, al, B, bo, bl = synth.dp_jobs(n=n, seed=seed, min_len=10, max_ratio=400)
    return band.DeviceJobs(ctx, A, B, mauvehip.make_jobs(ao, al, bo, bl))


def test_learned_band_matches_full_dp(ctx):
    tr = _set(ctx, 8000, 411)
    ev = _set(ctx, 3000, 412)
    try:
        tr.features()
        tr.align(None)
        _, labels = tr.results()
        mlp, delta, info = band.train(ctx, tr, labels, epochs=12, batch=512, lr=2e-3, seed=411)
        assert info["epoch_loss"][-1] < info["epoch_loss"][0]
        w_const, _ = band.best_constant_band(tr.la, tr.lb, labels)

        ev.align(None)
        full_res, full_dev = ev.results()
        full_paths = ev.paths()

        def run(bands):
            st = ev.align(bands)
            res, _ = ev.results()
            paths = ev.paths()
            sm = int((res["score"] != full_res["score"]).sum())
            pm = sum(mauvehip.path_of(paths, res, k) != mauvehip.path_of(full_paths, full_res, k)
                     for k in range(ev.n))
            return st, sm, pm

        ev.features()
        learned = mlp.predict_bands_device(ev.dF, ev.n, delta)
        st_l, sm_l, pm_l = run(learned)
        st_c, sm_c, pm_c = run(np.full(ev.n, w_const, dtype=np.uint32))
        print(f"learned: cells {st_l.cells} reruns {st_l.rerun_doubled}/{st_l.rerun_full}; "
              f"constant W={w_const}: cells {st_c.cells} reruns {st_c.rerun_doubled}/{st_c.rerun_full}; "
              f"full cells {ev.cells}")
        assert sm_l == 0 and pm_l == 0
        assert sm_c == 0 and pm_c == 0
        assert st_l.cells * 2 < ev.cells
        mlp.free()
    finally:
        tr.free()
        ev.free()
