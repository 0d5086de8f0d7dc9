"""Learned DP band: training, calibration and prediction around the C ABI.

The network (mh_mlp, 272-256-128-1, bf16 weights) predicts
y = log2(1 + label) for a gap, where the label is the full DP path's largest
distance from the corner-to-corner line, ceil(max |i*LB - j*LA| / LA).  A job
gets the band W = max(min_band, ceil(2^(y + delta))); delta is calibrated on
the training labels to minimise the modelled time of the banded run with its
edge fallback (W, then 2W, then the full matrix), from the cells of the
regions the library schedules (mh_dp_band_cells) and the single-wavefront
latency of the largest rerun.  The full DP stays the default
and the judge: every banded result whose path reaches the band edge is rerun.

Everything numeric runs on the GPU through the library: features
(mh_gap_features_device), labels (a full mh_dp_align_banded_device run),
training (mh_mlp_train: forward, backward and Adam kernels) and inference
(mh_band_predict).  This module only moves buffers and picks delta.
"""
import numpy as np

from . import (MLP_IN, Mlp, band_cells, gap_features_device, global_align_banded_device,
               DP_JOB_DTYPE, DP_RESULT_DTYPE)

MIN_BAND = 4
DELTA_GRID = np.round(np.arange(0.0, 3.01, 0.05), 3)
# grids measured by train_band.py on a calibration chunk
CALIBRATION_DELTAS = (0.6, 0.75, 0.9, 1.05, 1.2, 1.4, 1.7, 2.0)
CALIBRATION_CONSTANTS = (48, 64, 90, 117, 150, 200, 256)


def est_cells(la, lb, w):
    """DP cells of the region the library schedules for region band w."""
    w = np.broadcast_to(np.asarray(w, dtype=np.uint32), np.shape(la))
    return band_cells(la, lb, w).astype(np.float64)


# Time model of a banded run on one MI355X (seconds), from measurements of the
# DP kernels: each pass of the fallback costs its cells at the batch rate
# (~950 GCUPS of region cells) or its largest job at the rate of one
# wavefront (a job runs on one wavefront: ~1.7 GCUPS), whichever is longer;
# then the trace-back, whose longest walk takes ~0.6 us per step; plus a
# fixed cost for planning and synchronisation.
BATCH_CELLS_PER_S = 0.95e12
WAVE_CELLS_PER_S = 1.7e9
TRACE_STEP_SECONDS = 0.6e-6
PASS_SECONDS = 3e-3


def _pass_time(cells, la, lb):
    if len(cells) == 0:
        return 0.0
    walk = float((np.asarray(la, np.float64) + np.asarray(lb, np.float64)).max())
    return (max(float(cells.sum()) / BATCH_CELLS_PER_S, float(cells.max()) / WAVE_CELLS_PER_S)
            + walk * TRACE_STEP_SECONDS + PASS_SECONDS)


def expected_cost(la, lb, label, w):
    """Modelled seconds of banding at w with the edge fallback (region 2W
    accepted within W, region 4W accepted within 2W, full), given the true
    labels."""
    la = np.asarray(la)
    lb = np.asarray(lb)
    w = np.broadcast_to(np.asarray(w, dtype=np.uint32), np.shape(la)).astype(np.uint64)
    t = _pass_time(est_cells(la, lb, 2 * w), la, lb)
    miss1 = label >= w
    if miss1.any():
        t += _pass_time(est_cells(la[miss1], lb[miss1], 4 * w[miss1]), la[miss1], lb[miss1])
        miss2 = label >= 2 * w
        if miss2.any():
            t += _pass_time(la[miss2].astype(np.float64) * lb[miss2].astype(np.float64), la[miss2], lb[miss2])
    return t


def bands_from_output(y, delta, min_band=MIN_BAND):
    e = np.minimum(np.asarray(y, dtype=np.float32) + np.float32(delta), np.float32(30))
    return np.maximum(np.ceil(np.exp2(e)).astype(np.uint32), np.uint32(min_band))


CALIBRATION_BATCH = 100_000


def _batches(n, k=CALIBRATION_BATCH, seed=0):
    """Disjoint random batches of the deployment size covering the rows (the
    time model's fixed and latency terms do not scale with the batch, and
    averaging over batches weighs rare large reruns by their frequency)."""
    perm = np.random.default_rng(seed).permutation(n)
    nb = max(1, n // k)
    return [np.sort(perm[b * k:(b + 1) * k]) if n >= k else perm for b in range(nb)]


def _mean_cost(la, lb, label, w):
    la, lb, label, w = (np.asarray(v) for v in (la, lb, label, w))
    bs = _batches(len(la))
    return float(np.mean([expected_cost(la[r], lb[r], label[r], w[r]) for r in bs]))


def calibrate_delta(y, la, lb, label, min_band=MIN_BAND):
    """delta on the 0.05 grid of [0, 3] minimising expected_cost averaged over
    100k-gap batches of the training rows (coarse 0.2 search, then the 0.05
    steps around the best)."""
    def cost(d):
        return _mean_cost(la, lb, label, bands_from_output(y, d, min_band))
    coarse = {float(d): cost(d) for d in np.round(np.arange(0.0, 3.01, 0.2), 3)}
    c0 = min(coarse, key=coarse.get)
    fine = dict(coarse)
    for d in np.round(np.arange(max(0.0, c0 - 0.2), min(3.0, c0 + 0.2) + 1e-9, 0.05), 3):
        if float(d) not in fine:
            fine[float(d)] = cost(d)
    best = min(fine, key=fine.get)
    return best, fine[best]


def best_constant_band(la, lb, label, grid=None):
    """The method without the network: one band for every gap, chosen on the
    same labels by the same cost."""
    if grid is None:
        grid = np.unique(np.round(np.exp2(np.arange(2, 12, 0.125))).astype(np.uint32))
    costs = [_mean_cost(la, lb, label, np.full(len(la), w, dtype=np.uint32)) for w in grid]
    k = int(np.argmin(costs))
    return int(grid[k]), costs[k]


class DeviceJobs:
    """A job set resident on the device: operands, path/result/label buffers."""

    def __init__(self, ctx, A, B, jobs):
        self.ctx = ctx
        self.jobs = np.ascontiguousarray(jobs, dtype=DP_JOB_DTYPE)
        n = len(self.jobs)
        self.n = n
        self.la = self.jobs["a_len"].astype(np.int64)
        self.lb = self.jobs["b_len"].astype(np.int64)
        self.cells = int((self.la * self.lb).sum())
        self.dA = ctx.alloc(len(A) + 16)
        self.dB = ctx.alloc(len(B) + 16)
        ctx.h2d(self.dA, np.ascontiguousarray(A, dtype=np.uint8))
        ctx.h2d(self.dB, np.ascontiguousarray(B, dtype=np.uint8))
        self.dP = ctx.alloc(int((self.la + self.lb).sum()) + 16)
        self.dR = ctx.alloc(16 * n + 16)
        self.dD = ctx.alloc(4 * n + 16)
        self.dF = ctx.alloc(4 * MLP_IN * n + 16)

    def free(self):
        for p in (self.dA, self.dB, self.dP, self.dR, self.dD, self.dF):
            self.ctx.free(p)

    def features(self):
        gap_features_device(self.ctx, self.dA, self.dB, self.jobs, self.dF)

    def align(self, bands=None, flags=0):
        """Banded (or full, bands None) DP; returns (stats, results, deviations)."""
        st = global_align_banded_device(self.ctx, self.dA, self.dB, self.jobs, bands, self.dP, self.dR, self.dD,
                                        flags)
        self.ctx.synchronize()
        return st

    def results(self):
        res = np.zeros(self.n, dtype=DP_RESULT_DTYPE)
        dev = np.zeros(self.n, dtype=np.uint32)
        self.ctx.d2h(res, self.dR)
        self.ctx.d2h(dev, self.dD)
        return res, dev

    def paths(self):
        buf = np.zeros(int((self.la + self.lb).sum()) + 1, dtype=np.uint8)
        self.ctx.d2h(buf, self.dP)
        return buf


def train(ctx, dj, labels, epochs=6, batch=1024, lr=2e-3, weight_decay=0.0, seed=401):
    """Train a fresh network on a device job set whose features are computed
    and whose labels come from a full DP run.  Returns (mlp, delta, info)."""
    t = np.log2(1.0 + labels.astype(np.float32)).astype(np.float32)
    dT = ctx.alloc(4 * len(t) + 16)
    try:
        ctx.h2d(dT, t)
        mlp = Mlp(ctx, seed=seed)
        curve = mlp.train_device(dj.dF, dT, dj.n, epochs=epochs, batch=batch, lr=lr, weight_decay=weight_decay,
                                 seed=seed)
    finally:
        ctx.free(dT)
    y = predict_output(ctx, mlp, dj)
    delta, cost = calibrate_delta(y, dj.la, dj.lb, labels)
    info = {"epoch_loss": [float(v) for v in curve], "delta": delta, "train_expected_cells": cost,
            "train_rows": int(dj.n)}
    return mlp, delta, info


def predict_output(ctx, mlp, dj):
    dY = ctx.alloc(4 * dj.n + 16)
    try:
        mlp.forward_device(dj.dF, dj.n, dY)
        y = np.zeros(dj.n, dtype=np.float32)
        ctx.d2h(y, dY)
    finally:
        ctx.free(dY)
    return y


def save_model(path, params, delta, meta=None):
    meta = meta or {}
    np.savez(path, params=np.asarray(params, dtype=np.float32), delta=np.float32(delta),
             min_band=np.uint32(MIN_BAND), **{k: np.asarray(v) for k, v in meta.items()})


def load_model(path):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), float(z["delta"]), int(z["min_band"])


def count_mismatches(res_a, paths_a, res_b, paths_b):
    """(score mismatches, path mismatches) between two runs of one job set;
    both runs lay paths out in the same slots."""
    sm = int((res_a["score"] != res_b["score"]).sum())
    st_a = res_a["path_start"].astype(np.int64)
    ln_a = res_a["path_len"].astype(np.int64)
    same_shape = (st_a == res_b["path_start"].astype(np.int64)) & (ln_a == res_b["path_len"].astype(np.int64))
    n = len(paths_a)
    d = np.zeros(n + 1, dtype=np.int64)
    np.add.at(d, st_a, 1)
    np.add.at(d, st_a + ln_a, -1)
    inside = np.cumsum(d)[:n] > 0
    diff = inside & (paths_a[:n] != paths_b[:n])
    bad_pos = np.nonzero(diff)[0]
    owner = np.searchsorted(st_a[np.argsort(st_a)], bad_pos, side="right") - 1
    bad_jobs = set(np.argsort(st_a)[owner].tolist()) if len(bad_pos) else set()
    pm = int((~same_shape).sum()) + sum(1 for j in bad_jobs if same_shape[j])
    return sm, pm


def load_default(ctx):
    """The committed model: (Mlp, delta, min_band)."""
    import os
    params, delta, min_band = load_model(os.path.join(os.path.dirname(os.path.abspath(__file__)), "band_model.npz"))
    return Mlp(ctx, params=params), delta, min_band
