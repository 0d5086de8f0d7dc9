"""Learned DP band: training, calibration and prediction around the C ABI.

The network (mh_mlp, 272-256-128-1, bf16 weights) predicts
y = log2(1 + label) for a gap, where the label is the full DP path's largest
distance from the corner-to-corner line, ceil(max |i*LB - j*LA| / LA).  A job
gets the band W = max(min_band, ceil(2^(y + delta))); delta is calibrated on
the training labels to minimise the expected DP work including the edge
fallback (W, then 2W, then the full matrix).  The full DP stays the default
and the judge: every banded result whose path reaches the band edge is rerun.

Everything numeric runs on the GPU through the library: features
(mh_gap_features_device), labels (a full mh_dp_align_banded_device run),
training (mh_mlp_train: forward, backward and Adam kernels) and inference
(mh_band_predict).  This module only moves buffers and picks delta.
"""
import numpy as np

from . import (MLP_IN, Mlp, gap_features_device, global_align_banded_device,
               DP_JOB_DTYPE, DP_RESULT_DTYPE)

MIN_BAND = 4
DELTA_GRID = np.round(np.arange(0.0, 3.01, 0.05), 3)


def est_cells(la, lb, w):
    """Rough DP work of a band: the stripe region is about 2W + 256 wide."""
    la = np.asarray(la, dtype=np.float64)
    lb = np.asarray(lb, dtype=np.float64)
    return la * np.minimum(lb, 2.0 * np.asarray(w, dtype=np.float64) + 256.0)


def expected_cost(la, lb, label, w):
    """Work of banding at w with the edge fallback, given the true labels."""
    c = est_cells(la, lb, w)
    miss1 = label >= w
    c = c + np.where(miss1, est_cells(la, lb, 2 * w), 0.0)
    miss2 = label >= 2 * w
    return float((c + np.where(miss2, np.asarray(la, np.float64) * lb, 0.0)).sum())


def bands_from_output(y, delta, min_band=MIN_BAND):
    e = np.minimum(np.asarray(y, dtype=np.float32) + np.float32(delta), np.float32(30))
    return np.maximum(np.ceil(np.exp2(e)).astype(np.uint32), np.uint32(min_band))


def calibrate_delta(y, la, lb, label, min_band=MIN_BAND):
    costs = [expected_cost(la, lb, label, bands_from_output(y, d, min_band)) for d in DELTA_GRID]
    k = int(np.argmin(costs))
    return float(DELTA_GRID[k]), costs[k]


def best_constant_band(la, lb, label, grid=None):
    """The method without the network: one band for every gap, chosen on the
    same labels by the same cost."""
    if grid is None:
        grid = np.unique(np.round(np.exp2(np.arange(2, 14, 0.125))).astype(np.uint32))
    costs = [expected_cost(la, lb, label, np.full(len(la), w, dtype=np.uint32)) for w in grid]
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
