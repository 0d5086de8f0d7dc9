#!/usr/bin/env python3
"""Train the band network on the GPU and write the model file.

    python mauve-py_amd/mauvehip/train_band.py [--rows 1000000] [--seed 401] [--out PATH]

Training gaps: pairs drawn like the batched-DP workload (len(A) log-uniform
in 10-20000, divergence U(0.02, 0.30), substitutions:indels 9:1, indel
length Geom(0.3)), from PCG64(seed), generated and labelled in chunks of
100k: each chunk's features (mh_gap_features_device) and full-DP labels
(largest distance of the optimal path from the corner-to-corner line, from
mh_dp_align_banded_device) stay on the device.  Then mh_mlp_train runs Adam
over all rows, delta is calibrated on the training labels, and the
parameters, delta and training record go to an .npz file.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

DEFAULT_OUT = os.path.join(HERE, "band_model.npz")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=1_000_000)
    ap.add_argument("--chunk", type=int, default=100_000)
    ap.add_argument("--seed", type=int, default=401)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args(argv)

    from mauvehip import synth
    import mauvehip
    from mauvehip import band

    # generate every chunk first (spawned workers), then open the device
    t0 = time.time()
    chunks = []
    ss = np.random.SeedSequence(args.seed)
    kids = ss.spawn((args.rows + args.chunk - 1) // args.chunk)
    for c, kid in enumerate(kids):
        n = min(args.chunk, args.rows - c * args.chunk)
        cs = int(kid.generate_state(1)[0])
        chunks.append(synth.dp_jobs(n=n, seed=cs, workers=args.workers))
        log(f"[train] generated chunk {c + 1}/{len(kids)} ({n} gaps) at {time.time() - t0:.0f}s")

    ctx = mauvehip.Context(0)
    X = ctx.alloc(4 * mauvehip.MLP_IN * args.rows + 16)
    labels = np.zeros(args.rows, dtype=np.uint32)
    la_all = np.zeros(args.rows, dtype=np.int64)
    lb_all = np.zeros(args.rows, dtype=np.int64)
    cells = 0
    t1 = time.time()
    row = 0
    for c, (A, ao, al, B, bo, bl) in enumerate(chunks):
        dj = band.DeviceJobs(ctx, A, B, mauvehip.make_jobs(ao, al, bo, bl))
        dj.features()
        st = dj.align(None)
        _, dev = dj.results()
        n = dj.n
        # the chunk's features go into the training matrix (through the host)
        F = np.zeros((n, mauvehip.MLP_IN), dtype=np.float32)
        ctx.d2h(F, dj.dF)
        ctx.h2d(X + 4 * mauvehip.MLP_IN * row, F)
        labels[row:row + n] = dev
        la_all[row:row + n] = dj.la
        lb_all[row:row + n] = dj.lb
        cells += dj.cells
        row += n
        dj.free()
        log(f"[train] labelled chunk {c + 1}/{len(chunks)}: {st.cells / 1e12:.2f} Tcells, "
            f"{time.time() - t1:.0f}s")
    t = np.log2(1.0 + labels.astype(np.float32)).astype(np.float32)
    T = ctx.alloc(4 * args.rows + 16)
    ctx.h2d(T, t)
    t2 = time.time()
    mlp = mauvehip.Mlp(ctx, seed=args.seed)
    curve = mlp.train_device(X, T, args.rows, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    train_s = time.time() - t2
    log(f"[train] {args.epochs} epochs in {train_s:.1f}s, epoch loss {[round(float(v), 4) for v in curve]}")
    Y = ctx.alloc(4 * args.rows + 16)
    mlp.forward_device(X, args.rows, Y)
    y = np.zeros(args.rows, dtype=np.float32)
    ctx.d2h(y, Y)
    delta, cost = band.calibrate_delta(y, la_all, lb_all, labels)
    wc, cost_c = band.best_constant_band(la_all, lb_all, labels)
    resid = t - y
    info = {"rows": args.rows, "seed": args.seed, "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
            "epoch_loss": [float(v) for v in curve], "train_seconds": round(train_s, 2),
            "label_dp_cells": cells, "delta": delta, "expected_cells_learned": cost,
            "best_constant_band": wc, "expected_cells_constant": cost_c,
            "residual_std": float(resid.std()), "label_quantiles": {
                str(q): float(np.percentile(labels, q)) for q in (50, 90, 99, 100)}}
    band.save_model(args.out, mlp.params(), delta, {"info_json": json.dumps(info)})
    log(f"[train] delta {delta}, expected cells learned {cost:.3e} vs constant W={wc} {cost_c:.3e}")
    print(json.dumps(info), flush=True)
    for p in (X, T, Y):
        ctx.free(p)
    mlp.free()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
