#!/usr/bin/env python3
"""Train the band network on the GPU and write the model file.

    python mauve-py_amd/mauvehip/train_band.py [--rows 1000000] [--seed 401] [--out PATH]

Training gaps: pairs drawn like the batched-DP workload (len(A) log-uniform
in 10-20000, divergence U(0.02, 0.30), substitutions:indels 9:1, indel
length Geom(0.3)), from PCG64(seed), generated and labelled in chunks of
100k: each chunk's features (mh_gap_features_device) and full-DP labels
(largest distance of the optimal path from the corner-to-corner line, from
mh_dp_align_banded_device) stay on the device.  Then mh_mlp_train runs Adam
over all rows.  delta is calibrated by measurement: on the first 100k-gap
chunk of the training set the banded DP with its edge fallback is timed for
a grid of deltas, and among those whose results equal the full DP's (scores
and paths) and run within 5% of the fastest, the widest is kept (fewest
reruns); the best constant band (the method without the network) is chosen
the same way.  Parameters, delta and the training
record go to an .npz file.
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


def calibrate(ctx, mlp, chunk):
    """delta and the best constant band measured on one chunk: the banded DP
    with its fallback timed for every candidate, results compared with the
    full DP's; among settings with no mismatch and within 5% of the fastest,
    the widest."""
    import mauvehip
    from mauvehip import band
    A, ao, al, B, bo, bl = chunk
    cal = band.DeviceJobs(ctx, A, B, mauvehip.make_jobs(ao, al, bo, bl))
    cal.features()
    cal.align(None)
    ref_res, _ = cal.results()
    ref_paths = cal.paths()
    ycal = band.predict_output(ctx, mlp, cal)

    def measure(bands):
        cal.align(bands)                        # warm (arena, caches)
        t0 = time.perf_counter()
        st = cal.align(bands)
        el = time.perf_counter() - t0
        r, _ = cal.results()
        sm, pm = band.count_mismatches(ref_res, ref_paths, r, cal.paths())
        return el, sm + pm, st
    trials = {}
    for d in band.CALIBRATION_DELTAS:
        el, mm, st = measure(band.bands_from_output(ycal, d))
        trials[float(d)] = {"seconds": el, "mismatches": mm, "cells": int(st.cells),
                            "reruns": [int(st.rerun_doubled), int(st.rerun_full)]}
        log(f"[train] delta {d}: {el:.3f}s, mismatches {mm}, reruns {st.rerun_doubled}/{st.rerun_full}")
    # the rarer a rerun, the smaller the chance that a large gap's full-DP
    # rerun (one wavefront, a few hundred ms) lands in a batch: among settings
    # within 5% of the fastest, take the widest
    ok = [d for d, v in trials.items() if v["mismatches"] == 0] or list(trials)
    tbest = min(trials[d]["seconds"] for d in ok)
    delta = max(d for d in ok if trials[d]["seconds"] <= 1.05 * tbest)
    const = {}
    for w in band.CALIBRATION_CONSTANTS:
        el, mm, st = measure(np.full(cal.n, w, dtype=np.uint32))
        const[int(w)] = {"seconds": el, "mismatches": mm, "cells": int(st.cells)}
        log(f"[train] constant band {w}: {el:.3f}s, mismatches {mm}")
    okc = [w for w, v in const.items() if v["mismatches"] == 0] or list(const)
    tbc = min(const[w]["seconds"] for w in okc)
    wc = max(w for w in okc if const[w]["seconds"] <= 1.05 * tbc)
    cal.free()
    return delta, wc, {"delta": delta, "calibration_seconds_learned": trials[delta]["seconds"],
                       "best_constant_band": wc, "calibration_seconds_constant": const[wc]["seconds"],
                       "calibration_trials": {str(k): v for k, v in trials.items()},
                       "calibration_constants": {str(k): v for k, v in const.items()}}


def recalibrate(ctx, chunk, args):
    """--recalibrate: the model's parameters and training record kept, its
    calibration measured again."""
    import mauvehip
    from mauvehip import band
    params, _, _ = band.load_model(args.recalibrate)
    with np.load(args.recalibrate, allow_pickle=False) as z:
        info = json.loads(str(z["info_json"]))
    mlp = mauvehip.Mlp(ctx, params=params)
    delta, wc, cal = calibrate(ctx, mlp, chunk)
    info.update(cal)
    band.save_model(args.out, params, delta, {"info_json": json.dumps(info)})
    log(f"[train] recalibrated: delta {delta}: {cal['calibration_seconds_learned']:.3f}s; best constant band {wc}: "
        f"{cal['calibration_seconds_constant']:.3f}s (calibration chunk)")
    print(json.dumps(cal), flush=True)
    mlp.free()
    ctx.close()
    return 0


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
    ap.add_argument("--recalibrate", default=None,
                    help="keep this model's parameters and training record; measure delta and the best "
                         "constant band again under the current kernels (first chunk only)")
    args = ap.parse_args(argv)

    from mauvehip import synth
    import mauvehip
    from mauvehip import band

    # generate every chunk first (spawned workers), then open the device
    t0 = time.time()
    chunks = []
    ss = np.random.SeedSequence(args.seed)
    kids = ss.spawn((args.rows + args.chunk - 1) // args.chunk)
    if args.recalibrate:
        kids = kids[:1]
    for c, kid in enumerate(kids):
        n = min(args.chunk, args.rows - c * args.chunk)
        cs = int(kid.generate_state(1)[0])
        chunks.append(synth.dp_jobs(n=n, seed=cs, workers=args.workers))
        log(f"[train] generated chunk {c + 1}/{len(kids)} ({n} gaps) at {time.time() - t0:.0f}s")

    ctx = mauvehip.Context(0)
    if args.recalibrate:
        return recalibrate(ctx, chunks[0], args)
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
    resid = t - y
    # calibration by measurement on the first chunk
    delta, wc, cal = calibrate(ctx, mlp, chunks[0])
    info = {"rows": args.rows, "seed": args.seed, "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
            "epoch_loss": [float(v) for v in curve], "train_seconds": round(train_s, 2),
            "label_dp_cells": cells, **cal,
            "residual_std": float(resid.std()), "label_quantiles": {
                str(q): float(np.percentile(labels, q)) for q in (50, 90, 99, 100)}}
    band.save_model(args.out, mlp.params(), delta, {"info_json": json.dumps(info)})
    log(f"[train] delta {delta}: {cal['calibration_seconds_learned']:.3f}s; best constant band {wc}: "
        f"{cal['calibration_seconds_constant']:.3f}s (calibration chunk)")
    print(json.dumps(info), flush=True)
    for p in (X, T, Y):
        ctx.free(p)
    mlp.free()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
