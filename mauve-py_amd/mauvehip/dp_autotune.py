"""Autotuner of the batched DP: measures, on the GPU, the job shapes per width
class of B and the forward kernel variants, and writes the table the library
reads (mauve-py_amd/csrc/dp_tune_table.h; rebuild afterwards).

    python -m mauvehip.dp_autotune --out tune.json     # on the GPU
    python -m mauvehip.dp_autotune --apply tune.json   # writes the table

Two workloads stand for the DP's callers:
  whole   the whole pairwise path on the recoded 4.6 Mb pair (seed weight
          17): its two DP batches, the profile pass (anchored ranges of about
          95 x 95) and the refinement pass (200- and 500-column windows)
  batch   the 100k pairs of the batched-DP distribution (lengths log-uniform
          10-20000, divergence 2-30%, seed 301)
Each candidate runs in a child process of its own (the library reads
MH_DP_SHAPES once per process), with the variants set through
MH_DP_VARIANT_WIDE / MH_DP_VARIANT_NARROW.  The score of a candidate is the
sum over the workloads of its DP time (forward + trace-back, HIP events)
divided by the starting table's.  The search is coordinate-wise: the shape
table first, then the wide variant, then the narrow one.  Every candidate's
results are checked equal to the starting table's (rows of the whole path;
paths and scores of the batch).
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(os.path.dirname(HERE), "csrc", "dp_tune_table.h")

# width classes of B (upper bounds, inclusive; the last class is open), as in
# dp_tune_table.h
CLASSES = [32, 64, 96, 128, 160, 192, 224, 256, 384, 512, 1024]
NCLASS = len(CLASSES) + 1

# shape tables tried (0 = four jobs a wavefront on 16-lane rows, else stripes
# of 64 lanes x that many columns)
SHAPES = {
    "rows256_s8_s16": [0, 0, 0, 0, 0, 0, 0, 0, 8, 8, 16, 16],
    "rows128_s4_s8_s16": [0, 0, 0, 0, 4, 4, 4, 4, 8, 8, 16, 16],
    "rows128_s8_s16": [0, 0, 0, 0, 8, 8, 8, 8, 8, 8, 16, 16],
    "rows256_s8to1024_s16": [0, 0, 0, 0, 0, 0, 0, 0, 8, 8, 8, 16],
    "rows256_s16": [0, 0, 0, 0, 0, 0, 0, 0, 16, 16, 16, 16],
}
# forward variants (indices into dp.hip's table): wide ones hold every shape
WIDE = {0: "16 cols, 3 waves/SIMD, one workgroup per item",
        2: "16 cols, 2 waves/SIMD, one workgroup per item",
        3: "16 cols, 3 waves/SIMD, persistent",
        4: "16 cols, 2 waves/SIMD, persistent",
        6: "16 cols, 4 waves/SIMD, one workgroup per item (the 16-column shape spills)"}
NARROW = {1: "8 cols, 4 waves/SIMD, one workgroup per item",
          5: "8 cols, 4 waves/SIMD, persistent",
          0: "16 cols, 3 waves/SIMD, one workgroup per item"}


def _batch():
    from mauvehip import synth
    return synth.dp_jobs(n=100_000, seed=301, workers=min(16, os.cpu_count() or 1))


def _whole(ctx, reps):
    """DP milliseconds of the whole path on the recoded 4.6 Mb pair per step,
    and a digest of its rows."""
    import tempfile
    import mauvehip
    from mauvehip import pipeline, synth
    ref, mut, _ = synth.pair_recoded()
    g, r = np.asarray(mut).tobytes(), np.asarray(ref).tobytes()
    d = [ctx.alloc(len(g) + 16), ctx.alloc(len(r) + 16)]
    ctx.h2d(d[0], np.frombuffer(g, dtype=np.uint8))
    ctx.h2d(d[1], np.frombuffer(r, dtype=np.uint8))
    seed = mauvehip.get_seed(17, mauvehip.CODING_SEED)
    tmpd = tempfile.mkdtemp(prefix="mauvehip_tune_")
    xmfa = os.path.join(tmpd, "t.xmfa")
    rows = pipeline.align_pair(ctx, d, (g, r), seed, xmfa, ("g", "r"))[:2]     # warm-up
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.reset_kernel_times()
    for _ in range(reps):
        rows = pipeline.align_pair(ctx, d, (g, r), seed, xmfa, ("g", "r"))[:2]
    ctx.synchronize()
    fwd_ms, fwd_n = ctx.kernel_time("dp_forward")
    tb_ms, tb_n = ctx.kernel_time("dp_traceback")
    ctx.set_timing(False)
    os.remove(xmfa)
    os.rmdir(tmpd)
    for p in d:
        ctx.free(p)
    h = hashlib.sha256(bytes(rows[0]) + b"|" + bytes(rows[1]))
    return {"ms": (fwd_ms * fwd_n + tb_ms * tb_n) / reps, "fwd_ms": fwd_ms * fwd_n / reps, "digest": h.hexdigest()}


def child(reps):
    """Runs both workloads under the environment's configuration; prints one
    JSON line: per workload the mean DP milliseconds and a digest of the
    results."""
    import mauvehip
    ctx = mauvehip.Context(0)
    out = {"whole": _whole(ctx, reps)}
    A, ao, al, B, bo, bl = _batch()
    jobs = mauvehip.make_jobs(ao, al, bo, bl)
    dA, dB = ctx.alloc(len(A) + 16), ctx.alloc(len(B) + 16)
    ctx.h2d(dA, A)
    ctx.h2d(dB, B)
    dP = ctx.alloc(int((al.astype(np.int64) + bl).sum()) + 16)
    dR = ctx.alloc(16 * len(al) + 16)
    mauvehip.global_align_device(ctx, dA, dB, jobs, dP, dR)      # warm-up: arena, graphs
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(reps):
        mauvehip.global_align_device(ctx, dA, dB, jobs, dP, dR)
    ctx.synchronize()
    wall = (time.perf_counter() - t0) / reps
    fwd_ms, fwd_n = ctx.kernel_time("dp_forward")
    tb_ms, tb_n = ctx.kernel_time("dp_traceback")
    ctx.set_timing(False)
    res = np.zeros(len(al), dtype=mauvehip.DP_RESULT_DTYPE)
    ctx.d2h(res, dR)
    paths = np.zeros(int((al.astype(np.int64) + bl).sum()) + 16, dtype=np.uint8)
    ctx.d2h(paths, dP)
    h = hashlib.sha256(res["score"].tobytes())
    for j in range(0, len(al), max(1, len(al) // 500)):
        h.update(mauvehip.path_of(paths, res, j).encode())
    cells = int((al.astype(np.int64) * bl).sum())
    out["batch"] = {"ms": (fwd_ms * fwd_n + tb_ms * tb_n) / reps, "fwd_ms": fwd_ms * fwd_n / reps,
                    "wall_ms": 1e3 * wall, "gcups_fwd": cells / (fwd_ms * fwd_n / reps * 1e-3) / 1e9,
                    "digest": h.hexdigest()}
    for p in (dA, dB, dP, dR):
        ctx.free(p)
    ctx.close()
    print(json.dumps(out), flush=True)


def run_config(shapes, wide, narrow, reps):
    env = dict(os.environ)
    env["MH_DP_SHAPES"] = ",".join(str(x) for x in shapes)
    env["MH_DP_VARIANT_WIDE"] = str(wide)
    env["MH_DP_VARIANT_NARROW"] = str(narrow)
    env.pop("MH_DP_VARIANT", None)
    env["PYTHONPATH"] = os.path.dirname(HERE) + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    p = subprocess.run([sys.executable, "-m", "mauvehip.dp_autotune", "--child", "--reps", str(reps)],
                       env=env, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        raise RuntimeError(f"tuner child failed ({p.returncode}): {p.stderr[-2000:]}")
    return json.loads(p.stdout.strip().splitlines()[-1])


# smallest relative gain a candidate needs over the best so far (repeated
# runs of one configuration differ by about 0.5%)
MIN_GAIN = 0.01


def score(r, base):
    return sum(r[w]["ms"] / base[w]["ms"] for w in base)


def write_table(shapes, wide, narrow, record, best):
    lines = [
        "// Shapes of the full DP by the width of B, and the forward kernel variants",
        "// of the DP's rounds.  Written by mauvehip/dp_autotune.py from measurements",
        "// on one MI355X: every candidate's DP milliseconds per workload (forward +",
        f"// trace-back); chosen: shapes={best[0]} wide={best[1]} narrow={best[2]} (a",
        f"// change must gain more than {100 * MIN_GAIN:.0f}% over the best before it)",
    ]
    for k, v in record.items():
        lines.append(f"//   {k}: " + ", ".join(f"{w} {x['ms']:.3f} ms (forward {x['fwd_ms']:.3f})"
                                               for w, x in v.items()))
    lines += [
        "#pragma once",
        "#include <cstdint>",
        "",
        "namespace mh {",
        "// upper bounds (inclusive) of the classes of lb; the last class is open",
        "constexpr uint32_t DP_TUNE_LB_CLASS[] = {" + ", ".join(str(c) for c in CLASSES) + "};",
        "constexpr int DP_TUNE_NCLASS = (int)(sizeof(DP_TUNE_LB_CLASS) / sizeof(DP_TUNE_LB_CLASS[0])) + 1;",
        "// per class: 0 = four jobs a wavefront on 16-lane rows (the fewest columns",
        "// that hold the job; only for lb <= 256 and a + b < 1024, else stripes of 8),",
        "// 4, 8 or 16 = one job a wavefront in stripes of 64 lanes x that many columns",
        "constexpr uint8_t DP_TUNE_SHAPE[DP_TUNE_NCLASS] = {" + ", ".join(str(x) for x in shapes) + "};",
        "// forward kernel variants (indices into the variant table of dp.hip) of",
        "// rounds with a shape wider than 8 columns a lane, and of the other rounds",
        f"constexpr int DP_TUNE_WIDE_VARIANT = {wide};",
        f"constexpr int DP_TUNE_NARROW_VARIANT = {narrow};",
        "}  // namespace mh",
        "",
    ]
    with open(TABLE, "w") as f:
        f.write("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--child", action="store_true")
    ap.add_argument("--write", action="store_true", help="rewrite dp_tune_table.h with the winner")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=None, help="JSON record of every candidate")
    ap.add_argument("--apply", default=None, help="write dp_tune_table.h from a record written by --out")
    args = ap.parse_args()
    if args.child:
        child(args.reps)
        return 0
    if args.apply:
        rec = json.load(open(args.apply))
        # the search again over the recorded candidates, in the order it ran them
        start_rec = rec["record"]["shapes=rows256_s8_s16 wide=0 narrow=1"]
        best, best_score = ["rows256_s8_s16", 0, 1], 2.0
        for k, v in rec["record"].items():
            f = dict(x.split("=") for x in k.split())
            sc = score(v, start_rec)
            if sc < best_score * (1.0 - MIN_GAIN):
                best, best_score = [f["shapes"], int(f["wide"]), int(f["narrow"])], sc
        write_table(SHAPES[best[0]], best[1], best[2], rec["record"], best)
        print(f"[tune] wrote {TABLE}: {best}", flush=True)
        return 0
    record = {}
    start = ("rows256_s8_s16", 0, 1)
    base = run_config(SHAPES[start[0]], start[1], start[2], args.reps)
    record[f"shapes={start[0]} wide={start[1]} narrow={start[2]}"] = base
    print(f"[tune] start {start}: {json.dumps({w: round(v['ms'], 3) for w, v in base.items()})}", flush=True)
    best, best_score = start, score(base, base)

    def trial(cfg):
        nonlocal best, best_score
        r = run_config(SHAPES[cfg[0]], cfg[1], cfg[2], args.reps)
        for w in base:
            if r[w]["digest"] != base[w]["digest"]:
                raise RuntimeError(f"results differ under {cfg} on {w}")
        sc = score(r, base)
        record[f"shapes={cfg[0]} wide={cfg[1]} narrow={cfg[2]}"] = r
        print(f"[tune] {cfg}: score {sc:.4f} {json.dumps({w: round(v['ms'], 3) for w, v in r.items()})}",
              flush=True)
        # a change must beat the best by more than run-to-run noise
        if sc < best_score * (1.0 - MIN_GAIN):
            best, best_score = cfg, sc

    for name in SHAPES:
        if name != start[0]:
            trial((name, best[1], best[2]))
    for w in WIDE:
        if w != best[1]:
            trial((best[0], w, best[2]))
    for n in NARROW:
        if n != best[2]:
            trial((best[0], best[1], n))
    print(f"[tune] best {best}: score {best_score:.4f}", flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"best": best, "score": best_score, "record": record}, f, indent=1)
    if args.write:
        write_table(SHAPES[best[0]], best[1], best[2], record, best)
        print(f"[tune] wrote {TABLE}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
