#!/usr/bin/env python3
"""Benchmark of the seed-anchoring + inter-anchor DP path on one MI355X.

Headline (one JSON line on stdout): aligned Mb/s end-to-end on the recoded
genome pair -- 4.6 Mb reference with ORFs on both strands covering 88%, and a
derivative with 5% of codons synonymously recoded, 0.2% SNPs, 20 indels and
3 inversions -- with the seed weight-17 pattern.  One step, with both genomes
resident in HBM as ASCII:

    sorted mer list x2 (pack, keys, radix sort)   mh_sml_create_device
    pairwise seed matching (join, extension, MemHash replay)
    anchor chain -> inter-anchor DP jobs           mh_chain_gap_jobs
    NWSmall-exact DP + trace-back on the gaps      mh_dp_align_device

value = (len A + len B) / 1e6 / seconds per step.

Also reported: the batched DP on 100k pairs (lengths log-uniform 10-20000,
divergence 2-30%) as GCUPS, the roofline of the dominant kernel (HIP-event
timings inside this process), and the CPU oracle timed on one full step on
this host (1 core).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mauve-py_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_LANE_OPS = 256 * 4 * 16 * 2.4e9        # int32 lane-ops/s (256 CU x 4 SIMD x 16 lanes x 2.4 GHz)
DP_OPS_PER_CELL = 10                         # cost model of a DP cell, int32 VALU ops
DP_PEAK_GCUPS = VALU_LANE_OPS / DP_OPS_PER_CELL / 1e9
MAX_GAP = 5000                               # max_gap_length, LM/ProgressiveAligner.cpp:62
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and
# --pmc WRITE_SIZE on the recoded 4.6 Mb pair, scaled by the counters' reading
# on a plain 1 GiB copy kernel (FETCH_SIZE x2, WRITE_SIZE x1): the
# sorted-mer-list sort (one histogram kernel, then 4 onesweep passes of 9-bit
# digits over 4,599,984 keys), the join + solid-seed extension of the pair, and the
# whole path's NWSmall forward kernel (mean of the profile-pass and the
# refinement-pass launches)
MEASURED_TRAFFIC = {"sml_sort": 628711488, "extend": 782747872, "dp_forward_whole": 453269683}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dp-pairs", type=int, default=100_000, help="pairs in the batched DP leg (0 = skip)")
    ap.add_argument("--band-pairs", type=int, default=100_000,
                    help="held-out gaps in the learned-band leg (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--band-only", action="store_true",
                    help="run only the learned-band leg and print its record (for sweeps)")
    ap.add_argument("--dp-only", action="store_true",
                    help="run only the batched-DP leg and print its record (for profiling its kernels)")
    ap.add_argument("--whole-only", action="store_true",
                    help="run only the whole-path leg (for profiling its kernels on their own)")
    ap.add_argument("--unique-only", action="store_true",
                    help="run only the eight-genome UniqueMatchFinder leg and print its record (for profiling)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding of the path is design-only "
              "(see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    from mauvehip import synth
    if args.band_only:
        import mauvehip
        bandset = synth.dp_jobs(n=args.band_pairs, seed=402, workers=min(16, os.cpu_count() or 1))
        ctx = mauvehip.Context(0)
        print(json.dumps({"learned_band": learned_band_leg(ctx, bandset)}), flush=True)
        ctx.close()
        return 0
    if args.unique_only:
        import mauvehip
        ctx = mauvehip.Context(0)
        print(json.dumps({"multi_genome_unique": unique_leg(ctx, args.steps, args.warmup, not args.no_cpu_baseline)}),
              flush=True)
        ctx.close()
        return 0
    if args.dp_only:
        import mauvehip
        dp = synth.dp_jobs(n=args.dp_pairs, seed=301, workers=min(16, os.cpu_count() or 1))
        ctx = mauvehip.Context(0)
        print(json.dumps({"dp_batch": dp_batch_leg(ctx, dp, not args.no_cpu_baseline)}), flush=True)
        ctx.close()
        return 0
    t0 = time.time()
    ref, mut, _ = synth.pair_recoded()
    log(f"[bench] recoded pair {len(ref)} / {len(mut)} bp generated in {time.time() - t0:.1f}s")
    dp = None
    if args.dp_pairs > 0:
        t0 = time.time()
        dp = synth.dp_jobs(n=args.dp_pairs, seed=301, workers=min(16, os.cpu_count() or 1))
        log(f"[bench] {args.dp_pairs} DP pairs generated in {time.time() - t0:.1f}s")

    bandset = None
    if args.band_pairs > 0:
        t0 = time.time()
        bandset = synth.dp_jobs(n=args.band_pairs, seed=402, workers=min(16, os.cpu_count() or 1))
        log(f"[bench] {args.band_pairs} held-out band gaps generated in {time.time() - t0:.1f}s")

    import mauvehip
    ctx = mauvehip.Context(0)
    seed = mauvehip.get_seed(17, mauvehip.CODING_SEED)
    La, Lb = len(ref), len(mut)
    dA = ctx.alloc(La + 16)
    dB = ctx.alloc(Lb + 16)
    ctx.h2d(dA, ref)
    ctx.h2d(dB, mut)
    cA = ctx.alloc(La + 16)
    cB = ctx.alloc(Lb + 16)
    mauvehip.ascii_to_codes_device(ctx, dA, La, cA)
    mauvehip.ascii_to_codes_device(ctx, dB, Lb, cB)
    d_paths = ctx.alloc(2 * (La + Lb) + 16)
    d_res = ctx.alloc(16 * (La + Lb) + 16)
    state = {}
    stage_s = {"sml": 0.0, "match": 0.0, "chain": 0.0, "dp": 0.0, "free": 0.0}

    if args.whole_only:
        args.dp_pairs = args.band_pairs = 0
        dp = bandset = None

    def step():
        t = [time.perf_counter()]
        sa = mauvehip.SortedMerList.from_device(ctx, dA, La, seed)
        sb = mauvehip.SortedMerList.from_device(ctx, dB, Lb, seed)
        t.append(time.perf_counter())
        mt, stats = mauvehip.find_matches_raw(ctx, [sa, sb])
        t.append(time.perf_counter())
        chain, jobs = mauvehip.chain_gap_jobs_raw(mt, La, Lb, MAX_GAP)
        t.append(time.perf_counter())
        mauvehip.global_align_device(ctx, cA, cB, jobs, d_paths, d_res)
        ctx.synchronize()
        t.append(time.perf_counter())
        sa.free()
        sb.free()
        t.append(time.perf_counter())
        for k, name in enumerate(("sml", "match", "chain", "dp", "free")):
            stage_s[name] += t[k + 1] - t[k]
        state.update(stats=stats, jobs=jobs, chain=chain, matches=mt)

    if args.whole_only:
        out = {"metric": "aligned Mb/s end-to-end", "unit": "Mb/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
               "dtype": "int32", "data": "synthetic", "config": {}}
    else:
        for _ in range(args.warmup):
            step()
        ctx.set_timing(True)
        ctx.reset_kernel_times()
        ctx.synchronize()
        for k in stage_s:
            stage_s[k] = 0.0
        t0 = time.perf_counter()
        for k in range(args.steps):
            step()
            log(f"[bench] step {k + 1}/{args.steps} done")
        ctx.synchronize()
        el = time.perf_counter() - t0
        ms_step = 1e3 * el / args.steps
        value = (La + Lb) / 1e6 / (el / args.steps)
        jobs = state["jobs"]
        cells = int((jobs["a_len"].astype(np.int64) * jobs["b_len"]).sum())

        # per-kernel device time over the timed region (HIP events on the library stream)
        N = La - mauvehip.get_seed_length(seed) + 1
        w = mauvehip.get_seed_weight(seed)
        passes = (2 * w + 1 + 9) // 10        # digits of up to 10 bits, the fewest passes
        kern = {}
        for name in ("sml_pack", "sml_keys", "sml_sort", "join", "join_write", "extend", "partition", "dp_forward",
                     "dp_traceback", "host_wait_seeds", "host_partition", "host_replay", "host_list", "host_dp_plan", "host_dp_sort"):
            ms, n = ctx.kernel_time(name)
            kern[name] = {"avg_ms": ms, "launches": n, "ms_per_step": ms * n / args.steps}
        nseeds = int(state["stats"].seed_tuples)
        # algorithmic work per launch
        sort_bytes = 24 * N * passes + 8 * N                      # per genome: passes + one histogram read
        k3_bytes = 2 * 12 * N + 8 * nseeds + 2 * La // 4          # join + extension, per pair
        lines = {
            "sml_sort": ("hbm", sort_bytes / (kern["sml_sort"]["avg_ms"] * 1e-3) / 1e9 if kern["sml_sort"]["avg_ms"] else 0,
                         HBM_PEAK_GBS, "GB/s", sort_bytes),
            "extend": ("hbm", k3_bytes / ((kern["join"]["avg_ms"] + kern["join_write"]["avg_ms"]
                                           + kern["extend"]["avg_ms"]) * 1e-3) / 1e9
                       if kern["extend"]["avg_ms"] else 0, HBM_PEAK_GBS, "GB/s", k3_bytes),
            "dp_forward": ("valu", cells / (kern["dp_forward"]["ms_per_step"] * 1e-3) / 1e9
                           if kern["dp_forward"]["ms_per_step"] else 0, DP_PEAK_GCUPS, "GCUPS", cells),
        }
        dom = max(("sml_sort", "extend", "dp_forward"), key=lambda k: kern[k]["ms_per_step"])
        b, ach, peak, unit, work = lines[dom]
        roofline = {"kernel": dom, "bound": b, "achieved": round(ach, 2), "peak": round(peak, 1), "unit": unit,
                    "frac": round(ach / peak, 4) if peak else None,
                    "traffic": MEASURED_TRAFFIC.get(dom) if (La, Lb) == (4600000, 4600339) else None,
                    "work_per_launch": work}
        log(f"[bench] headline {value:.2f} Mb/s, {ms_step:.1f} ms/step; kernels {json.dumps(kern)}")

        out = {
            "metric": "aligned Mb/s end-to-end", "value": round(value, 3), "unit": "Mb/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
            "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic",
            "config": {"workload": "recoded genome pair, 4.6 Mb, seed weight 17: sorted mer lists x2 + pairwise "
                                   "matching + anchor-chain gap DP",
                       "genome_bp": [La, Lb], "seed": hex(seed), "seed_tuples": nseeds,
                       "matches": int(state["stats"].matches), "chain_anchors": int(len(state["chain"])),
                       "dp_jobs": int(len(jobs)), "dp_cells": cells, "max_gap": MAX_GAP},
            "roofline": roofline,
            "rooflines": {k: {"bound": v[0], "achieved": round(v[1], 2), "peak": round(v[2], 1), "unit": v[3],
                              "frac": round(v[1] / v[2], 4) if v[2] else None, "work_per_launch": v[4],
                              "traffic": MEASURED_TRAFFIC.get(k) if (La, Lb) == (4600000, 4600339) else None}
                          for k, v in lines.items()},
            "kernels": {k: {"avg_ms": round(v["avg_ms"], 4), "launches_per_step": v["launches"] / args.steps}
                        for k, v in kern.items()},
            "stages_ms_per_step": {k: round(1e3 * v / args.steps, 3) for k, v in stage_s.items()},
        }

    if dp is not None:
        out["dp_batch"] = dp_batch_leg(ctx, dp, not args.no_cpu_baseline)

    if bandset is not None:
        out["learned_band"] = learned_band_leg(ctx, bandset)
        out["learned_band"]["c2_gaps"] = c2_gap_band_leg(ctx, ref, mut, state["jobs"])

    if not args.whole_only:
        out["multi_genome"] = multi_leg(ctx, ref, seed, args.steps, args.warmup, not args.no_cpu_baseline)
        out["multi_genome_unique"] = unique_leg(ctx, args.steps, args.warmup, not args.no_cpu_baseline)

    # the whole path (stand-in LCB chain, recursive anchoring, refinement
    # passes, XMFA and index LUT) is the headline
    full = full_leg(ctx, mut, ref, seed, args.steps, args.warmup)
    if not args.whole_only:
        out["core"] = {"value": out["value"], "ms_per_step": out["ms_per_step"],
                       "workload": out["config"]["workload"], "roofline": out["roofline"],
                       "rooflines": out.pop("rooflines"),
                       "stages_ms_per_step": out.pop("stages_ms_per_step"), "dp_jobs": out["config"]["dp_jobs"],
                       "dp_cells": out["config"]["dp_cells"]}
    out["value"] = full["value"]
    out["ms_per_step"] = full["ms_per_step"]
    out["roofline"] = full["roofline"]
    out["stages_ms_per_step"] = full["stages_ms_per_step"]
    out["config"] = {"workload": "recoded genome pair, 4.6 Mb, seed weight 17, genome 1 = recoded derivative: "
                                 "sorted mer lists x2, pairwise matching, anchor chain (stand-in for the LCB pass), "
                                 "recursive anchoring of gaps > 200, profile and refinement passes (batched DP), "
                                 "XMFA file, buildIndex LUT",
                     "genome_bp": [len(mut), len(ref)], "seed": hex(seed), **full["config"]}
    out["kernels"] = full["kernels"]

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_full(mut, ref, seed, full)
        log(f"[bench] CPU oracle (whole path, one thread) {out['cpu_baseline']['seconds']:.2f}s per step")
    if not args.no_cpu_baseline and not args.whole_only:
        from oracle import pyoracle as po
        t0 = time.perf_counter()
        st_o, ln_o, _ = po.find_matches([ref.tobytes(), mut.tobytes()], seed)
        codes_a, codes_b = synth.ascii_to_codes(ref), synth.ascii_to_codes(mut)
        for j in jobs:
            po.nwsmall(codes_a[j["a_off"]:j["a_off"] + j["a_len"]], codes_b[j["b_off"]:j["b_off"] + j["b_len"]])
        cpu_s = time.perf_counter() - t0
        mt = state["matches"]
        same = (len(ln_o) == len(mt) and np.array_equal(ln_o, mt["length"])
                and np.array_equal(st_o, mt["start"][:, :2]))
        out["core"]["cpu_baseline"] = {"value": round((La + Lb) / 1e6 / cpu_s, 4), "unit": "Mb/s", "cores": 1,
                               "kind": "port", "seconds": round(cpu_s, 2),
                               "sample": "one full step of the same workload (oracle sorted mer lists, "
                                         "match finding and NWSmall over the same gap jobs), single thread",
                               "matches_identical": bool(same)}
        log(f"[bench] CPU oracle step {cpu_s:.2f}s")
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def multi_leg(ctx, ref, seed, steps, warmup, cpu):
    """Seed matching over four genomes (PairwiseMatchFinder): the first 1 Mb of
    the reference and three relatives of it (1% SNPs, a 60 kb inversion and a
    short deletion each), resident in HBM; a step builds the four sorted mer
    lists and finds all pairwise matches.  The oracle times the same search on
    one core and its list is compared."""
    import mauvehip
    from mauvehip import synth
    rng = np.random.default_rng(77)
    base = np.asarray(ref[:1_000_000])
    seqs = [base]
    for _ in range(3):
        m = base.copy()
        hit = rng.random(len(m)) < 0.01
        m[hit] = synth.BASES[(np.searchsorted(synth.BASES, m[hit]) + rng.integers(1, 4, hit.sum())) % 4]
        at = int(rng.integers(0, len(m) - 60_000))
        m[at:at + 60_000] = synth.revcomp(m[at:at + 60_000])
        cut = int(rng.integers(0, len(m) - 50))
        seqs.append(np.concatenate([m[:cut], m[cut + int(rng.integers(1, 30)):]]))
    dev = []
    for q in seqs:
        d = ctx.alloc(len(q) + 16)
        ctx.h2d(d, q)
        dev.append(d)
    total = sum(len(q) for q in seqs)

    def step():
        smls = [mauvehip.SortedMerList.from_device(ctx, d, len(q), seed) for d, q in zip(dev, seqs)]
        out = mauvehip.find_matches_raw(ctx, smls)
        for x in smls:
            x.free()
        return out

    for _ in range(max(1, warmup)):
        mt, st = step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        mt, st = step()
    ctx.synchronize()
    el = (time.perf_counter() - t0) / steps
    res = {"genomes": 4, "genome_bp": [len(q) for q in seqs], "seed": hex(seed),
           "value": round(total / 1e6 / el, 3), "unit": "Mb/s", "ms_per_step": round(1e3 * el, 3),
           "seed_tuples": int(st.seed_tuples), "matches": int(st.matches), "collisions": int(st.collisions)}
    log(f"[bench] four-genome matching {res['value']:.1f} Mb/s, {res['ms_per_step']:.1f} ms/step")
    if cpu:
        from oracle import pyoracle as po
        t0 = time.perf_counter()
        os_, ol, _ = po.find_matches([q.tobytes() for q in seqs], seed)
        cs = time.perf_counter() - t0
        same = (len(ol) == len(mt) and np.array_equal(ol, mt["length"])
                and np.array_equal(os_, mt["start"][:, :4]))
        res["cpu_baseline"] = {"value": round(total / 1e6 / cs, 4), "unit": "Mb/s", "cores": 1, "kind": "port",
                               "seconds": round(cs, 2), "sample": "the same four-genome search (oracle sorted mer "
                               "lists and PairwiseMatchFinder), single thread", "matches_identical": bool(same)}
        log(f"[bench] CPU oracle four-genome matching {cs:.2f}s")
    for d in dev:
        ctx.free(d)
    return res


UNIQUE_CPU_LENGTH = 2_000_000


def pair_gap_jobs(starts, lengths, lens, b_base):
    """The inter-anchor DP jobs of every genome pair (0, g) of a k-genome
    match list: the matches holding genomes 0 and g, chained by the stand-in
    chain (mh_chain_gap_jobs, gaps of 1..MAX_GAP on both sides); B offsets
    shifted to genome g's place in the concatenated B codes.  The pairs chain
    on host threads at once (the library call releases the interpreter)."""
    import mauvehip
    from concurrent.futures import ThreadPoolExecutor

    def pair(g):
        m = (starts[:, 0] != 0) & (starts[:, g] != 0)
        _, jobs = mauvehip.chain_gap_jobs(np.ascontiguousarray(starts[m][:, [0, g]]), lengths[m], lens[0], lens[g],
                                          MAX_GAP)
        jobs = jobs.copy()
        jobs["b_off"] += np.uint64(b_base[g])
        return jobs

    with ThreadPoolExecutor(max_workers=starts.shape[1] - 1) as ex:
        return np.concatenate(list(ex.map(pair, range(1, starts.shape[1]))))


def unique_leg(ctx, steps, warmup, cpu):
    """Eight 5 Mb genomes (45% GC) derived from one ancestor along a random
    binary tree (1-5% SNPs, 0.2% indels, two inversions and a 2-20 kb deletion
    per branch).  A step builds the eight sorted mer lists, finds the matches
    with UniqueMatchFinder (the finder progressiveMauve runs for more than
    four genomes) under the default seed of weight 19 (rank 0, spaced), chains
    each genome pair (0, g) with the stand-in chain and runs the DP of every
    pair's inter-anchor gaps as one batch.  Genomes resident in HBM.  The CPU
    baseline is the oracle's matching on the same generator at
    UNIQUE_CPU_LENGTH bases, whose list is compared with the GPU's; a sample
    of the DP jobs is checked against the oracle's NWSmall."""
    import mauvehip
    from mauvehip import synth
    fam = synth.genome_family()
    seed = mauvehip.get_seed(19, 0)
    total = sum(len(q) for q in fam)

    def run(seqs, dev):
        smls = [mauvehip.SortedMerList.from_device(ctx, d, len(q), seed) for d, q in zip(dev, seqs)]
        try:
            return mauvehip.find_matches_unique(ctx, smls)
        finally:
            for x in smls:
                x.free()

    def upload(seqs):
        dev = []
        for q in seqs:
            d = ctx.alloc(len(q) + 16)
            ctx.h2d(d, q)
            dev.append(d)
        return dev

    dev = upload(fam)
    # DP operands: genome 0's codes as A, genomes 1..7 concatenated as B
    lens = [len(q) for q in fam]
    b_base = [0] + list(np.cumsum([0] + lens[1:-1]))
    cA = ctx.alloc(lens[0] + 16)
    cB = ctx.alloc(sum(lens[1:]) + 16)
    mauvehip.ascii_to_codes_device(ctx, dev[0], lens[0], cA)
    for g in range(1, len(fam)):
        mauvehip.ascii_to_codes_device(ctx, dev[g], lens[g], cB + int(b_base[g]))
    st_, ln_, stats = run(fam, dev)
    jobs = pair_gap_jobs(st_, ln_, lens, b_base)
    d_paths = ctx.alloc(2 * int((jobs["a_len"].astype(np.int64) + jobs["b_len"]).sum()) + 16)
    d_res = ctx.alloc(16 * 2 * len(jobs) + 16)
    stage = {"match": 0.0, "chain": 0.0, "dp": 0.0}

    def step():
        t = [time.perf_counter()]
        st_, ln_, stats = run(fam, dev)
        t.append(time.perf_counter())
        jobs = pair_gap_jobs(st_, ln_, lens, b_base)
        t.append(time.perf_counter())
        mauvehip.global_align_device(ctx, cA, cB, jobs, d_paths, d_res)
        ctx.synchronize()
        t.append(time.perf_counter())
        for k, name in enumerate(("match", "chain", "dp")):
            stage[name] += t[k + 1] - t[k]
        return st_, ln_, stats, jobs

    for _ in range(max(1, warmup)):
        step()
    for k in stage:
        stage[k] = 0.0
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        st_, ln_, stats, jobs = step()
        log(f"[bench] eight-genome step {k + 1}/{steps} done")
    ctx.synchronize()
    el = (time.perf_counter() - t0) / steps
    cells = int((jobs["a_len"].astype(np.int64) * jobs["b_len"]).sum())
    # a seeded sample of the DP jobs against the oracle's NWSmall
    from oracle import pyoracle as po
    res_dp = np.zeros(len(jobs), dtype=mauvehip.DP_RESULT_DTYPE)
    ctx.d2h(res_dp, d_res)
    codes = [np.frombuffer(q, dtype=np.uint8) if isinstance(q, bytes) else np.asarray(q) for q in fam]
    lut = np.full(256, 4, dtype=np.uint8)
    for c, v in zip(b"ACGTacgt", (0, 1, 2, 3, 0, 1, 2, 3)):
        lut[c] = v
    A0 = lut[codes[0]]
    Bc = lut[np.concatenate(codes[1:])]
    small = np.nonzero(jobs["a_len"].astype(np.int64) * jobs["b_len"] <= 1_000_000)[0]
    pick = np.random.default_rng(5).choice(small, size=min(300, len(small)), replace=False)
    same = all(int(po.nwsmall(A0[int(jobs["a_off"][j]):int(jobs["a_off"][j] + jobs["a_len"][j])],
                              Bc[int(jobs["b_off"][j]):int(jobs["b_off"][j] + jobs["b_len"][j])])[1])
               == int(res_dp["score"][j]) for j in pick)
    for d in dev + [cA, cB, d_paths, d_res]:
        ctx.free(d)
    res = {"genomes": len(fam), "genome_bp": lens, "seed": hex(seed),
           "value": round(total / 1e6 / el, 3), "unit": "Mb/s", "ms_per_step": round(1e3 * el, 3),
           "stages_ms_per_step": {k: round(1e3 * v / steps, 2) for k, v in stage.items()},
           "seed_tuples": int(stats.seed_tuples), "matches": int(stats.matches),
           "collisions": int(stats.collisions), "repeat_runs": int(stats.repeat_runs),
           "dp": {"pairs": len(fam) - 1, "jobs": int(len(jobs)), "cells": cells,
                  "gcups": round(cells / (stage["dp"] / steps) / 1e9, 1) if stage["dp"] else None,
                  "oracle_sample_jobs": int(len(pick)), "oracle_sample_scores_identical": bool(same)}}
    log(f"[bench] eight genomes: {res['value']:.1f} Mb/s, {res['ms_per_step']:.1f} ms/step "
        f"({json.dumps(res['stages_ms_per_step'])}), DP {len(jobs)} jobs {cells / 1e9:.2f} Gcells")
    if cpu:
        from oracle import pyoracle as po
        small = synth.genome_family(length=UNIQUE_CPU_LENGTH)
        dev = upload(small)
        gs, gl, _ = run(small, dev)
        for d in dev:
            ctx.free(d)
        t0 = time.perf_counter()
        os_, ol, _ = po.find_matches_unique([q.tobytes() for q in small], seed)
        cs = time.perf_counter() - t0
        stot = sum(len(q) for q in small)
        res["cpu_baseline"] = {"value": round(stot / 1e6 / cs, 4), "unit": "Mb/s", "cores": 1, "kind": "port",
                               "seconds": round(cs, 2),
                               "sample": f"the same generator with a {UNIQUE_CPU_LENGTH // 1_000_000} Mb ancestor "
                                         f"({stot / 1e6:.1f} Mb in all), oracle sorted mer lists and "
                                         f"UniqueMatchFinder, single thread",
                               "matches_identical": bool(np.array_equal(gs, os_) and np.array_equal(gl, ol))}
        log(f"[bench] CPU oracle unique matching {cs:.2f}s")
    return res


def full_leg(ctx, g, r, seed, steps, warmup):
    """The whole pairwise path (mauvehip.pipeline.align_pair) on genome 1 =
    g, genome 2 = r, both resident in HBM; XMFA written to a temporary file."""
    import tempfile
    import mauvehip
    from mauvehip import pipeline
    g, r = np.asarray(g).tobytes(), np.asarray(r).tobytes()
    d = [ctx.alloc(len(g) + 16), ctx.alloc(len(r) + 16)]
    ctx.h2d(d[0], np.frombuffer(g, dtype=np.uint8))
    ctx.h2d(d[1], np.frombuffer(r, dtype=np.uint8))
    tmpd = tempfile.mkdtemp(prefix="mauvehip_bench_")
    xmfa = os.path.join(tmpd, "pair.xmfa")
    names = ["recoded.fa", "reference.fa"]
    for _ in range(max(1, warmup)):
        pipeline.align_pair(ctx, d, (g, r), seed, xmfa, names)
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.reset_kernel_times()
    stages = {}
    rs = mauvehip.RefineStats()
    t0 = time.perf_counter()
    for k in range(steps):
        rs.profile_cells = rs.window_cells = 0
        ra, rb, lut, info = pipeline.align_pair(ctx, d, (g, r), seed, xmfa, names, timings=stages, refine_stats=rs)
        log(f"[bench] whole-path step {k + 1}/{steps} done")
    ctx.synchronize()
    el = (time.perf_counter() - t0) / steps
    cells = int(rs.profile_cells + rs.window_cells)
    kern = {}
    raw = {}
    for name in ("sml_pack", "sml_keys", "sml_sort", "join", "join_write", "extend", "partition", "dp_forward",
                 "dp_traceback", "host_replay", "host_dp_plan", "host_refine_prefix", "host_refine_pieces", "host_refine_plan",
                 "host_refine_dp", "host_refine_splice"):
        ms, n = ctx.kernel_time(name)
        raw[name] = ms * n / steps
        kern[name] = {"avg_ms": round(ms, 4), "launches_per_step": n / steps, "ms_per_step": round(ms * n / steps, 3)}
    fwd = raw["dp_forward"]
    gk = cells / (fwd * 1e-3) / 1e9 if fwd else 0.0
    roofline = {"kernel": "dp_forward", "bound": "valu", "achieved": round(gk, 2), "peak": round(DP_PEAK_GCUPS, 1),
                "unit": "GCUPS", "frac": round(gk / DP_PEAK_GCUPS, 4),
                "traffic": MEASURED_TRAFFIC["dp_forward_whole"] if (len(g), len(r)) == (4600339, 4600000) else None,
                "work_per_launch": cells // max(1, int(kern["dp_forward"]["launches_per_step"]))}
    value = (len(g) + len(r)) / 1e6 / el
    log(f"[bench] whole path {value:.2f} Mb/s, {1e3 * el:.1f} ms/step; stages "
        f"{ {k: round(1e3 * v / steps, 2) for k, v in stages.items()} }")
    # inputs of the sampled CPU baseline
    t = {"value": round(value, 3), "ms_per_step": round(1e3 * el, 3), "roofline": roofline, "kernels": kern,
         "stages_ms_per_step": {k: round(1e3 * v / steps, 3) for k, v in stages.items()},
         "config": {**info, "profile_dp_cells": int(rs.profile_cells), "window_dp_cells": int(rs.window_cells),
                    "lut_mapped": round(float((lut >= 0).mean()), 4), "xmfa_bytes": os.path.getsize(xmfa)},
         "rows": (ra, rb), "lut": lut}
    for p in d:
        ctx.free(p)
    os.remove(xmfa)
    os.rmdir(tmpd)
    return t


def cpu_full(g, r, seed, full):
    """The whole path on this host's CPU, one thread, every stage over the
    whole pair (no sampling, no scaling): the oracle's sorted mer lists and
    match finding, the library's host chain (the stand-in for the LCB pass,
    host code in both paths), the oracle's recursive anchor search, then the
    compiled restatement of the profile and refinement passes, the XMFA file
    and the buildIndex LUT (oracle/cpu_path.cpp, -O3).  The rows and the LUT
    are compared with the GPU path's."""
    import tempfile
    import mauvehip
    from oracle import pyoracle as po
    g, r = np.asarray(g).tobytes(), np.asarray(r).tobytes()
    t = {}
    t0 = time.perf_counter()
    st, ln, _ = po.find_matches([g, r], seed)
    t["match"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    m = np.zeros(len(ln), dtype=mauvehip.MATCH_DTYPE)
    m["start"][:, :2] = st
    m["length"] = ln
    chain, _ = mauvehip.chain_gap_jobs_raw(m, len(g), len(r), 0)
    ra, rb = mauvehip.chain_rows(m, chain, g, r)
    t["chain"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    extra = []
    for lend, rend in po.c_recursion_gaps(ra, rb, 1, 1, 200):
        gs, gl = po.anchor_gap(g, r, lend, rend)
        if len(gl):
            e = np.zeros(len(gl), dtype=mauvehip.MATCH_DTYPE)
            e["start"][:, :2] = gs
            e["length"] = gl
            extra.append(e)
    if extra:
        m = np.concatenate([m] + extra)
        chain, _ = mauvehip.chain_gap_jobs_raw(m, len(g), len(r), 0)
        ra, rb = mauvehip.chain_rows(m, chain, g, r)
    t["recurse"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    a2, b2, cells = po.c_refine_gapped(ra, rb)
    t["refine"] = time.perf_counter() - t0
    tmpd = tempfile.mkdtemp(prefix="mauvehip_cpu_")
    xmfa = os.path.join(tmpd, "cpu.xmfa")
    t0 = time.perf_counter()
    lut, nbytes = po.c_xmfa_lut(xmfa, ("genome1.fas", "genome2.fas"), a2, b2, len(g), len(r), g, r)
    t["xmfa_lut"] = time.perf_counter() - t0
    os.remove(xmfa)
    os.rmdir(tmpd)
    total = sum(t.values())
    gr = full.get("rows")
    glut = full.get("lut")
    return {"value": round((len(g) + len(r)) / 1e6 / total, 4), "unit": "Mb/s", "cores": 1, "kind": "port",
            "seconds": round(total, 2), "stages_seconds": {k: round(v, 3) for k, v in t.items()},
            "dp_cells": int(cells), "xmfa_bytes": int(nbytes),
            "rows_identical": bool(gr is not None and bytes(gr[0]) == a2.encode() and bytes(gr[1]) == b2.encode()),
            "lut_identical": bool(glut is not None and np.array_equal(np.asarray(glut), lut)),
            "sample": "the whole step on one thread, every stage over the whole pair: oracle sorted mer lists + "
                      "matching, host chain, oracle recursive anchor search, compiled restatement of the "
                      "profile and refinement passes, XMFA file and LUT over every column"}


CPU_DP_SECONDS = 12.0


def cpu_dp_batch(dp, gpu_scores):
    """The oracle's NWSmall (-O3, one job per call of or_nwsmall) on this
    host's CPU over a sample of the batch: jobs in a seeded random order,
    taken until about CPU_DP_SECONDS of work, on one thread and then on up to
    16 threads (ctypes releases the GIL; one job per task, largest first).
    The sample's scores are compared with the GPU's."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import pyoracle as po
    A, ao, al, B, bo, bl = dp
    order = np.random.default_rng(5).permutation(len(al))
    pa, pl = np.ascontiguousarray(ao[order], dtype=np.uint64), np.ascontiguousarray(al[order], dtype=np.uint32)
    pb, ql = np.ascontiguousarray(bo[order], dtype=np.uint64), np.ascontiguousarray(bl[order], dtype=np.uint32)
    cum = np.cumsum(pl.astype(np.int64) * ql)
    A = np.ascontiguousarray(A, dtype=np.uint8)
    B = np.ascontiguousarray(B, dtype=np.uint8)
    scores = np.zeros(len(al), dtype=np.float32)

    def run(lo, hi):
        return po.c_nwsmall_batch(A, B, pa, pl, pb, ql, lo, hi, scores)

    # rate probe on the first jobs, then one thread for about CPU_DP_SECONDS
    t0 = time.perf_counter()
    n0 = int(np.searchsorted(cum, 2e8)) + 1
    run(0, n0)
    rate = cum[n0 - 1] / max(time.perf_counter() - t0, 1e-6)
    n1 = int(min(len(al), np.searchsorted(cum, rate * CPU_DP_SECONDS) + 1))
    t0 = time.perf_counter()
    c1 = run(0, n1)
    s1 = time.perf_counter() - t0
    threads = max(1, min(16, os.cpu_count() or 1))
    nt = int(min(len(al), np.searchsorted(cum, c1 / s1 * threads * CPU_DP_SECONDS) + 1))
    big_first = np.argsort(-(pl[:nt].astype(np.int64) * ql[:nt]), kind="stable")
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        ct = sum(ex.map(lambda k: run(int(k), int(k) + 1), big_first))
    st = time.perf_counter() - t0
    m = max(n1, nt)
    same = bool(np.array_equal(np.rint(scores[:m]).astype(np.int64), np.asarray(gpu_scores)[order[:m]]))
    return {"value": round(c1 / s1 / 1e9, 4), "unit": "GCUPS", "cores": 1, "kind": "port",
            "seconds": round(s1, 2), "cells": int(c1), "jobs": int(n1),
            "threads": {"value": round(ct / st / 1e9, 4), "cores": threads, "seconds": round(st, 2),
                        "cells": int(ct), "jobs": int(nt)},
            "scores_identical": same,
            "sample": f"oracle NWSmall (compiled, -O3) over the first jobs of a seeded random order of the "
                      f"{len(al)}-pair batch: {n1} jobs on one thread, {nt} jobs on {threads} threads; "
                      f"their scores compared with the GPU's"}


BAND_REPS = 3


def dp_batch_leg(ctx, dp, cpu):
    """The batched full DP over the 100k-pair distribution (C3): operands
    resident, one warm-up call over the whole batch, then one timed call."""
    import mauvehip
    out = {}
    A, ao, al, B, bo, bl = dp
    jobs_c3 = mauvehip.make_jobs(ao, al, bo, bl)
    cells3 = int((al.astype(np.int64) * bl).sum())
    dA3, dB3 = ctx.alloc(len(A) + 16), ctx.alloc(len(B) + 16)
    ctx.h2d(dA3, A)
    ctx.h2d(dB3, B)
    dP3 = ctx.alloc(int((al + bl).sum()) + 16)
    dR3 = ctx.alloc(16 * len(al) + 16)
    # warm-up over the whole batch: the trace arena (sized once per call
    # to the largest round, up to 70% of free HBM) is allocated here, not
    # in the timed call
    mauvehip.global_align_device(ctx, dA3, dB3, jobs_c3, dP3, dR3)
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.reset_kernel_times()
    t0 = time.perf_counter()
    mauvehip.global_align_device(ctx, dA3, dB3, jobs_c3, dP3, dR3)
    ctx.synchronize()
    el3 = time.perf_counter() - t0
    fwd_ms, fwd_n = ctx.kernel_time("dp_forward")
    tb_ms, tb_n = ctx.kernel_time("dp_traceback")
    gc = cells3 / el3 / 1e9
    gk = cells3 / (fwd_ms * fwd_n * 1e-3) / 1e9 if fwd_n else 0
    out["dp_batch"] = {"pairs": len(al), "cells": cells3, "seconds": round(el3, 3), "gcups": round(gc, 1),
                       "forward_kernel_gcups": round(gk, 1), "rounds": fwd_n,
                       "traceback_ms": round(tb_ms * tb_n, 2),
                       "roofline": {"kernel": "dp_forward", "bound": "valu", "achieved": round(gk, 1),
                                    "peak": round(DP_PEAK_GCUPS, 1), "unit": "GCUPS",
                                    "frac": round(gk / DP_PEAK_GCUPS, 4)}}
    log(f"[bench] DP batch {len(al)} pairs {cells3 / 1e12:.2f} Tcells: {gc:.1f} GCUPS end-to-end, "
        f"{gk:.1f} GCUPS forward kernel")
    if cpu:
        res3 = np.zeros(len(al), dtype=mauvehip.DP_RESULT_DTYPE)
        ctx.d2h(res3, dR3)
        out["dp_batch"]["cpu_baseline"] = cpu_dp_batch(dp, res3["score"])
        log(f"[bench] CPU NWSmall {json.dumps(out['dp_batch']['cpu_baseline'])}")
    for p in (dA3, dB3, dP3, dR3):
        ctx.free(p)
    return out["dp_batch"]


def c2_gap_band_leg(ctx, ref, mut, jobs):
    """The learned band on the recoded 4.6 Mb pair's own inter-anchor gaps
    (the core leg's chain gaps, A = the reference, B = the derivative): full DP
    against learned band (features + network + banded DP + fallback), best of
    three each, with mismatches against the full DP."""
    import mauvehip
    from mauvehip import band
    lut = np.full(256, 4, dtype=np.uint8)
    for c, v in zip(b"ACGTacgt", (0, 1, 2, 3, 0, 1, 2, 3)):
        lut[c] = v
    A = lut[np.frombuffer(np.asarray(ref).tobytes(), dtype=np.uint8)]
    B = lut[np.frombuffer(np.asarray(mut).tobytes(), dtype=np.uint8)]
    dj = band.DeviceJobs(ctx, A, B, jobs)
    mlp, delta, min_band = band.load_default(ctx)
    dj.align(None)
    t_full = []
    for _ in range(3):
        ctx.synchronize()
        t0 = time.perf_counter()
        dj.align(None)
        t_full.append(time.perf_counter() - t0)
    res_f, _ = dj.results()
    paths_f = dj.paths()
    t_l = []
    for _ in range(3):
        ctx.synchronize()
        t0 = time.perf_counter()
        dj.features()
        st = dj.align(mlp.predict_bands_device(dj.dF, dj.n, delta, min_band))
        t_l.append(time.perf_counter() - t0)
    res_l, _ = dj.results()
    sm, pm = band.count_mismatches(res_f, paths_f, res_l, dj.paths())
    out = {"jobs": int(dj.n), "full_dp_cells": int(dj.cells), "largest": [int(jobs["a_len"].max()), int(jobs["b_len"].max())],
           "full_seconds": round(min(t_full), 5), "learned_seconds": round(min(t_l), 5),
           "banded_jobs": int(st.banded), "learned_cells_computed": int(st.cells),
           "reruns_2w": int(st.rerun_doubled), "reruns_full": int(st.rerun_full),
           "score_mismatches": sm, "path_mismatches": pm}
    log(f"[bench] learned band on the 4.6 Mb pair's gaps: {json.dumps(out)}")
    mlp.free()
    dj.free()
    return out


def learned_band_leg(ctx, bandset):
    """Full DP vs learned band (features + network + banded DP with edge
    fallback, all timed) vs the best constant band, on held-out gaps."""
    import mauvehip
    from mauvehip import band
    A, ao, al, B, bo, bl = bandset
    dj = band.DeviceJobs(ctx, A, B, mauvehip.make_jobs(ao, al, bo, bl))
    mlp, delta, min_band = band.load_default(ctx)
    info = json.loads(str(np.load(os.path.join(ROOT, "mauve-py_amd", "mauvehip", "band_model.npz"),
                                  allow_pickle=False)["info_json"]))
    w_const = int(info["best_constant_band"])
    # warm-up of every kernel on a small prefix
    small = band.DeviceJobs(ctx, A, B, mauvehip.make_jobs(ao[:2000], al[:2000], bo[:2000], bl[:2000]))
    small.features()
    small.align(mlp.predict_bands_device(small.dF, small.n, delta, min_band))
    small.align(None)
    small.free()

    dj.align(None)      # grows the trace arena to its working size
    ctx.synchronize()
    t0 = time.perf_counter()
    dj.align(None)
    t_full = time.perf_counter() - t0
    res_f, lab = dj.results()
    paths_f = dj.paths()

    def learned():
        ctx.reset_kernel_times()
        ctx.synchronize()
        t0 = time.perf_counter()
        dj.features()
        bands = mlp.predict_bands_device(dj.dF, dj.n, delta, min_band)
        st = dj.align(bands)
        t = time.perf_counter() - t0
        passes = {k: round(ctx.kernel_time(k)[0] * 1e-3, 4) for k in ("band_pass1", "band_pass2", "band_pass3")}
        return t, st, ctx.kernel_time("dp_forward"), passes

    def constant():
        ctx.synchronize()
        t0 = time.perf_counter()
        st = dj.align(np.full(dj.n, w_const, dtype=np.uint32))
        return time.perf_counter() - t0, st

    # alternating repeats, the fastest of each kept; the results checked
    # after the first constant run and the last learned one
    runs_l, runs_c = [], []
    for r in range(BAND_REPS):
        runs_c.append(constant())
        if r == 0:
            res_c, _ = dj.results()
            sm_c, pm_c = band.count_mismatches(res_f, paths_f, res_c, dj.paths())
        runs_l.append(learned())
    res_l, dev_l = dj.results()
    sm_l, pm_l = band.count_mismatches(res_f, paths_f, res_l, dj.paths())
    t_learned, st_l, (fwd_ms, fwd_n), passes_l = min(runs_l, key=lambda r: r[0])
    t_const, st_c = min(runs_c, key=lambda r: r[0])

    cells = dj.cells
    gk = st_l.cells / (fwd_ms * fwd_n * 1e-3) / 1e9 if fwd_n else 0.0
    out = {"pairs": int(dj.n), "seed": 402, "full_dp_cells": cells,
           "full_seconds": round(t_full, 4), "full_gcups": round(cells / t_full / 1e9, 1),
           "learned_seconds": round(t_learned, 4), "learned_effective_gcups": round(cells / t_learned / 1e9, 1),
           "speedup_vs_full_dp": round(t_full / t_learned, 2),
           "learned_cells_computed": int(st_l.cells), "learned_reruns_2w": int(st_l.rerun_doubled),
           "learned_reruns_full": int(st_l.rerun_full), "learned_pass_seconds": passes_l,
           "learned_runs_seconds": [round(r[0], 4) for r in runs_l],
           "constant_runs_seconds": [round(r[0], 4) for r in runs_c],
           "score_mismatches": sm_l, "path_mismatches": pm_l, "delta": delta,
           "constant_band": w_const, "constant_seconds": round(t_const, 4),
           "constant_effective_gcups": round(cells / t_const / 1e9, 1),
           "constant_cells_computed": int(st_c.cells), "constant_score_mismatches": sm_c,
           "constant_path_mismatches": pm_c,
           "band_kernel_computed_gcups": round(gk, 1),
           "label_quantiles": {str(q): float(np.percentile(lab, q)) for q in (50, 90, 99, 100)}}
    log(f"[bench] learned band: full {t_full:.3f}s, learned {t_learned:.3f}s ({t_full / t_learned:.1f}x), "
        f"constant W={w_const} {t_const:.3f}s; mismatches score {sm_l} path {pm_l}")
    mlp.free()
    dj.free()
    return out


if __name__ == "__main__":
    sys.exit(main())
