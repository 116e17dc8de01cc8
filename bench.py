#!/usr/bin/env python3
"""Benchmark of the alias-table hot path on one MI355X.

Workload (the headline configuration): N = 1e8 FP64 weights i.i.d.
Pareto(shape 1.1, scale 1) from Philox seed 42, plain PSA+ build (W sum,
light/heavy partition + PSA+ with fused leftover prefix sums, split, pack),
then 1e9 sectioned samples (2^12-row sections) and the rejection-sampling
variant.  One step = one full build over the resident weights.

Prints one JSON line.  `value` is build throughput (Gitems/s, weights already in
HBM); `roofline` is the dominant build kernel's algorithmic bytes over its mean
duration measured with HIP events on the build stream; `cpu_baseline` is the CPU
oracle (the same canonical build) on the host's cores over a bounded sample.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "alias-table-gpu_amd"))

METRIC = "build Gitems/s + Gsamples/s at N=1e8 (%HBM peak); learned/batched vs plain"
HBM_PEAK_GBS = 8000.0
# HBM bytes per build phase measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (calibrated on a 1 GiB device copy) for the headline workload at this shape;
# profiles/build_pareto1e8_s2048_c24_p1562500_hbm_traffic.md has the commands
# and counts.
MEASURED_TRAFFIC = {(100_000_000, 2048, 24, 1562500): {"sum": 804_000_000, "partition": 4_249_000_000,
                                                       "split": 3_007_000_000, "pack": 1_821_000_000}}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def phase_bytes(n, R, p, row_bytes=16, w_bytes=8):
    """Algorithmic HBM bytes of one launch of each build phase.

    Leftover lists are SoA: item (4 B) + weight (8 B) + prefix (8 B) written,
    item + weight (12 B) read by the pack."""
    return {
        "sum": n * w_bytes,
        "partition": n * w_bytes + (n - R) * row_bytes + R * 20,
        "split": p * 2 * 16,  # the final probe pair per split; the search itself is latency-bound
        "pack": R * 12 + R * row_bytes,
    }


def cpu_baseline(n_cpu, args):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import atg
    import oracle_binding as oracle
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(cores, 64))
    w = atg.generate_weights("pareto", n_cpu, seed=42, param=1.1)
    times = []
    for rep in range(3):
        t0 = time.perf_counter()
        oracle.build(w, psa_plus=True, chains=args.chains, section_size=args.section_size,
                     num_splits=args.num_splits or 81920, threads=threads, want_lists=False)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    # faithful single-thread restatement of the reference CPU PSA (AliasTableSplit::build)
    t0 = time.perf_counter()
    oracle.ref_split_build(w, 1400)
    t_ref = time.perf_counter() - t0
    return {"value": round(n_cpu / t / 1e9, 4), "unit": "Gitems/s", "cores": threads, "kind": "port",
            "sample": f"canonical PSA+ build of N={n_cpu:.0e} Pareto(1.1) FP64 weights (same generator, seed 42), "
                      f"OpenMP over {threads} threads, median of 3",
            "single_thread_reference_psa_gitems_s": round(n_cpu / t_ref / 1e9, 4)}


def time_builds(t, stream, steps, warmup):
    """Mean ms of `steps` whole builds (HIP events on the build stream)."""
    import torch
    for _ in range(warmup):
        t.build(stream)
    torch.cuda.synchronize()
    t.sync()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        t.build(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t.sync()
    return e0.elapsed_time(e1) / steps


def bench_learned(args, stream):
    """Learned split index vs bisection on N=1e8 weights where 1% of the items
    carry 90% of the mass (lognormal(10, 1) heavies rescaled, U(0,1) lights):
    whole-build times of each method at its own autotuned shape and of the
    learned index at the plain build's shape, whose table must be bit-identical
    to the plain table."""
    import numpy as np
    import atg
    n = args.n
    w = atg.generate_weights("heavy_mix", n, seed=42, param=0.9)
    res = {"workload": f"N={n:.0e} heavy_mix (1% heavy, 90% mass), {args.learned_segments} segments; "
                       f"each split method at its own autotuned shape"}
    tabs = {}
    for split in ("bisect", "learned"):
        if args.no_autotune:
            shape = {"section_size": args.section_size, "chains": args.chains, "num_splits": args.learned_p}
        else:
            shape, _ = atg.autotune(w, split=split, learned_segments=args.learned_segments)
        t = atg.Table(n, psa_plus=True, chains=shape["chains"], section_size=shape["section_size"],
                      num_splits=shape["num_splits"], split=split, learned_segments=args.learned_segments,
                      record_phases=True)
        t.set_weights(w, stream)
        ms = time_builds(t, stream, args.steps, args.warmup)
        st = t.stats()
        res[split] = {"build_ms": round(ms, 4), "split_ms": round(st["phase_ms"][3], 4),
                      "gitems_s": round(n / (ms * 1e-3) / 1e9, 3), "section_size": shape["section_size"],
                      "chains": shape["chains"], "p": st["p"]}
        if split == "learned":
            res[split]["learned_fallbacks"] = st["learned_fallbacks"]
            res[split]["monotone"] = st["monotone"]
        if split == "bisect":
            tabs[split] = t.table()
        del t
    res["speedup"] = round(res["bisect"]["build_ms"] / res["learned"]["build_ms"], 4)
    res["split_speedup"] = round(res["bisect"]["split_ms"] / max(res["learned"]["split_ms"], 1e-9), 3)
    # Bit identity: the learned index at the plain build's shape must give the
    # plain build's table.
    sh = res["bisect"]
    t = atg.Table(n, psa_plus=True, chains=sh["chains"], section_size=sh["section_size"], num_splits=sh["p"],
                  split="learned", learned_segments=args.learned_segments, record_phases=True)
    t.set_weights(w, stream)
    res["learned_same_shape_build_ms"] = round(time_builds(t, stream, args.steps, args.warmup), 4)
    res["speedup_same_shape"] = round(res["bisect"]["build_ms"] / res["learned_same_shape_build_ms"], 4)
    a, b = tabs["bisect"], t.table()
    del t
    res["bit_identical"] = bool(np.array_equal(a[0].view(np.uint64), b[0].view(np.uint64)) and
                                np.array_equal(a[1], b[1]))
    log(f"[bench] learned vs plain: {res}")
    return res


def bench_batch(args, stream, name, ntables, lo, hi, dist, dtype, split, per_table, param=0.0):
    """Batched build of many tables (one launch per phase over all tables), plus
    `per_table` samples from every table; per-table latencies from device
    timestamps."""
    import numpy as np
    import torch
    import atg
    offsets, w = atg.generate_batch(dist, ntables, lo, hi, seed=42, param=param, dtype=dtype)
    total = int(offsets[-1])
    b = atg.Batch(offsets, dtype=dtype, split=split, section_size=args.batch_section, chains=args.batch_chains,
                  num_splits=-args.batch_items_per_split)
    b.set_weights(w, stream)
    del w
    out = torch.empty(ntables * per_table, dtype=torch.int32, device="cuda")
    for _ in range(args.warmup):
        b.build(stream)
    torch.cuda.synchronize()
    b.sync()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    K = args.steps
    build_ms, sample_ms = [], []
    for _ in range(K):
        ev[0].record(stream)
        b.build(stream)
        ev[1].record(stream)
        b.sample(per_table, seed=7, out=out, stream=stream)
        ev[2].record(stream)
        torch.cuda.synchronize()
        build_ms.append(ev[0].elapsed_time(ev[1]))
        sample_ms.append(ev[1].elapsed_time(ev[2]))
    b.sync()
    lat = b.latencies_ns() / 1e3
    ms = statistics.mean(build_ms)
    row_bytes = 16
    w_bytes = 4 if dtype == "f32" else 8
    res = {"workload": name, "tables": ntables, "items": total, "dtype": dtype, "split": split,
           "section_size": args.batch_section, "chains": args.batch_chains,
           "items_per_split": args.batch_items_per_split,
           "build_ms": round(ms, 4), "gitems_s": round(total / (ms * 1e-3) / 1e9, 3),
           "floor_hbm_frac": round(total * (2 * w_bytes + row_bytes) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "samples_per_table": per_table, "sample_ms": round(statistics.mean(sample_ms), 4),
           "latency_us_p50": round(float(np.percentile(lat, 50)), 2),
           "latency_us_p99": round(float(np.percentile(lat, 99)), 2),
           "latency_us_max": round(float(lat.max()), 2)}
    log(f"[bench] batched {name}: {res}")
    del b, out
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=100_000_000)
    ap.add_argument("--samples", type=int, default=1_000_000_000)
    ap.add_argument("--section-size", type=int, default=2048,
                    help="PSA+ section (0 = the reference's min(1000, N/10))")
    ap.add_argument("--chains", type=int, default=32)
    ap.add_argument("--num-splits", type=int, default=2621440, help="p (0 = the reference's AUTO, 320 per CU)")
    ap.add_argument("--split", choices=["bisect", "learned"], default="bisect")
    ap.add_argument("--learned-segments", type=int, default=4096)
    ap.add_argument("--sample-section", type=int, default=4096)
    ap.add_argument("--learned-p", type=int, default=2621440, help="p of the learned-vs-plain comparison")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-section", type=int, default=2048, help="section size of every batched table")
    ap.add_argument("--batch-chains", type=int, default=32)
    ap.add_argument("--batch-items-per-split", type=int, default=38)
    ap.add_argument("--batch-learned", action="store_true",
                    help="also the 16,384-table FP32 Zipf batch with the learned index (2.4e9 items)")
    ap.add_argument("--cpu-n", type=int, default=10_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sampling", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy a device buffer of this many MiB 3 times (scale for HBM byte counters)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --section-size/--chains/--num-splits instead of the autotuner's choice")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path would shard into "
              f"independent per-GPU tables)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import atg

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calibrate_copy_mib > 0:
        # A copy of known size: its FETCH_SIZE / WRITE_SIZE counters give the
        # scale that turns the build kernels' counters into bytes.
        src = torch.empty(args.calibrate_copy_mib << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"[bench] calibration copies of {args.calibrate_copy_mib} MiB done")
    n = args.n
    log(f"[bench] generating N={n} Pareto(1.1) weights")
    w = atg.generate_weights("pareto", n, seed=42, param=1.1)
    stream = torch.cuda.current_stream()
    tuned = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        shape, tms = atg.autotune(w, split=args.split, learned_segments=args.learned_segments)
        tuned = dict(shape, median_ms=round(tms, 4), tuning_s=round(time.perf_counter() - t0, 2))
        args.section_size, args.chains, args.num_splits = shape["section_size"], shape["chains"], shape["num_splits"]
        log(f"[bench] autotuned shape {tuned}")
    t = atg.Table(n, psa_plus=True, chains=args.chains, section_size=args.section_size,
                  num_splits=args.num_splits, split=args.split, learned_segments=args.learned_segments)
    t.set_weights(w, stream)
    for _ in range(args.warmup):
        t.build(stream)
    torch.cuda.synchronize()
    t.sync()
    log(f"[bench] warmup done; timing {args.steps} builds")

    K = args.steps
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        e = evs[k]
        e[0].record(stream)
        for i, ph in enumerate(atg.Table.PHASES):
            t.build_phase(ph, stream)
            e[i + 1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    t.sync()
    st = t.stats()
    ms_per_step = (t1 - t0) * 1e3 / K
    phase_ms = {ph: statistics.mean(evs[k][i].elapsed_time(evs[k][i + 1]) for k in range(K))
                for i, ph in enumerate(atg.Table.PHASES)}
    build_ms_events = statistics.mean(evs[k][0].elapsed_time(evs[k][4]) for k in range(K))
    R, p = st["R"], st["p"]
    pb = phase_bytes(n, R, p)
    dom = max(("sum", "partition", "pack"), key=lambda ph: phase_ms[ph])
    achieved = pb[dom] / (phase_ms[dom] * 1e-3) / 1e9
    b_build = n * (2 * 8 + 16) + R * (2 * 12 + 8)
    log(f"[bench] build {ms_per_step:.3f} ms/step (events {build_ms_events:.3f}); phases {phase_ms}; R/N={R / n:.3f} p={p}")

    result = {
        "metric": METRIC,
        "value": round(n * K / (t1 - t0) / 1e9, 3),
        "unit": "Gitems/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain PSA+ alias-table build, N=1e8 FP64 Pareto(1.1) weights (seed 42), "
                               "then 1e9 sectioned samples (2^12-row sections) + rejection variant",
                   "n": n, "section_size": st["section_len"], "chains": args.chains, "p": p, "layout": "aos",
                   "psa_plus": True, "split": args.split, "autotuned": tuned},
        "roofline": {"bound": "hbm", "kernel": {"partition": "partition_local+partition_place", "sum": "sum_sections",
                                                "pack": "pack_staged"}[dom],
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((n, args.section_size, args.chains, args.num_splits),
                                                     {}).get(dom),
                     "algorithmic_bytes": pb[dom], "mean_ms": round(phase_ms[dom], 4)},
        "build": {"phase_ms": {k: round(v, 4) for k, v in phase_ms.items()},
                  "R": R, "R_over_N": round(R / n, 4), "P_w": 2, "B_build": b_build,
                  "build_hbm_frac": round(b_build / (build_ms_events * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                  "floor_fraction": round(n * 32 / (build_ms_events * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                  "W": st["W"], "monotone": st["monotone"], "split_fallbacks": st["split_fallbacks"]},
    }

    if not args.no_sampling:
        S = args.samples
        out = torch.empty(S, dtype=torch.int32, device=dev)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        t.sample(S, seed=1, kind="sectioned", section_rows=args.sample_section, out=out, stream=stream)
        ts = []
        for rep in range(3):
            ev[0].record(stream)
            t.sample(S, seed=100 + rep, kind="sectioned", section_rows=args.sample_section, out=out, stream=stream)
            ev[1].record(stream)
            torch.cuda.synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        ms = statistics.median(ts)
        sb = 4 * S + 16 * n
        result["sampling"] = {"sectioned_gsamples_s": round(S / (ms * 1e-3) / 1e9, 3), "sectioned_ms": round(ms, 3),
                              "samples": S, "section_rows": args.sample_section,
                              "sectioned_hbm_frac": round(sb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"[bench] sectioned sampling {S} in {ms:.3f} ms")
        t.rejection_build(stream)  # first call allocates A
        torch.cuda.synchronize()
        ts = []
        for rep in range(3):
            ev[0].record(stream)
            t.rejection_build(stream)
            ev[1].record(stream)
            torch.cuda.synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        rb_ms = statistics.median(ts)
        A = t.rejection_size()
        t.rejection_sample(S, seed=3, out=out, stream=stream)
        ts = []
        for rep in range(3):
            ev[0].record(stream)
            t.rejection_sample(S, seed=10 + rep, out=out, stream=stream)
            ev[1].record(stream)
            torch.cuda.synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        rms = statistics.median(ts)
        rbytes = 4 * S + (A / n) * S * 16  # one 16-byte slot per trial
        result["sampling"].update({"rejection_build_ms": round(rb_ms, 3), "rejection_A_over_N": round(A / n, 4),
                                   "rejection_gsamples_s": round(S / (rms * 1e-3) / 1e9, 3),
                                   "rejection_hbm_frac": round(rbytes / (rms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
        log(f"[bench] rejection build {rb_ms:.3f} ms, sampling {rms:.3f} ms")
        del out

    if not args.no_learned:
        result["learned"] = bench_learned(args, stream)
    if not args.no_batch:
        result["batched"] = [bench_batch(args, stream, "65,536 tables, sizes log-uniform [16, 2^16], Exp(1) FP64",
                                         65536, 16, 1 << 16, "exponential", "f64", "bisect", 64)]
        if args.batch_learned:
            result["batched"].append(bench_batch(
                args, stream, "16,384 tables, sizes log-uniform [2^10, 2^20], FP32 rank-Zipf(1.2) shuffled, learned",
                16384, 1 << 10, 1 << 20, "zipf_shuffled", "f32", "learned", 256, param=1.2))

    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline on N={args.cpu_n}")
        result["cpu_baseline"] = cpu_baseline(args.cpu_n, args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
