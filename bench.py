#!/usr/bin/env python3
"""Benchmark of the alias-table hot path on one MI355X.

Headline workload: N = 1e8 FP64 weights i.i.d. Pareto(shape 1.1, scale 1) from
Philox seed 42, PSA+ build (W sum, light/heavy partition with PSA+ and the
fused leftover prefix sums, split, pack) at an explicit build shape (section
2048, 16 chains, n/p = 64; the shape changes the table, so it is an input,
chosen once with atg_autotune_shape), with the knobs that leave the table
unchanged (split method, knots, split/pack interleave) tuned at run time.
One step = one whole build over the resident weights.

Also measured: 1e9 sectioned samples (4096-row sections), basic sampling and
the rejection variant on that table; the four fills of the rejection array
(direct, bisection, p-ary, learned inverse-CDF index) on the Pareto and the
heavy-mix weights; the pack variants at the headline shape; the learned and
the p-ary split against bisection on N = 1e8 heavy-mix weights; the two
batched workloads (65,536 FP64 tables and 16,384 FP32 tables with the learned
split) with graph replay and per-table latencies; the implied-probability
error of the tables; the build at the reference's default shape; and the CPU
baselines on the host's cores.

Prints one JSON line.  `value` is build throughput (Gitems/s, weights already
in HBM); `roofline` is the dominant build kernel's algorithmic bytes over its
mean duration measured with HIP events on the build stream.
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
KERNEL_PHASES = ("sum", "partition_local", "partition_place", "split", "pack")
KERNEL_NAMES = {"sum": "sum_sections+sum_finish", "partition_local": "partition_local",
                "partition_place": "block_tree+partition_place", "split": "split_bisect or split_knots+split_learned",
                "pack": "pack_staged"}
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE + WRITE_SIZE
# (scaled by the counters of a device copy of known size) for the headline
# workload at the given shape; the file holds the commands and the counts.
MEASURED_TRAFFIC = {
    (100_000_000, 2048, 16, 1_562_500, "partition_local"): (
        2_312_409_000, "profiles/build_pareto1e8_s2048_c16_p1562500_hbm_traffic.md"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(n, R, p, row_bytes=16, w_bytes=8):
    """Algorithmic HBM bytes of one launch of each build kernel (group):
    weights read, packed rows written; leftover lists (item 4 B + weight 8 B +
    prefix 8 B) written once and read (item + weight) by the pack."""
    return {
        "sum": n * w_bytes,
        "partition_local": n * w_bytes + (n - R) * row_bytes,
        "partition_place": R * 20,
        "split": p * 2 * 16,  # the final probe pair per split; the search itself is latency-bound
        "pack": R * 12 + R * row_bytes,
    }


def time_events(stream, fn, reps):
    """Median and mean ms of `reps` calls of fn() between HIP events on `stream`."""
    import torch
    ts = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(reps):
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return statistics.median(ts), statistics.mean(ts)


def headline(args, stream, w):
    """The headline build: explicit shape, run-time knobs, `value` from the
    wall clock over K builds, per-kernel means from HIP events."""
    import torch
    import atg
    n = args.n
    shape = dict(section_size=args.section_size, chains=args.chains, num_splits=n // args.items_per_split)
    knobs = {"split": args.split, "learned_segments": args.learned_segments, "interleave": False}
    if not args.no_autotune:
        t0 = time.perf_counter()
        knobs, tms = atg.autotune(w, reps=3, **shape)
        knobs["median_ms"] = round(tms, 4)
        knobs["tuning_s"] = round(time.perf_counter() - t0, 2)
        log(f"[bench] run-time knobs {knobs}")
    t = atg.Table(n, split=knobs["split"], learned_segments=knobs["learned_segments"],
                  interleave=knobs["interleave"], **shape)
    t.set_weights(w, stream)
    for _ in range(args.warmup):
        t.build(stream)
    torch.cuda.synchronize()
    t.sync(stream)
    K = args.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        t.build(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    t.sync(stream)
    ms_per_step = (t1 - t0) * 1e3 / K
    # Per-kernel times: the same build phase by phase with events between.
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(len(KERNEL_PHASES) + 1)] for _ in range(K)]
    for k in range(K):
        evs[k][0].record(stream)
        for i, ph in enumerate(KERNEL_PHASES):
            t.build_phase(ph, stream)
            evs[k][i + 1].record(stream)
    torch.cuda.synchronize()
    t.sync(stream)
    kms = {ph: statistics.mean(evs[k][i].elapsed_time(evs[k][i + 1]) for k in range(K))
           for i, ph in enumerate(KERNEL_PHASES)}
    st = t.stats()
    log(f"[bench] build {ms_per_step:.3f} ms/step; kernels {kms}; R/N={st['R'] / n:.3f} p={st['p']}")
    return t, shape, knobs, ms_per_step, (t1 - t0), kms, st


def bench_sampling(args, stream, t, n):
    import torch
    S = args.samples
    out = torch.empty(S, dtype=torch.int32, device="cuda")
    res = {"samples": S, "section_rows": args.sample_section}
    for kind in ("sectioned", "basic"):
        t.sample(S, seed=1, kind=kind, section_rows=args.sample_section, out=out, stream=stream)
        seeds = iter(range(100, 200))
        ms, _ = time_events(stream, lambda: t.sample(S, seed=next(seeds), kind=kind, section_rows=args.sample_section,
                                                     out=out, stream=stream), 3)
        sb = 4 * S + (16 * n if kind == "sectioned" else 16 * S)
        res[f"{kind}_ms"] = round(ms, 3)
        res[f"{kind}_gsamples_s"] = round(S / (ms * 1e-3) / 1e9, 3)
        res[f"{kind}_hbm_frac"] = round(sb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        log(f"[bench] {kind} sampling {S} in {ms:.3f} ms")
    t.rejection_build(stream)  # first call allocates A
    torch.cuda.synchronize()
    rb_ms, _ = time_events(stream, lambda: t.rejection_build(stream), 3)
    A = t.rejection_size()
    t.rejection_sample(S, seed=3, out=out, stream=stream)
    seeds = iter(range(10, 100))
    rms, _ = time_events(stream, lambda: t.rejection_sample(S, seed=next(seeds), out=out, stream=stream), 3)
    rbytes = 4 * S + (A / n) * S * 16  # one 16-byte slot per trial
    res.update({"rejection_build_ms": round(rb_ms, 3), "rejection_A_over_N": round(A / n, 4),
                "rejection_ms": round(rms, 3), "rejection_gsamples_s": round(S / (rms * 1e-3) / 1e9, 3),
                "rejection_hbm_frac": round(rbytes / (rms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
    log(f"[bench] rejection build {rb_ms:.3f} ms, sampling {rms:.3f} ms")
    # By LDS-staged sections of A: the first call computes the section
    # masses (kept for later calls with the same section size); timed apart.
    slots = args.sample_section
    fm, _ = time_events(stream, lambda: t.rejection_sample_sectioned(S, seed=4, section_slots=slots, out=out,
                                                                     stream=stream), 1)
    seeds = iter(range(200, 300))
    sms, _ = time_events(stream, lambda: t.rejection_sample_sectioned(S, seed=next(seeds), section_slots=slots,
                                                                      out=out, stream=stream), 3)
    sbytes = 4 * S + 16 * A  # A read once, an int32 per sample
    res.update({"rejection_sectioned_ms": round(sms, 3), "rejection_sectioned_first_call_ms": round(fm, 3),
                "rejection_sectioned_gsamples_s": round(S / (sms * 1e-3) / 1e9, 3),
                "rejection_sectioned_hbm_frac": round(sbytes / (sms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "rejection_section_slots": slots})
    log(f"[bench] sectioned rejection sampling {sms:.3f} ms (first call {fm:.3f} ms)")
    del out
    torch.cuda.empty_cache()
    return res


def bench_reference_shape(args, stream, w):
    """The build at the reference's default shape: sections min(1000, N/10),
    16 chains, p = AUTO (320 per CU), bisection."""
    import torch
    import atg
    t = atg.Table(args.n)
    t.set_weights(w, stream)
    t.build(stream)
    torch.cuda.synchronize()
    ms, _ = time_events(stream, lambda: t.build(stream), 5)
    st = t.stats()
    del t
    torch.cuda.empty_cache()
    return {"section_size": st["section_len"], "chains": 16, "p": st["p"], "split": "bisect",
            "build_ms": round(ms, 4), "gitems_s": round(args.n / (ms * 1e-3) / 1e9, 3), "R_over_N": round(st["R"] / args.n, 4)}


def bench_pack_methods(args, stream, w, shape):
    """The pack variants at the headline shape: whole builds and pack phase
    times, the LDS-staged default against one lane per chain from HBM (PSA+
    on), and both against the weights-through-items pack (PSA+ off)."""
    import torch
    import atg
    res = {}
    for name, pack, psa in (("chunked", "chunked", True), ("basic", "basic", True),
                            ("chunked_no_psa", "chunked", False), ("no_weights_no_psa", "no_weights", False)):
        t = atg.Table(args.n, pack=pack, psa_plus=psa, record_phases=True, **shape)
        t.set_weights(w, stream)
        t.build(stream)
        torch.cuda.synchronize()
        ms, _ = time_events(stream, lambda: t.build(stream), max(3, args.steps // 4))
        res[name] = {"build_ms": round(ms, 4), "pack_ms": round(t.stats()["phase_ms"][4], 4)}
        t.close()
    torch.cuda.empty_cache()
    log(f"[bench] pack variants: {res}")
    return res


def bench_learned(args, stream):
    """Learned split index vs bisection on N = 1e8 weights where 1% of the
    items carry 90% of the mass (lognormal(10, 1) heavies rescaled, U(0,1)
    lights), at one explicit shape: whole-build times, split times, the
    table's bit identity and its implied-probability error."""
    import numpy as np
    import torch
    import atg
    n = args.n
    w = atg.generate_weights("heavy_mix", n, seed=42, param=0.9)
    shape = dict(section_size=args.section_size, chains=args.learned_chains, num_splits=n // args.items_per_split)
    res = {"workload": f"N={n:.0e} heavy_mix (1% heavy, 90% mass), {args.learned_segments} segments",
           "section_size": shape["section_size"], "chains": shape["chains"], "items_per_split": args.items_per_split}
    tabs = {}
    for name, split, knots in (("bisect", "bisect", args.learned_segments),
                               ("learned", "learned", args.learned_segments),
                               ("learned_65536", "learned", 65536), ("pary", "pary", args.learned_segments)):
        t = atg.Table(n, split=split, learned_segments=knots, record_phases=True, **shape)
        t.set_weights(w, stream)
        for _ in range(args.warmup):
            t.build(stream)
        torch.cuda.synchronize()
        ms, mean = time_events(stream, lambda: t.build(stream), args.steps)
        st = t.stats()
        res[name] = {"build_ms": round(ms, 4), "split_ms": round(st["phase_ms"][3], 4),
                     "gitems_s": round(n / (ms * 1e-3) / 1e9, 3), "p": st["p"]}
        if split != "bisect":
            res[name]["search_fallbacks"] = st["learned_fallbacks"]
            res[name]["monotone"] = st["monotone"]
        if name == "learned":
            rel, ok = t.post_build(verify=True)
            res["max_rel_err"] = rel
            res["reference_check_ok"] = ok
        tabs[name] = t.table()
        del t
        torch.cuda.empty_cache()
    res["speedup"] = round(res["bisect"]["build_ms"] / res["learned"]["build_ms"], 4)
    res["split_speedup"] = round(res["bisect"]["split_ms"] / max(res["learned"]["split_ms"], 1e-9), 3)
    res["speedup_65536_knots"] = round(res["bisect"]["build_ms"] / res["learned_65536"]["build_ms"], 4)
    res["pary_speedup"] = round(res["bisect"]["build_ms"] / res["pary"]["build_ms"], 4)
    a = tabs["bisect"]
    res["bit_identical"] = all(bool(np.array_equal(a[0].view(np.uint64), b[0].view(np.uint64)) and
                                    np.array_equal(a[1], b[1]))
                               for b in (tabs["learned"], tabs["learned_65536"], tabs["pary"]))
    log(f"[bench] learned vs plain: {res}")
    if not args.no_rejection_fills:
        res["rejection_fills"] = bench_rejection_fills(args, stream, w, f"N={n:.0e} heavy_mix (1% heavy, 90% mass)")
    return res


def bench_rejection_fills(args, stream, w, label):
    """The fills of the rejection array on one weight set (RejectionSampling::
    fillA by subclass, plus the learned inverse-CDF index): the whole
    rejection build and fillA alone per fill, the learned model's error
    bounds and fallbacks."""
    import numpy as np
    import torch
    import atg
    n, K = len(w), args.learned_segments
    t = atg.Table(n)
    t.set_weights(w, stream)
    res = {"workload": label, "knots": K}
    reps = max(3, args.steps // 4)
    for fill in ("direct", "bisect", "pary", "learned"):
        t.rejection_build(stream, fill=fill, knots=K)
        torch.cuda.synchronize()
        bms, _ = time_events(stream, lambda: t.rejection_build(stream, fill=fill, knots=K), reps)
        fms, _ = time_events(stream, lambda: t.rejection_fill(stream), reps)
        A = t.rejection_size()
        r = {"build_ms": round(bms, 4), "fill_ms": round(fms, 4), "fill_gslots_s": round(A / (fms * 1e-3) / 1e9, 3),
             # A written once (16 B a slot) over the fill's time
             "fill_A_write_frac": round(16 * A / (fms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if fill == "learned":
            r["fallbacks"] = t.rejection_fill_fallbacks()
            _, e = t.rejection_model(K)
            r["err_bound_median"], r["err_bound_max"] = int(np.median(e)), int(e.max())
        res[fill] = r
        log(f"[bench] rejection fill {fill}: build {bms:.3f} ms, fillA {fms:.3f} ms")
    res["A_size"] = A
    res["learned_vs_bisect_fill"] = round(res["bisect"]["fill_ms"] / res["learned"]["fill_ms"], 3)
    res["learned_vs_bisect_build"] = round(res["bisect"]["build_ms"] / res["learned"]["build_ms"], 3)
    res["pary_vs_bisect_fill"] = round(res["bisect"]["fill_ms"] / res["pary"]["fill_ms"], 3)
    t.close()
    return res


def bench_batch(args, stream, name, ntables, lo, hi, dist, dtype, split, per_table, tune_split, param=0.0):
    """Batched build of many tables: launch groups (and for the FP64 batch the
    split method) tuned in place, the build captured in a hipGraph and
    replayed; makespan from events, per-table latencies from device
    timestamps, `per_table` samples from every table."""
    import numpy as np
    import torch
    import atg
    offsets, w = atg.generate_batch(dist, ntables, lo, hi, seed=42, param=param, dtype=dtype)
    total = int(offsets[-1])
    b = atg.Batch(offsets, dtype=dtype, split=split, section_size=args.batch_section, chains=args.batch_chains,
                  num_splits=-args.batch_items_per_split)
    b.set_weights(w, stream)
    del w
    tuned = b.autotune(reps=3, makespan_slack=args.batch_slack, tune_split=tune_split)
    log(f"[bench] batch {name}: tuned {tuned}")
    b.use_graph(True)
    out = torch.empty(ntables * per_table, dtype=torch.int32, device="cuda")
    for _ in range(args.warmup):
        b.build(stream)
    torch.cuda.synchronize()
    b.sync(stream)
    lat = []
    build_ms, sample_ms = [], []
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    for _ in range(args.steps):
        e[0].record(stream)
        b.build(stream)
        e[1].record(stream)
        b.sample(per_table, seed=7, out=out, stream=stream)
        e[2].record(stream)
        torch.cuda.synchronize()
        build_ms.append(e[0].elapsed_time(e[1]))
        sample_ms.append(e[1].elapsed_time(e[2]))
        lat.append(b.latencies_ns() / 1e3)
    b.sync(stream)
    lat = np.concatenate(lat)
    ms = statistics.mean(build_ms)
    w_bytes = 4 if dtype == "f32" else 8
    _, ngroups = b.groups()
    res = {"workload": name, "tables": ntables, "items": total, "dtype": dtype, "split": tuned["split"],
           "section_size": args.batch_section, "chains": args.batch_chains,
           "items_per_split": args.batch_items_per_split, "graph": True,
           "groups": ngroups, "group_items": tuned["group_items"],
           "build_ms": round(ms, 4), "gitems_s": round(total / (ms * 1e-3) / 1e9, 3),
           "floor_hbm_frac": round(total * (2 * w_bytes + 16) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "samples_per_table": per_table, "sample_ms": round(statistics.mean(sample_ms), 4),
           "latency_us_p50": round(float(np.percentile(lat, 50)), 2),
           "latency_us_p99": round(float(np.percentile(lat, 99)), 2),
           "latency_us_max": round(float(lat.max()), 2)}
    log(f"[bench] batched {name}: {res}")
    del b, out
    torch.cuda.empty_cache()
    return res


def cpu_baseline(args, w):
    """The oracle (the same canonical construction, restated in C++) and the
    faithful single-thread restatements of the reference's CPU builds and
    samplers, timed on this host's cores over bounded samples."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import atg
    import oracle_binding as oracle
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(aff, 16))
    n = args.cpu_n
    x = w[:n] if n <= len(w) else atg.generate_weights("pareto", n, seed=42, param=1.1)
    shape = dict(chains=args.chains, section_size=args.section_size, num_splits=n // args.items_per_split)
    times = []
    ref = None
    for rep in range(args.cpu_reps + 1):  # the first run warms up
        t0 = time.perf_counter()
        ref = oracle.build(x, psa_plus=True, threads=cores, want_lists=False, **shape)
        if rep:
            times.append(time.perf_counter() - t0)
    tb = statistics.median(times)
    res = {"value": round(n / tb / 1e9, 4), "unit": "Gitems/s", "cores": cores, "kind": "port",
           "sample": f"canonical PSA+ build (the oracle) of the headline N={n:.0e} Pareto(1.1) FP64 weights at the "
                     f"headline shape, OpenMP over {cores} threads, median of {args.cpu_reps} after a warm-up",
           "host_cpus": aff}

    def timed(fn):
        t0 = time.perf_counter()
        r = fn()
        return time.perf_counter() - t0, r

    # Faithful single-thread restatements of the reference's CPU builds.
    ts, _ = timed(lambda: oracle.ref_split_build(x, 1400))
    res["ref_split_build_1t_gitems_s"] = round(n / ts / 1e9, 4)  # AliasTableSplit::build
    ts, _ = timed(lambda: oracle.ref_sweep_build(x))
    res["ref_sweep_build_1t_gitems_s"] = round(n / ts / 1e9, 4)  # AliasTableSweep::build
    ts, _ = timed(lambda: oracle.ref_stack_build(x))
    res["ref_stack_build_1t_gitems_s"] = round(n / ts / 1e9, 4)  # AliasTableStack::build
    # Sampling from the headline table: 1e7 basic samples, one thread and all cores.
    S = args.cpu_samples
    for th in (1, cores):
        oracle.set_threads(th)
        ts, _ = timed(lambda: oracle.sample_basic(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], 5, S))
        res[f"sample_basic_{th}t_gsamples_s"] = round(S / ts / 1e9, 4)
    # Rejection: the build (one thread, as RejectionSampling::build) and 1e7 samples.
    ts, (_, A) = timed(lambda: oracle.rejection_build(x, ref["stats"]["W"]))
    res["rejection_build_1t_gitems_s"] = round(n / ts / 1e9, 4)
    for th in (1, cores):
        oracle.set_threads(th)
        ts, _ = timed(lambda: oracle.sample_rejection(A, x, ref["stats"]["W"], 6, S))
        res[f"rejection_sample_{th}t_gsamples_s"] = round(S / ts / 1e9, 4)
    del A, ref
    # The reference's CPU configuration: N = 2^20 weights 1 - u in (0, 1] from
    # seed 42, the stack (Vose-like) and the PSA build, 1e7 samples, histogram.
    n1 = 1 << 20
    w1 = atg.generate_weights("unit_uniform", n1, seed=42)
    ts_stack, (W1, rw1, ra1) = timed(lambda: oracle.ref_stack_build(w1))
    oracle.set_threads(1)
    ts_psa, r1 = timed(lambda: oracle.build(w1, psa_plus=True, threads=1, want_lists=False))
    ts_s, s1 = timed(lambda: oracle.sample_basic(r1["row_w"], r1["row_alias"], r1["stats"]["W_N"], 42, 10_000_000))
    hist = np.bincount(s1, minlength=n1).astype(np.float64)
    expected = w1 / w1.sum() * hist.sum()
    from scipy import stats as sstats
    p = float(sstats.chisquare(hist, expected).pvalue)
    oracle.set_threads(cores)
    res["reference_cpu_config"] = {
        "workload": "N=2^20 weights 1-u in (0,1] (seed 42): stack and PSA builds, 1e7 samples, one thread",
        "stack_build_ms": round(ts_stack * 1e3, 2), "psa_build_ms": round(ts_psa * 1e3, 2),
        "sample_1e7_ms": round(ts_s * 1e3, 2), "chi2_pvalue": round(p, 4)}
    log(f"[bench] CPU baseline {res}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=100_000_000)
    ap.add_argument("--samples", type=int, default=1_000_000_000)
    ap.add_argument("--section-size", type=int, default=2048, help="PSA+ section (the explicit build shape)")
    ap.add_argument("--chains", type=int, default=16)
    ap.add_argument("--items-per-split", type=int, default=64, help="p = N / this")
    ap.add_argument("--split", choices=["bisect", "learned"], default="bisect",
                    help="split method when the run-time tuner is off")
    ap.add_argument("--no-autotune", action="store_true", help="no run-time tuning of the table-neutral knobs")
    ap.add_argument("--learned-segments", type=int, default=4096)
    ap.add_argument("--learned-chains", type=int, default=12)
    ap.add_argument("--sample-section", type=int, default=4096)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-pack-methods", action="store_true", help="skip the pack variants")
    ap.add_argument("--no-rejection-fills", action="store_true",
                    help="skip the fills of the rejection array (direct, bisection, p-ary, learned)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-batch-learned", action="store_true", help="skip the 16,384-table FP32 batch")
    ap.add_argument("--batch-section", type=int, default=2048, help="section size of every batched table")
    ap.add_argument("--batch-chains", type=int, default=32)
    ap.add_argument("--batch-items-per-split", type=int, default=38)
    ap.add_argument("--batch-slack", type=float, default=0.10,
                    help="batch tuner: smallest p99 latency within this makespan slack of the fastest")
    ap.add_argument("--cpu-n", type=int, default=100_000_000)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--cpu-samples", type=int, default=10_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sampling", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="skip the implied-probability checks")
    ap.add_argument("--no-reference-shape", action="store_true", help="skip the build at the reference's default shape")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy a device buffer of this many MiB 3 times (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path would shard into "
              f"independent per-GPU tables)", file=sys.stderr)
        sys.exit(2)

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
    t, shape, knobs, ms_per_step, wall, kms, st = headline(args, stream, w)
    K = args.steps
    R, p = st["R"], st["p"]
    kb = kernel_bytes(n, R, p)
    dom = max(KERNEL_PHASES, key=lambda ph: kms[ph])
    achieved = kb[dom] / (kms[dom] * 1e-3) / 1e9
    traffic, traffic_src = MEASURED_TRAFFIC.get((n, shape["section_size"], shape["chains"], shape["num_splits"], dom),
                                                (None, None))
    build_ms_events = sum(kms.values())
    b_build = n * (2 * 8 + 16) + R * (2 * 16 + 8)  # N (2 s_w + s_row) + R (2 s_LH + 8), s_LH = 16
    result = {
        "metric": METRIC,
        "value": round(n * K / wall / 1e9, 3),
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
        "config": {"workload": "PSA+ alias-table build, N=1e8 FP64 Pareto(1.1) weights (seed 42); also 1e9 "
                               "sectioned samples (2^12-row sections) + rejection variant",
                   "n": n, "section_size": shape["section_size"], "section_len": st["section_len"],
                   "chains": shape["chains"], "p": p, "layout": "aos", "psa_plus": True,
                   "shape": "explicit input (chosen once by atg_autotune_shape)",
                   "run_time_knobs": knobs},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": kb[dom], "mean_ms": round(kms[dom], 4)},
        "build": {"kernel_ms": {k: round(v, 4) for k, v in kms.items()},
                  "kernel_hbm_frac": {k: round(kb[k] / (kms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                      for k in KERNEL_PHASES},
                  "R": R, "R_over_N": round(R / n, 4), "P_w": 2, "s_LH": 16, "B_build": b_build,
                  "build_hbm_frac": round(b_build / (build_ms_events * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                  "floor_fraction": round(n * 32 / (build_ms_events * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                  "W": st["W"], "monotone": st["monotone"], "split_fallbacks": st["split_fallbacks"],
                  "learned_fallbacks": st["learned_fallbacks"]},
    }
    if not args.no_precision:
        rel, ok = t.post_build(verify=True)
        result["precision"] = {"headline_max_rel_err": rel, "headline_reference_check_ok": ok}
        log(f"[bench] headline table max relative implied-probability error {rel:.3e} (reference check {ok})")
    if not args.no_sampling:
        result["sampling"] = bench_sampling(args, stream, t, n)
    del t
    torch.cuda.empty_cache()
    if not args.no_rejection_fills:
        result["rejection_fills"] = bench_rejection_fills(args, stream, w, f"N={n:.0e} Pareto(1.1)")
    torch.cuda.empty_cache()
    if not args.no_pack_methods:
        result["pack_methods"] = bench_pack_methods(args, stream, w, shape)
    if not args.no_reference_shape:
        result["reference_shape"] = bench_reference_shape(args, stream, w)
        log(f"[bench] reference shape: {result['reference_shape']}")
    if not args.no_learned:
        result["learned"] = bench_learned(args, stream)
        if "precision" in result:
            result["precision"]["heavy_mix_learned_max_rel_err"] = result["learned"].get("max_rel_err")
    if not args.no_batch:
        result["batched"] = [bench_batch(args, stream, "65,536 tables, sizes log-uniform [16, 2^16], Exp(1) FP64",
                                         65536, 16, 1 << 16, "exponential", "f64", "bisect", 64, True)]
        if not args.no_batch_learned:
            result["batched"].append(bench_batch(
                args, stream, "16,384 tables, sizes log-uniform [2^10, 2^20], FP32 rank-Zipf(1.2) shuffled, learned",
                16384, 1 << 10, 1 << 20, "zipf_shuffled", "f32", "learned", 256, False, param=1.2))
    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline on N={args.cpu_n}")
        result["cpu_baseline"] = cpu_baseline(args, w)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
