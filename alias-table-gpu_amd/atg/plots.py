"""Measurement sweeps in the schema of the reference's plots/ harness
(plots/PlotConstruction.cuh, plots/PlotSampling.cuh): a header line, then
one semicolon-separated row per measurement, so that the reference's figure
scripts read the output unchanged.  Build and sampling times come from HIP
events on the stream the work runs on (the reference's Timer,
utils/Timer.cu:3-22); phase times from atg_stats.phase_ms (TableSplitTimer,
utils/Timer.cuh:17-45, printed as TimingResult::print, utils/Timer.cu:42-48).

    python -m atg.plots <sweep> [--quick] [--out FILE]

Sweeps (reference function in brackets):
  build_soa            type;duration;N            [plotBuildSpeedStructOfArrays]
  build_distribution   type;duration;distribution [plotBuildSpeedWeightDistribution]
  thread_number        type;duration;p;N          [plotThreadNumber]
  split_variants       type;duration;splits       [plotBuildSpeedSplitVariants]
  with_transfer        type;duration;N            [plotBuildSpeedWithTransfer]
  light_vs_heavy       step;duration;fractionHeavy [plotBuildSpeedLightVsHeavy]
  sampler_method       type;N;numSamples;gSamplesPerSec [plotSamplingSpeedSamplerMethod]
  sampling_distribution_speed type;N;gSamplesPerSec;distribution [plotSamplingSpeedWeightDistribution]
  number_of_samples    type;N;numSamples;gSamplesPerSec [plotNumberOfSamples]
  sampling_speed       type;N;gSamplesPerSec      [plotSamplingSpeed]
  distribution         type;item;sampled          [plotSamplingDistribution]
  rejection_performance method;N;weight0;time     [plotRejectionPerformance]
"""
import argparse
import sys

import numpy as np

from . import Table, generate_weights

ITERATIONS = 20  # builds per measurement (BENCHMARK_BUILD_SPEED_ITERATIONS)

# The reference's WeightDistribution enum values by name (SamplingAlgorithm.cuh).
DIST_IDS = {"sine": 1, "best_worst": 2, "uniform": 3, "ramp": 4, "powerlaw1": 5, "powerlaw2": 6,
            "one_heavy": 7, "one_light": 8, "powerlaw1_shuffled": 9, "powerlaw05_shuffled": 10}


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("atg.plots measures on the GPU")
    return torch


def table_name(t):
    return "AliasTableSplitGpu<{}>(psa={},split={},p={})".format(
        "StructOfArrays" if t.layout == "soa" else "ArrayOfStructs", int(t.psa_plus), t.split, t.stats()["p"])


class _Events:
    def __init__(self, torch):
        self.torch = torch
        self.stream = torch.cuda.current_stream()

    def time(self, fn, reps):
        e0 = self.torch.cuda.Event(enable_timing=True)
        e1 = self.torch.cuda.Event(enable_timing=True)
        e0.record(self.stream)
        for _ in range(reps):
            fn()
        e1.record(self.stream)
        self.torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps


def make_table(n, dist="uniform", layout="aos", psa_plus=True, split="bisect", num_splits=0, w=None, **kw):
    t = Table(n, layout=layout, psa_plus=psa_plus, split=split, num_splits=num_splits, **kw)
    t.psa_plus, t.split = psa_plus, split
    t.set_weights(generate_weights(dist, n, seed=42) if w is None else w)
    return t


def build_speed(ev, t, append, out, iterations=ITERATIONS):
    """plotBuildSpeed (PlotConstruction.cuh:36-48): name;mean build ms;append."""
    t.build(ev.stream)
    ms = ev.time(lambda: t.build(ev.stream), iterations)
    t.sync()
    out.write(f"{table_name(t)};{ms};{append}\n")


def sweep_build_soa(ev, out, quick):
    out.write("type;duration;N\n")
    ns = [500_000, 1_000_000] if quick else range(500_000, 10_000_001, 500_000)
    for n in ns:
        for layout in ("aos", "soa"):
            build_speed(ev, make_table(n, layout=layout), n, out)


def sweep_build_distribution(ev, out, quick):
    out.write("type;duration;distribution\n")
    n = 1_000_000 if quick else 10_000_000
    for dist in ("uniform", "powerlaw2", "ramp"):
        w = generate_weights(dist, n, seed=42)
        for psa in (False, True):
            build_speed(ev, make_table(n, psa_plus=psa, w=w), DIST_IDS[dist], out)
        t = make_table(n, w=w)
        t.rejection_build(ev.stream)
        ms = ev.time(lambda: t.rejection_build(ev.stream), ITERATIONS)
        out.write(f"RejectionSampling;{ms};{DIST_IDS[dist]}\n")


def sweep_thread_number(ev, out, quick):
    out.write("type;duration;p;N\n")
    for n in ([1_000_000] if quick else [1_000_000, 10_000_000]):
        ps = [256, 512] if quick else range(256, 4 * 1024, 256)
        for p in ps:
            build_speed(ev, make_table(n, psa_plus=False, num_splits=p), f"{p};{n}", out)


def sweep_split_variants(ev, out, quick):
    """Split-phase time of bisection and the learned index against p."""
    out.write("type;duration;splits\n")
    n = 1_000_000 if quick else 10_000_000
    w = generate_weights("uniform", n, seed=42)
    splits = [1024, 2048] if quick else range(1024, 10 * 1024 + 1, 64)
    for p in splits:
        for split in ("bisect", "learned"):
            t = make_table(n, split=split, num_splits=p, w=w, record_phases=True)
            ph = []
            for _ in range(10):
                t.build(ev.stream)
                t.sync()
                ph.append(t.stats()["phase_ms"][3])
            out.write(f"{table_name(t)};{float(np.mean(ph))};{p}\n")


def sweep_with_transfer(ev, out, quick):
    out.write("type;duration;N\n")
    ns = [100_000, 200_000] if quick else range(100_000, 6_000_001, 100_000)
    for n in ns:
        w = generate_weights("uniform", n, seed=42)
        t = make_table(n, w=w)
        ms = ev.time(lambda: t.build(ev.stream), ITERATIONS)
        out.write(f"WithoutTransfer;{ms};{n}\n")
        ms = ev.time(lambda: (t.set_weights(w, ev.stream), t.build(ev.stream)), ITERATIONS)
        out.write(f"WithTransfer;{ms};{n}\n")
        ms = ev.time(lambda: (t.set_weights(w, ev.stream), t.build(ev.stream), t.table()), ITERATIONS)
        out.write(f"FullTransfer;{ms};{n}\n")


def sweep_light_vs_heavy(ev, out, quick, random_positions=True):
    """Per-phase times against the fraction of heavy items (weights 1 and 2)."""
    out.write("step;duration;fractionHeavy\n")
    n = 1_000_000 if quick else 10_000_000
    rng = np.random.default_rng(42)
    for frac in ([2, 50] if quick else range(2, 100, 4)):
        w = np.ones(n)
        k = int(0.01 * frac * n)
        idx = rng.choice(n, k, replace=False) if random_positions else np.arange(k)
        w[idx] = 2.0
        t = make_table(n, w=w, record_phases=True)
        ph = []
        for _ in range(10 if quick else 250):
            t.build(ev.stream)
            t.sync()
            ph.append(t.stats()["phase_ms"])
        m = np.mean(np.array(ph), axis=0)
        for name, v in zip(("Sum", "Partition", "Prefix sum", "Split", "Pack"), m):
            out.write(f"{name};{v};{frac}\n")


def _gsps(ev, t, kind, count):
    torch = ev.torch
    buf = torch.empty(count, dtype=torch.int32, device="cuda")
    if kind == "rejection":
        fn = lambda: t.rejection_sample(count, seed=1, out=buf, stream=ev.stream)  # noqa: E731
    else:
        fn = lambda: t.sample(count, seed=1, kind=kind, out=buf, stream=ev.stream)  # noqa: E731
    fn()
    ms = ev.time(fn, 5)
    return count / (ms * 1e-3) / 1e9


SAMPLER_NAMES = {"basic": "SamplerGpuBasic", "sectioned": "SamplerGpuSectioned", "rejection": "SamplerRejection"}


def sweep_sampler_method(ev, out, quick):
    out.write("type;N;numSamples;gSamplesPerSec\n")
    n = 1_000_000 if quick else 10_000_000
    t = make_table(n)
    t.build(ev.stream)
    t.sync()
    counts = [100_000, 1_100_000] if quick else range(100_000, 50_000_001, 1_000_000)
    for s in counts:
        for kind in ("basic", "sectioned"):
            out.write(f"{SAMPLER_NAMES[kind]};{n};{s};{_gsps(ev, t, kind, s)}\n")


def sweep_sampling_distribution_speed(ev, out, quick):
    out.write("type;N;gSamplesPerSec;distribution\n")
    n = 1_000_000 if quick else 10_000_000
    s = 1_000_000 if quick else 100_000_000
    for dist, did in DIST_IDS.items():
        t = make_table(n, dist=dist)
        t.build(ev.stream)
        t.sync()
        for kind in ("basic", "sectioned"):
            out.write(f"{SAMPLER_NAMES[kind]};{n};{_gsps(ev, t, kind, s)};{did}\n")
        if quick:
            break


def sweep_number_of_samples(ev, out, quick):
    out.write("type;N;numSamples;gSamplesPerSec\n")
    for n in ([1_000_000] if quick else [1_000_000, 10_000_000, 100_000_000]):
        t = make_table(n)
        t.build(ev.stream)
        t.sync()
        for s in ([1_000_000] if quick else [10 ** k for k in range(5, 10)]):
            out.write(f"{SAMPLER_NAMES['sectioned']};{n};{s};{_gsps(ev, t, 'sectioned', s)}\n")


def sweep_sampling_speed(ev, out, quick):
    out.write("type;N;gSamplesPerSec\n")
    s = 1_000_000 if quick else 100_000_000
    for n in ([1_000_000] if quick else [10 ** k for k in range(3, 9)]):
        t = make_table(n)
        t.build(ev.stream)
        t.rejection_build(ev.stream)
        t.sync()
        for kind in ("sectioned", "basic", "rejection"):
            out.write(f"{SAMPLER_NAMES[kind]};{n};{_gsps(ev, t, kind, s)}\n")


def sweep_distribution(ev, out, quick):
    """Histogram of the sampled items (plotSamplingDistribution)."""
    torch = ev.torch
    out.write("type;item;sampled\n")
    n = 100
    t = make_table(n, dist="sine", psa_plus=False)
    t.build(ev.stream)
    t.sync()
    hist = torch.zeros(n, dtype=torch.int32, device="cuda")
    t.sample(100_000 if quick else 10_000_000, seed=1, kind="sectioned", hist=hist, stream=ev.stream)
    torch.cuda.synchronize()
    for i, c in enumerate(hist.cpu().numpy()):
        out.write(f"{SAMPLER_NAMES['sectioned']};{i};{int(c)}\n")


REJECTION_FILL_NAMES = {"bisect": "binarySearch", "direct": "direct", "pary": "pArySearch", "learned": "learned"}


def sweep_rejection_performance(ev, out, quick):
    """plotRejectionPerformance (PlotConstruction.cuh:18-34): sine weights with
    item 0 raised to weight0, the mean time of fillA alone after one build
    (RejectionSampling::benchmarkBuild, RejectionSampling.cu:58-70), per fill."""
    out.write("method;N;weight0;time\n")
    for n in ([1_000_000] if quick else [1_000_000, 10_000_000]):
        base = generate_weights("sine", n)
        for weight0 in ([1e6, 2e7] if quick else range(1_000_000, 20_000_001, 1_000_000)):
            w = base.copy()
            w[0] = weight0
            t = make_table(n, w=w)
            for fill, name in REJECTION_FILL_NAMES.items():
                t.rejection_build(ev.stream, fill=fill)
                ms = ev.time(lambda: t.rejection_fill(ev.stream), ITERATIONS)
                out.write(f"{name};{n};{int(weight0)};{ms}\n")
            t.close()


SWEEPS = {
    "build_soa": sweep_build_soa,
    "build_distribution": sweep_build_distribution,
    "thread_number": sweep_thread_number,
    "split_variants": sweep_split_variants,
    "with_transfer": sweep_with_transfer,
    "light_vs_heavy": sweep_light_vs_heavy,
    "sampler_method": sweep_sampler_method,
    "sampling_distribution_speed": sweep_sampling_distribution_speed,
    "number_of_samples": sweep_number_of_samples,
    "sampling_speed": sweep_sampling_speed,
    "distribution": sweep_distribution,
    "rejection_performance": sweep_rejection_performance,
}


def run(name, out, quick=False):
    ev = _Events(_torch())
    SWEEPS[name](ev, out, quick)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("sweep", choices=sorted(SWEEPS))
    ap.add_argument("--quick", action="store_true", help="a few points of each grid")
    ap.add_argument("--out", default="-")
    args = ap.parse_args(argv)
    out = sys.stdout if args.out == "-" else open(args.out, "w")
    try:
        run(args.sweep, out, args.quick)
    finally:
        if out is not sys.stdout:
            out.close()


if __name__ == "__main__":
    main()
