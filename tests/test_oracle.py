"""Pins of the CPU oracle against facts that do not come from the oracle itself:
published known-answer vectors, an independent Philox implementation, closed
forms, exact (integer / dyadic) arithmetic where every association agrees, the
faithful restatement of the reference's CPU build, brute force on tiny inputs,
and statistical tests of the samplers."""
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy import stats as sstats

import atg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key, expected).
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,exp", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, exp):
    assert oracle.philox(ctr, key) == exp


def test_philox_matches_rocrand(oracle, tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "pxc"
    r = subprocess.run([hipcc, "-O1", "-x", "hip", "--cuda-host-only",
                        os.path.join(ROOT, "tests", "golden", "philox_xcheck.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    rng = np.random.default_rng(1)
    cases = rng.integers(0, 2**32, size=(200, 6), dtype=np.uint64)
    inp = "\n".join(" ".join(f"{int(v):x}" for v in c) for c in cases) + "\n"
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for c, line in zip(cases, out):
        got = oracle.philox([int(v) for v in c[:4]], [int(v) for v in c[4:]])
        assert [int(x, 16) for x in line.split()] == got


# ---- canonical sums --------------------------------------------------------

def test_section_layout_matches_reference_formula(oracle):
    # gridDim = N / sectionSize + 1, section length N / gridDim + 1
    # (AliasTableSplitGpu.cu:356-357, PsaPlusNormal.cu:27-29).
    for n, s in [(1, 1), (10, 1), (100, 10), (10**8, 1000), (10**8, 4096), (777, 1000)]:
        G, Ls = oracle.section_layout(n, s)
        assert G == n // s + 1 and Ls == n // G + 1
        assert G * Ls >= n and Ls <= max(s, 1)


def test_sum_closed_form(oracle):
    # Adder::unitTest (Adder.cu:159-171): sum_{i<1000} i = 499500.
    w = np.arange(1000, dtype=np.float64)
    for s in (1, 7, 100, 1000, 4096):
        assert oracle.weight_sum(w, s)[0] == 499500.0


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4095, 100003, 1 << 20])
def test_sum_exact_on_integers(oracle, n):
    # Integer weights with partial sums < 2^53: every association is exact.
    w = np.random.default_rng(n).integers(0, 1 << 20, size=n).astype(np.float64)
    for s in (1, 100, 1000, 4096):
        W, sums = oracle.weight_sum(w, s)
        assert W == float(int(w.astype(np.int64).sum()))


def _pairwise64(v):
    """Pairwise trees of consecutive groups of 64 (zero-padded): t[m] = t[m] +
    t[m + width] for width 1, 2, 4, ..., 32; returns t[0] of every group."""
    v = np.concatenate([v, np.zeros((-len(v)) % 64)])
    t = v.reshape(-1, 64).copy()
    width = 1
    while width < 64:
        t[:, 0::2 * width] = t[:, 0::2 * width] + t[:, width::2 * width]
        width *= 2
    return t[:, 0]


def _canonical_sum_numpy(w, section_param, reverse_lanes=False):
    """The canonical W written out with numpy from its definition: sections
    of the PSA+ layout; in a section lane l adds items l, l + 64, ... left to
    right; the 64 lane sums meet in a pairwise tree; the section sums meet in
    recursive pairwise trees of 64.  Returns W and the section sums.
    reverse_lanes: each lane adds its items right to left instead (a changed
    association, to show that the check below can see one)."""
    n = len(w)
    G = n // section_param + 1
    Ls = n // G + 1
    sums = np.empty(G)
    for s in range(G):
        x = w[min(n, s * Ls):min(n, (s + 1) * Ls)]
        acc = np.zeros(64)
        starts = list(range(0, len(x), 64))
        for k in (starts[::-1] if reverse_lanes else starts):
            chunk = x[k:k + 64]
            acc[:len(chunk)] = acc[:len(chunk)] + chunk
        sums[s] = _pairwise64(acc)[0]
    top = sums.copy()
    while len(top) > 1:
        top = _pairwise64(top)
    return top[0], sums


def _sequential_sum(w):
    s = 0.0
    for v in w.tolist():
        s = s + v
    return s


@pytest.mark.parametrize("n,section_param", [(2583, 10**6), (64 * 64 * 5 + 17, 997), (300_007, 1000)])
def test_sum_association_fixed(oracle, n, section_param):
    # Values spread over 2^24 (no single term dominates a section, but the
    # order of the additions shows in the low bits): the oracle's W and every
    # section sum equal the definition written out independently with numpy,
    # while a plain left-to-right sum (the reference's sumWeightsCpu,
    # SamplingAlgorithm.cu:14-19) and lanes summed right to left give other
    # bits, so a changed association fails this check.
    rng = np.random.default_rng(n)
    w = rng.random(n) * 2.0 ** rng.integers(0, 24, n)
    W, sums = oracle.weight_sum(w, section_param)
    W_np, sums_np = _canonical_sum_numpy(w, section_param)
    assert np.array_equal(sums.view(np.uint64), sums_np.view(np.uint64))
    assert W == W_np
    assert W != _sequential_sum(w)
    _, sums_rev = _canonical_sum_numpy(w, section_param, reverse_lanes=True)
    assert np.any(sums_rev != sums)
    G, Ls = oracle.section_layout(n, section_param)
    if G == 1:
        assert Ls > n  # one section holds every item


# ---- construction ------------------------------------------------------------

def _verify(oracle, w, ref):
    return oracle.verify(w, ref["stats"]["W"], ref["stats"]["W_N"], ref["row_w"], ref["row_alias"])


@pytest.mark.parametrize("n,p", [(1000, 100), (10000, 55), (65536, 1024), (100003, 1400)])
def test_plain_build_equals_reference_cpu_build_on_integers(oracle, n, p):
    # With integer weights every sum is exact, so the canonical build without
    # PSA+ must reproduce the faithful restatement of AliasTableSplit::build
    # (sequential W and scans, SplitMethodBasic, PackMethodSweep) bit for bit.
    w = np.random.default_rng(n).integers(0, 64, size=n).astype(np.float64)
    ref = oracle.build(w, psa_plus=False, num_splits=p)
    W, rw, ra = oracle.ref_split_build(w, p)
    assert W == ref["stats"]["W"]
    assert np.array_equal(rw, ref["row_w"])
    assert np.array_equal(ra, ref["row_alias"])


@pytest.mark.parametrize("psa", [False, True])
@pytest.mark.parametrize("dist", ["pareto", "unit_uniform", "exponential", "sine", "powerlaw1_shuffled",
                                  "heavy_mix", "zipf_shuffled", "ramp", "powerlaw2", "uniform"])
def test_tables_pass_reference_check(oracle, dist, psa):
    n = 200003
    w = atg.generate_weights(dist, n, seed=1, param=1.1 if dist == "pareto" else 0.0)
    ref = oracle.build(w, psa_plus=psa, num_splits=81920)
    v = _verify(oracle, w, ref)
    assert v["reference_ok"], v
    # The reference's absolute EPSILON (1e-10, SamplingAlgorithm.cuh:14) may
    # finish a heavy row up to EPSILON above W/N and move that mass to the next
    # heavy; spills carry the rounding of prefix sums of magnitude W.
    W_N = ref["stats"]["W_N"]
    assert v["max_abs"] < 2 * 1e-10 + 64 * np.spacing(ref["stats"]["W"]), v
    if W_N > 1e-3:
        assert v["max_rel"] < 1e-7, v
    assert ref["stats"]["split_fallbacks"] == 0


def corrupt_row_weight(ref):
    """A copy of the table with one row weight moved by a quarter of W/N (a
    row well below W/N whose alias is an item): mass moves between the row's
    item and its alias."""
    rw, ra = ref["row_w"].copy(), ref["row_alias"].copy()
    W_N = ref["stats"]["W_N"]
    i = int(np.nonzero((rw < 0.5 * W_N) & (ra >= 0) & (ra < len(rw)) & (ra != np.arange(len(rw))))[0][0])
    rw[i] += 0.25 * W_N
    return rw, ra


def corrupt_alias(ref):
    """A copy of the table with one alias pointed at another item (a row whose
    alias carries at least W/N / 4 of mass)."""
    rw, ra = ref["row_w"].copy(), ref["row_alias"].copy()
    W_N = ref["stats"]["W_N"]
    n = len(rw)
    i = int(np.nonzero((rw < 0.75 * W_N) & (ra >= 0) & (ra < n))[0][0])
    ra[i] = (ra[i] + 1) % n if (ra[i] + 1) % n != i else (ra[i] + 2) % n
    return rw, ra


@pytest.mark.parametrize("corrupt", [corrupt_row_weight, corrupt_alias])
def test_verify_rejects_corrupted_tables(oracle, corrupt):
    # The table check that every PSA+ pin relies on must fail on a wrong row:
    # the reference criterion (AliasTable.cu:24-65) and the relative error.
    n = 20000
    w = atg.generate_weights("sine", n)
    ref = oracle.build(w, num_splits=1000)
    good = _verify(oracle, w, ref)
    assert good["reference_ok"] and good["max_rel"] < 1e-9
    rw, ra = corrupt(ref)
    bad = oracle.verify(w, ref["stats"]["W"], ref["stats"]["W_N"], rw, ra)
    assert not bad["reference_ok"], bad
    assert bad["max_rel"] > 1e-3, bad


PURPOSE_BASIC = 0x0000B001


def basic_row_and_coin(oracle, seed, k, n):
    """Row and coin of basic sample k: Philox4x32-10 (pinned above) at counter
    (k, PURPOSE_BASIC), key = seed; row = floor(x64 n / 2^64) from words 0-1,
    coin = the top 53 bits of words 2-3 as a double in [0, 1)."""
    x = oracle.philox([k & 0xFFFFFFFF, k >> 32, PURPOSE_BASIC, 0], [seed & 0xFFFFFFFF, seed >> 32])
    row = (((x[1] << 32) | x[0]) * n) >> 64
    coin = (((x[3] << 32) | x[2]) >> 11) * 2.0 ** -53
    return row, coin


def decision_boundary_table(oracle, n, seed, count):
    """A table of n rows with W/N = 1 built around the first `count` basic
    samples: the row each sample hits gets one of five rows that sit on the
    boundaries of the rule of SamplerGpuBasic.cu:28-45 (take the alias iff
    row.w < W/N and coin * W/N >= row.w; an alias that is not an item keeps
    the row).  Returns the rows and, for every sample alone on its row, the
    item the rule must return."""
    rw = np.ones(n)
    ra = np.arange(n, dtype=np.int32)
    hits = [basic_row_and_coin(oracle, seed, k, n) for k in range(count)]
    rows = [r for r, _ in hits]
    expected = {}
    for k, (r, coin) in enumerate(hits):
        if rows.count(r) > 1:
            continue
        alias = (r + 1 + k) % n
        cat = k % 5
        if cat == 0:    # coin * W/N == row.w exactly: alias (>=)
            rw[r], ra[r], exp = coin, alias, alias
        elif cat == 1:  # row.w one ulp above coin * W/N: row
            rw[r], ra[r], exp = np.nextafter(coin, 2.0), alias, r
        elif cat == 2:  # row.w == W/N: row (strict <)
            rw[r], ra[r], exp = 1.0, alias, r
        elif cat == 3:  # always-alias weight but the alias is the heavy sentinel N: row
            rw[r], ra[r], exp = 0.0, n, r
        else:           # row.w = 0: alias whatever the coin
            rw[r], ra[r], exp = 0.0, alias, alias
        expected[k] = exp
    return rw, ra, expected


def test_basic_sampler_decision_boundaries(oracle):
    n, seed, count = 1 << 16, 77, 200
    rw, ra, expected = decision_boundary_table(oracle, n, seed, count)
    assert len(expected) > 150
    out = oracle.sample_basic(rw, ra, 1.0, seed, count)
    for k, exp in expected.items():
        assert out[k] == exp, (k, k % 5, out[k], exp)


@pytest.mark.parametrize("k", [10, 14, 16])
def test_dyadic_tables_are_exact(oracle, k):
    # N = 2^k integer weights: W/N is exact and every spill is exact, so the
    # implied probabilities equal the weights exactly.
    n = 1 << k
    w = np.random.default_rng(k).integers(0, 32, size=n).astype(np.float64)
    for psa in (False, True):
        ref = oracle.build(w, psa_plus=psa, num_splits=max(2, n // 64))
        v = _verify(oracle, w, ref)
        assert v["max_rel"] == 0.0 and v["reference_ok"]


def test_closed_form_tables(oracle):
    n = 10000
    # All equal: every item is light, one leftover class -> rows {w, self}.
    ref = oracle.build(np.full(n, 2.5))
    assert ref["stats"]["degenerate"] == 1
    assert np.all(ref["row_w"] == 2.5) and np.array_equal(ref["row_alias"], np.arange(n))
    # OneHeavy (SamplingAlgorithm.cu:37-39): item 0 = 1.5, rest 1; exact masses.
    w = atg.generate_weights("one_heavy", n)
    for psa in (False, True):
        ref = oracle.build(w, psa_plus=psa)
        v = _verify(oracle, w, ref)
        assert v["reference_ok"] and v["max_rel"] < 1e-12
    # Regression case of UnitTests.cu:47-49: N=505711, p=16576, OneLight.
    w = atg.generate_weights("one_light", 505711)
    ref = oracle.build(w, psa_plus=False, num_splits=16576)
    assert _verify(oracle, w, ref)["reference_ok"]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 9, 10, 11, 25])
def test_tiny_inputs(oracle, n):
    rng = np.random.default_rng(n)
    for _ in range(20):
        w = rng.random(n) + 0.01
        for psa in (False, True):
            ref = oracle.build(w, psa_plus=psa)
            assert np.all(ref["row_alias"] >= 0), "every row written"
            assert _verify(oracle, w, ref)["reference_ok"]


def test_split_bisection_matches_brute_force(oracle):
    # For small leftover lists, enumerate every j in [0, min(n', numH) - 1] and
    # collect the positions where the FOUND predicate of SplitMethodBasic.cuh:23-39
    # holds; with a strictly increasing sigma there is exactly one and the
    # bisection must return it.
    rng = np.random.default_rng(11)
    for trial in range(50):
        n = int(rng.integers(20, 400))
        w = rng.random(n) * rng.choice([1, 10, 100], size=n)
        ref = oracle.build(w, psa_plus=bool(trial % 2), num_splits=int(rng.integers(2, 40)))
        st = ref["stats"]
        if st["p"] == 0:
            continue
        Lp, Hp, Hw = ref["L_pref"], ref["H_pref"], ref["H_w"]
        numL, numH, p = st["numL"], st["numH"], st["p"]
        for k in range(1, p):
            n_ = int(np.ceil((st["nToSplit"] * k) / p))
            target = (st["wToSplit"] * n_) / st["nToSplit"]
            found = []
            for j in range(0, min(n_, numH)):
                i = min(n_ - j, numL - 1)
                if j + 1 >= numH or i - 1 < 0 or i >= numL:
                    continue
                s, sn = Lp[i] + Hp[j], Lp[i - 1] + Hp[j + 1]
                if s <= target < sn:
                    found.append((i, j))
            assert len(found) == 1
            assert (ref["split_i"][k], ref["split_j"][k]) == found[0]


def test_psa_plus_off_reproduces_plain(oracle):
    w = atg.generate_weights("pareto", 100003, param=1.1)
    a = oracle.build(w, psa_plus=False)
    b = oracle.build(w, psa_plus=True, chains=16)
    assert a["stats"]["R"] == 100003 and b["stats"]["R"] < 100003
    assert a["stats"]["W"] == b["stats"]["W"]


def test_multithreaded_oracle_is_bit_identical(oracle):
    w = atg.generate_weights("pareto", 300007, param=1.1)
    a = oracle.build(w, threads=1)
    b = oracle.build(w, threads=8)
    for k in ("row_w", "row_alias", "L_pref", "H_pref", "split_spill"):
        assert np.array_equal(np.asarray(a[k]).view(np.uint8), np.asarray(b[k]).view(np.uint8)), k


def test_cpu_baseline_builds_are_valid(oracle):
    w = atg.generate_weights("sine", 10000)
    for fn in (oracle.ref_sweep_build, oracle.ref_stack_build):
        W, rw, ra = fn(w)
        v = oracle.verify(w, W, W / len(w), rw, ra)
        assert v["reference_ok"], fn.__name__


# ---- samplers --------------------------------------------------------------

def _chi2(counts, w):
    expected = w / w.sum() * counts.sum()
    keep = expected >= 5
    obs = np.append(counts[keep], counts[~keep].sum())
    exp = np.append(expected[keep], expected[~keep].sum())
    if exp[-1] == 0:
        obs, exp = obs[:-1], exp[:-1]
    return sstats.chisquare(obs, exp * obs.sum() / exp.sum()).pvalue


def test_samplers_chi_squared(oracle):
    n = 5000
    w = atg.generate_weights("sine", n)
    ref = oracle.build(w)
    S = 2_000_000
    b = oracle.sample_basic(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], 7, S)
    s = oracle.sample_sectioned(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], 512, S, 7)
    W = oracle.weight_sum(w, 500)[0]
    incl, A = oracle.rejection_build(w, W)
    r, trials = oracle.sample_rejection(A, w, W, 7, S)
    for out in (b, s, r):
        assert _chi2(np.bincount(out, minlength=n).astype(float), w) > 1e-4
    assert trials / S < 2.0


def test_section_counts_exact_total_and_distribution(oracle):
    n, rows, total = 1_000_003, 4096, 10**9
    c = oracle.section_counts(n, rows, total, seed=3).astype(np.float64)
    assert c.sum() == total
    lens = np.minimum(n, (np.arange(len(c)) + 1) * rows) - np.arange(len(c)) * rows
    assert _chi2(c, lens.astype(float)) > 1e-4
    # small totals go through exact Bernoulli trials
    c = oracle.section_counts(1000, 64, 7, seed=1)
    assert c.sum() == 7


def test_binomial_moments(oracle):
    # Normal-approximation branch (n > 32) and the exact branch (n <= 32).
    for n, q in [(10_000, 0.3), (20, 0.5)]:
        x = np.array([oracle.binomial(n, q, seed=9, node=i) for i in range(20000)], dtype=float)
        assert abs(x.mean() - n * q) < 4 * np.sqrt(n * q * (1 - q) / len(x)) + 1e-9
        assert abs(x.var() / (n * q * (1 - q)) - 1) < 0.05


def test_rejection_array_brute_force(oracle):
    w = atg.generate_weights("pareto", 5000, param=1.1)
    W = float(np.sum(w))
    incl, A = oracle.rejection_build(w, W)
    counts = np.floor((len(w) * w) / W).astype(np.int64) + 1
    assert np.array_equal(incl, np.cumsum(counts))
    assert np.array_equal(A, np.repeat(np.arange(len(w)), counts))


def _outlier_sine(n, weight0):
    # The reference's rejection construction plot: sine weights with item 0
    # raised to weightOfOutlier (RejectionSampling.cu:58-60, PlotConstruction.cuh:18-27).
    w = atg.generate_weights("sine", n)
    w[0] = weight0
    return w


@pytest.mark.parametrize("case", ["pareto", "outlier", "uniform", "one_heavy", "tiny"])
def test_rejection_fill_bisect_equals_runs(oracle, case):
    # The reference's per-slot bisection (RejectionSamplingBinarySearch.cu:12-33)
    # and the run-writing fill give one A, also with a run of most of A.
    w = {"pareto": lambda: atg.generate_weights("pareto", 30000, param=1.1),
         "outlier": lambda: _outlier_sine(10000, 2e7),
         "uniform": lambda: np.full(4096, 0.5),
         "one_heavy": lambda: atg.generate_weights("one_heavy", 5000),
         "tiny": lambda: np.array([3.0])}[case]()
    W = float(np.sum(w))
    incl, A = oracle.rejection_build(w, W)
    assert np.array_equal(oracle.rejection_fill_bisect(incl), A)


def _model_pred(q, A_size, K, s):
    """pred(s) of the inverse-CDF model from its knot items (definition in
    oracle_rejection_model), vectorised, and the segment of each slot."""
    M = A_size - 1
    s = s.astype(np.uint64)
    k = np.zeros_like(s) if M == 0 else np.minimum(K - 1, s * np.uint64(K) // np.uint64(M))
    kslot = lambda kk: (kk * np.uint64(M) + np.uint64(K - 1)) // np.uint64(K)
    sk, sk1 = kslot(k), kslot(k + np.uint64(1))
    qk, qk1 = q[k.astype(np.int64)].astype(np.uint64), q[(k + np.uint64(1)).astype(np.int64)].astype(np.uint64)
    span = np.maximum(sk1 - sk, np.uint64(1))
    pred = np.where(sk1 > sk, qk + (s - sk) * (qk1 - qk) // span, qk)
    return k.astype(np.int64), pred.astype(np.int64)


@pytest.mark.parametrize("case,K", [("pareto", 64), ("pareto", 4096), ("outlier", 37), ("heavy_mix", 512),
                                    ("tiny", 8), ("uniform", 1000)])
def test_rejection_model_knots_and_error_bounds(oracle, case, K):
    # The knots are the items of the knot slots (a searchsorted over the
    # inclusive counts); every slot's item lies within its segment's bound
    # of the prediction, and every non-empty segment attains its bound.
    w = {"pareto": lambda: atg.generate_weights("pareto", 40000, param=1.1),
         "outlier": lambda: _outlier_sine(3000, 1e6),
         "heavy_mix": lambda: atg.generate_weights("heavy_mix", 50000),
         "tiny": lambda: np.array([1.0, 2.0]),
         "uniform": lambda: np.ones(777)}[case]()
    W = float(np.sum(w))
    incl, A = oracle.rejection_build(w, W)
    q, e = oracle.rejection_model(incl, K)
    M = len(A) - 1
    ks = (np.arange(K + 1, dtype=np.uint64) * np.uint64(M) + np.uint64(K - 1)) // np.uint64(K)
    assert np.array_equal(q, np.searchsorted(incl, ks, side="right"))
    assert np.array_equal(q, A[ks.astype(np.int64)])
    k, pred = _model_pred(q, len(A), K, np.arange(len(A)))
    dev = np.abs(A.astype(np.int64) - pred)
    worst = np.zeros(K, np.int64)
    np.maximum.at(worst, k, dev)
    assert np.array_equal(worst, e.astype(np.int64))


# ---- batches (test infrastructure for the engine's batched build) ------------

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_oracle_equals_single_builds(oracle, dtype):
    sizes = [1, 2, 17, 100, 1000, 4097, 20011, 65536]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    w = np.concatenate([atg.generate_weights("exponential", n, seed=5 + t) for t, n in enumerate(sizes)]).astype(dtype)
    kw = dict(chains=32, section_size=2048, num_splits=-38)
    rw, ra, W = oracle.build_batch(w, offsets, **kw)
    for t, n in enumerate(sizes):
        lo, hi = offsets[t], offsets[t + 1]
        ref = oracle.build(w[lo:hi].astype(np.float64), chains=32, section_size=2048,
                           num_splits=max(2, n // 38), want_lists=False)
        assert W[t] == ref["stats"]["W"]
        assert np.array_equal(rw[lo:hi], ref["row_w"]) and np.array_equal(ra[lo:hi], ref["row_alias"]), t
    # a subset leaves the other tables untouched
    rw2, ra2, W2 = oracle.build_batch(w, offsets, which=[7, 3], **kw)
    for t in range(len(sizes)):
        lo, hi = offsets[t], offsets[t + 1]
        if t in (3, 7):
            assert np.array_equal(rw2[lo:hi], rw[lo:hi]) and W2[t] == W[t]
        else:
            assert np.all(ra2[lo:hi] == -1)
    Wc, ok, mabs = oracle.check_batch(w, offsets, rw, ra, W, **kw)
    assert np.array_equal(Wc, W) and ok.all()
    # a changed alias in one table fails that table's check only
    t = 6
    ref6 = {"row_w": rw[offsets[t]:offsets[t + 1]], "row_alias": ra[offsets[t]:offsets[t + 1]],
            "stats": {"W_N": W[t] / sizes[t]}}
    bw, ba = corrupt_alias(ref6)
    ra_bad = ra.copy()
    ra_bad[offsets[t]:offsets[t + 1]] = ba
    _, ok, _ = oracle.check_batch(w, offsets, rw, ra_bad, W, **kw)
    assert not ok[t] and ok.sum() == len(sizes) - 1


def test_sectioned_subset_equals_full_run(oracle):
    n = 100_003
    w = atg.generate_weights("pareto", n, seed=2, param=1.1)
    ref = oracle.build(w, want_lists=False)
    total, rows, seed = 3_000_000, 4096, 5
    full = oracle.sample_sectioned(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], rows, total, seed)
    nsec = (n + rows - 1) // rows
    sel = [0, 5, nsec - 1, 11]
    out, first, cnt = oracle.sample_sectioned_subset(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], rows,
                                                     total, seed, sel)
    pos = 0
    for f, c in zip(first.tolist(), cnt.tolist()):
        assert np.array_equal(out[pos:pos + c], full[f:f + c])
        pos += c
    assert pos == len(out)


# ---- sectioned rejection sampling ------------------------------------------

def test_rejection_sectioned_masses_and_distribution(oracle):
    n = 20000
    w = atg.generate_weights("pareto", n, seed=9, param=1.1)
    W = oracle.weight_sum(w, 1000)[0]
    incl, A = oracle.rejection_build(w, W)
    SS = 1024
    pre = oracle.rejection_section_prefix(A, w, W, SS)
    # The acceptance masses add up to N (sum over items of N w_i / W) up to
    # the 2^-53 coin granularity and rounding.
    assert abs(pre[-1] - n) < 1e-9 * n
    assert np.all(np.diff(pre) > 0)
    # Section masses against a direct numpy statement: first slot of a run
    # accepts with frac(N w / W), the others always.
    first = np.ones(len(A), bool)
    first[1:] = A[1:] != A[:-1]
    xv = (n * w[A]) / W
    acc = np.where(first, xv - np.floor(xv), 1.0)
    prob = np.where(acc >= 1.0, 1.0, (np.floor(acc * 2.0**53) + 1.0) * 2.0**-53)
    masses = np.add.reduceat(prob, np.arange(0, len(A), SS))
    np.testing.assert_allclose(np.diff(pre), masses, rtol=1e-12)
    # Exact totals and a chi-squared fit of the samples to the weights.
    c = oracle.rejection_section_counts(pre, 3_000_000, 5)
    assert int(c.sum()) == 3_000_000
    out = oracle.sample_rejection_sectioned(A, w, W, SS, 2_000_000, 7)
    assert _chi2(np.bincount(out, minlength=n).astype(float), w) > 1e-4
    # A subset of sections reproduces the full run at the same positions.
    sel = [0, 3, len(pre) - 2]
    sub, f, k = oracle.sample_rejection_sectioned(A, w, W, SS, 2_000_000, 7, sections=sel)
    for a, b in zip(f.tolist(), k.tolist()):
        assert np.array_equal(sub[a:a + b], out[a:a + b])


def test_split_bisection_matches_inverse_split(oracle):
    # SplitMethodInverse (SplitMethodInverse.cu:3-31) as a second brute force:
    # where the split certificate holds (max light increment < min heavy
    # increment: one FOUND position per split), its (i, j, spill) equal the
    # bisection's bit for bit, for plain and PSA+ leftovers.
    rng = np.random.default_rng(23)
    checked = 0
    for trial in range(60):
        n = int(rng.integers(50, 3000))
        dist = ["pareto", "exponential", "sine", "heavy_mix", "unit_uniform"][trial % 5]
        w = atg.generate_weights(dist, n, seed=100 + trial, param=1.1 if dist == "pareto" else 0.9)
        ref = oracle.build(w, psa_plus=bool(trial % 2), num_splits=int(rng.integers(2, 60)))
        st = ref["stats"]
        if st["p"] == 0 or st["monotone"] != 1:
            continue
        si, sj, ss = oracle.split_inverse(st["p"], st["nToSplit"], st["wToSplit"], ref["L_pref"], ref["H_pref"],
                                          ref["H_w"])
        assert np.array_equal(si, ref["split_i"]) and np.array_equal(sj, ref["split_j"]), trial
        assert np.array_equal(ss.view(np.uint64), ref["split_spill"].view(np.uint64)), trial
        checked += 1
    assert checked >= 30
