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


def test_sum_tree_shape(oracle):
    # 64 items in one section: the lane sums are the items themselves and the
    # total is the pairwise tree ((x0+x1)+(x2+x3))+... .  Built with numpy.
    x = np.random.default_rng(3).random(64) * np.logspace(0, 15, 64)
    t = x.copy()
    while len(t) > 1:
        t = t[0::2] + t[1::2]
    W, _ = oracle.weight_sum(x, 64)
    assert W == t[0]


def test_scan_closed_form(oracle):
    # PrefixSum::unitTest (PrefixSum.cu:265-296): input i % 7, exclusive prefix.
    v = (np.arange(1000) % 7).astype(np.float64)
    ex, agg = oracle.local_scan(v)
    assert np.array_equal(ex, np.concatenate([[0.0], np.cumsum(v)[:-1]]))
    assert agg == v.sum()


def test_scan_total_is_pairwise_chunk_tree(oracle):
    x = np.random.default_rng(5).random(256) * np.logspace(0, 12, 256)
    _, agg = oracle.local_scan(x)
    carry = 0.0
    for c in range(4):
        t = x[64 * c:64 * (c + 1)].copy()
        while len(t) > 1:
            t = t[0::2] + t[1::2]
        carry = carry + t[0]
    assert agg == carry


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
