"""GPU sampling parity: basic, sectioned and rejection samplers produce the
oracle's items on the same Philox stream; histograms pass chi-squared."""
import numpy as np
import pytest
from scipy import stats as sstats

import atg

pytestmark = pytest.mark.gpu


def build_pair(oracle, n, dist="pareto", seed=42, **kw):
    w = atg.generate_weights(dist, n, seed=seed, param=1.1 if dist == "pareto" else 0.0)
    t = atg.Table(n, **kw)
    t.set_weights(w)
    t.build()
    t.sync()
    ref = oracle.build(w, num_splits=81920, want_lists=False)
    tw, ta = t.table()
    assert np.array_equal(tw.view(np.uint64), ref["row_w"].view(np.uint64))
    assert np.array_equal(ta, ref["row_alias"])
    return w, t, ref


def chi2_pvalue(counts, w):
    expected = w / w.sum() * counts.sum()
    keep = expected >= 5
    obs = np.append(counts[keep], counts[~keep].sum())
    exp = np.append(expected[keep], expected[~keep].sum())
    if exp[-1] == 0:
        obs, exp = obs[:-1], exp[:-1]
    return sstats.chisquare(obs, exp * obs.sum() / exp.sum()).pvalue


@pytest.mark.parametrize("n,count", [(1000, 100000), (100003, 2000000), (1 << 20, 3000000)])
def test_basic_matches_oracle(gpu, oracle, n, count):
    torch = gpu
    w, t, ref = build_pair(oracle, n)
    out = torch.empty(count, dtype=torch.int32, device="cuda")
    cs = t.sample(count, seed=123, kind="basic", out=out, checksum=True)
    got = out.cpu().numpy()
    exp = oracle.sample_basic(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], 123, count)
    np.testing.assert_array_equal(got, exp)
    assert cs == int(exp.astype(np.int64).sum())


@pytest.mark.parametrize("n,count,rows", [(1000, 100000, 64), (100003, 2000000, 4096), (1 << 20, 3000000, 4096),
                                          (777777, 1000, 1024), (5000, 7, 4096)])
def test_sectioned_matches_oracle(gpu, oracle, n, count, rows):
    torch = gpu
    w, t, ref = build_pair(oracle, n)
    out = torch.empty(max(count, 1), dtype=torch.int32, device="cuda")
    t.sample(count, seed=99, kind="sectioned", section_rows=rows, out=out)
    got = out.cpu().numpy()[:count]
    exp = oracle.sample_sectioned(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], rows, count, 99)
    np.testing.assert_array_equal(got, exp)


def test_sectioned_section_sizes_on_one_table(gpu, oracle):
    """Calls with more sections than any before on the same table (the
    per-section counts buffer grows), then fewer again."""
    torch = gpu
    n, count = 100003, 500000
    w, t, ref = build_pair(oracle, n)
    out = torch.empty(count, dtype=torch.int32, device="cuda")
    for rows, seed in ((4096, 7), (64, 8), (4096, 9), (1, 10)):
        t.sample(count, seed=seed, kind="sectioned", section_rows=rows, out=out)
        got = out.cpu().numpy()
        exp = oracle.sample_sectioned(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], rows, count, seed)
        np.testing.assert_array_equal(got, exp, err_msg=f"section_rows={rows}")


def test_histograms_chi_squared(gpu, oracle):
    torch = gpu
    n = 10000
    w = atg.generate_weights("sine", n)
    t = atg.Table(n)
    t.set_weights(w)
    t.build()
    t.sync()
    for kind in ("basic", "sectioned"):
        hist = torch.zeros(n, dtype=torch.int32, device="cuda")
        t.sample(20_000_000, seed=5, kind=kind, section_rows=2500, hist=hist)
        h = hist.cpu().numpy().astype(np.float64)
        assert h.sum() == 20_000_000
        assert chi2_pvalue(h, w) > 1e-4, kind
    t.rejection_build()
    hist = torch.zeros(n, dtype=torch.int32, device="cuda")
    t.rejection_sample(20_000_000, seed=6, hist=hist)
    h = hist.cpu().numpy().astype(np.float64)
    assert chi2_pvalue(h, w) > 1e-4


@pytest.mark.parametrize("n,dist", [(1000, "sine"), (100003, "pareto"), (1 << 20, "exponential"),
                                    (50_000, "one_heavy"), (300_001, "heavy_mix"), (1, "unit_uniform"),
                                    (4097, "powerlaw2")])
def test_rejection_matches_oracle(gpu, oracle, n, dist):
    torch = gpu
    w = atg.generate_weights(dist, n, seed=11, param=1.1 if dist == "pareto" else 0.0)
    t = atg.Table(n)
    t.set_weights(w)
    t.rejection_build()
    W = oracle.weight_sum(w, max(1, min(1000, n // 10)))[0]
    incl, A = oracle.rejection_build(w, W)
    assert t.rejection_size() == len(A)
    np.testing.assert_array_equal(t.rejection_array(), A)
    count = 1_000_000
    out = torch.empty(count, dtype=torch.int32, device="cuda")
    t.rejection_sample(count, seed=8, out=out)
    exp, _ = oracle.sample_rejection(A, w, W, 8, count)
    np.testing.assert_array_equal(out.cpu().numpy(), exp)


@pytest.mark.parametrize("n,dist", [(100003, "pareto"), (70001, "one_heavy")])
def test_rejection_f32_weights(gpu, oracle, n, dist):
    """FP32 weights: the counts and acceptance fractions use the exact FP64
    values of the FP32 inputs."""
    torch = gpu
    w = atg.generate_weights(dist, n, seed=13, param=1.1 if dist == "pareto" else 0.0).astype(np.float32)
    w64 = w.astype(np.float64)
    t = atg.Table(n, dtype="f32")
    t.set_weights(w)
    t.rejection_build()
    W = oracle.weight_sum(w64, max(1, min(1000, n // 10)))[0]
    incl, A = oracle.rejection_build(w64, W)
    assert t.rejection_size() == len(A)
    np.testing.assert_array_equal(t.rejection_array(), A)
    count = 500_000
    out = torch.empty(count, dtype=torch.int32, device="cuda")
    t.rejection_sample(count, seed=21, out=out)
    exp, _ = oracle.sample_rejection(A, w64, W, 21, count)
    np.testing.assert_array_equal(out.cpu().numpy(), exp)


def test_basic_decision_boundaries(gpu, oracle):
    """Hand-built rows on the boundaries of the rule of SamplerGpuBasic.cu:28-45
    (a row weight equal to coin * W/N, one ulp above it, equal to W/N, an
    always-alias weight whose alias is the heavy sentinel N, and a zero weight),
    loaded with atg_set_table: the GPU returns the items the rule fixes and
    the oracle's items."""
    import test_oracle as pins
    torch = gpu
    n, seed, count = 1 << 16, 77, 200
    rw, ra, expected = pins.decision_boundary_table(oracle, n, seed, count)
    t = atg.Table(n)
    t.set_table(rw, ra, float(n))  # W/N = 1
    out = torch.empty(count, dtype=torch.int32, device="cuda")
    t.sample(count, seed=seed, kind="basic", out=out)
    got = out.cpu().numpy()
    for k, exp in expected.items():
        assert got[k] == exp, (k, k % 5, got[k], exp)
    np.testing.assert_array_equal(got, oracle.sample_basic(rw, ra, 1.0, seed, count))


@pytest.mark.parametrize("n,dist,slots,count", [(100003, "pareto", 4096, 3_000_000), (1 << 20, "exponential", 4096, 4_000_000),
                                                (50000, "one_heavy", 1024, 1_000_000), (4097, "powerlaw2", 8192, 500_000),
                                                (300001, "heavy_mix", 2048, 2_000_000), (1, "unit_uniform", 4096, 1000)])
def test_rejection_sectioned_matches_oracle(gpu, oracle, n, dist, slots, count):
    """Rejection sampling by LDS-staged sections of A: the oracle's items on
    the same Philox stream (section masses, prefix, counts and in-section
    trials all exact), twice on one structure (cached masses), then with
    another section size."""
    torch = gpu
    w = atg.generate_weights(dist, n, seed=5, param=1.1 if dist == "pareto" else 0.9)
    t = atg.Table(n)
    t.set_weights(w)
    t.rejection_build()
    W = t.rejection_weight_sum()
    _, A = oracle.rejection_build(w, W)
    out = torch.empty(count, dtype=torch.int32, device="cuda")
    for seed, ss in ((11, slots), (12, slots), (13, 512)):
        cs = t.rejection_sample_sectioned(count, seed=seed, section_slots=ss, out=out, checksum=True)
        exp = oracle.sample_rejection_sectioned(A, w, W, ss, count, seed)
        np.testing.assert_array_equal(out.cpu().numpy(), exp, err_msg=f"section_slots={ss} seed={seed}")
        assert cs == int(exp.astype(np.int64).sum())


def test_rejection_sectioned_chi_squared(gpu):
    torch = gpu
    n = 20000
    w = atg.generate_weights("sine", n)
    t = atg.Table(n)
    t.set_weights(w)
    t.rejection_build()
    hist = torch.zeros(n, dtype=torch.int32, device="cuda")
    t.rejection_sample_sectioned(20_000_000, seed=6, section_slots=1024, hist=hist)
    h = hist.cpu().numpy().astype(np.float64)
    assert h.sum() == 20_000_000
    assert chi2_pvalue(h, w) > 1e-4
