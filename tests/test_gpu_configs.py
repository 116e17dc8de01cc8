"""The benchmark workloads at their full sizes, against the oracle.

* 65,536 tables of log-uniform sizes in [16, 2^16] with Exp(1) FP64 weights
  (5.1e8 items), at the benchmark's shape, in launch groups and from a
  replayed graph: every table's W and rows equal the oracle's.
* 16,384 tables of log-uniform sizes in [2^10, 2^20] with FP32 rank-Zipf(1.2)
  weights (2.5e9 items, so table offsets pass 2^31), learned split: every
  table's W equals the canonical sum and passes the reference's table check;
  the rows of every table that starts beyond item 2^31 and of the 64 largest
  tables equal the oracle's.
* N = 1e8 Pareto(1.1): 1e9 sectioned samples, compared section by section
  with the oracle on a few hundred sections (the first and the last among
  them); the rejection array A bit for bit and 1e7 rejection samples; 1e7
  basic samples.
* N = 1e8 with 1% of the items holding 90% of the mass: the rejection array
  through the learned inverse-CDF fill (4096 segments), the p-ary and the
  bisection fills, each equal to the oracle's; the model equal to the oracle's.
"""
import numpy as np
import pytest

import atg

pytestmark = pytest.mark.gpu

BATCH_SHAPE = dict(section_size=2048, chains=32, num_splits=-38)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def log(*a):
    print("[configs]", *a, flush=True)


def test_batch_65536_tables_exp_f64(gpu, oracle):
    offsets, w = atg.generate_batch("exponential", 65536, 16, 1 << 16, seed=42)
    total = int(offsets[-1])
    log(f"65,536 tables, {total} items")
    b = atg.Batch(offsets, group_items=total // 4, **BATCH_SHAPE)
    b.set_weights(w)
    b.use_graph(True)
    for _ in range(2):
        b.build()
    b.sync()
    rw, ra, W = b.tables()
    b.close()
    ref_rw, ref_ra, ref_W = oracle.build_batch(w, offsets, **BATCH_SHAPE)
    bad_W = np.nonzero(bits(W) != bits(ref_W))[0]
    assert bad_W.size == 0, f"{bad_W.size} tables with another W, first {bad_W[:5]}"
    bad = np.nonzero((bits(rw) != bits(ref_rw)) | (ra != ref_ra))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"
    # 64 samples per table against the oracle's stream for a few hundred tables
    torch = gpu
    b = atg.Batch(offsets, **BATCH_SHAPE)
    b.set_weights(w)
    b.build()
    out = torch.empty(65536 * 64, dtype=torch.int32, device="cuda")
    b.sample(64, seed=7, out=out)
    got = out.cpu().numpy()
    for t in list(range(0, 65536, 257)) + [65535]:
        lo, hi = offsets[t], offsets[t + 1]
        exp = oracle.sample_batch_table(ref_rw[lo:hi], ref_ra[lo:hi], ref_W[t] / (hi - lo), 7, 64, t * 64)
        assert np.array_equal(got[t * 64:(t + 1) * 64], exp), t


def test_batch_16384_tables_zipf_f32_learned(gpu, oracle):
    offsets, w = atg.generate_batch("zipf_shuffled", 16384, 1 << 10, 1 << 20, seed=42, param=1.2, dtype="f32")
    total = int(offsets[-1])
    assert total > 2**31
    log(f"16,384 FP32 tables, {total} items")
    b = atg.Batch(offsets, dtype="f32", split="learned", **BATCH_SHAPE)
    b.set_weights(w)
    b.build()
    b.sync()
    rw, ra, W = b.tables()
    b.close()
    log("copied out")
    Wc, ok, mabs = oracle.check_batch(w, offsets, rw, ra, W, **BATCH_SHAPE)
    bad_W = np.nonzero(bits(W) != bits(Wc))[0]
    assert bad_W.size == 0, f"{bad_W.size} tables with another W, first {bad_W[:5]}"
    assert ok.all(), f"{np.count_nonzero(~ok)} tables fail the table check, first {np.nonzero(~ok)[0][:5]}"
    sizes = np.diff(offsets)
    beyond = np.nonzero(offsets[:-1] >= 2**31)[0]
    largest = np.argsort(-sizes, kind="stable")[:64]
    which = np.union1d(beyond, largest)
    assert beyond.size > 1000
    log(f"oracle on {which.size} tables")
    ref_rw, ref_ra, ref_W = oracle.build_batch(w, offsets, which=which, **BATCH_SHAPE)
    for t in which.tolist():
        lo, hi = offsets[t], offsets[t + 1]
        assert bits([W[t]])[0] == bits([ref_W[t]])[0], t
        assert np.array_equal(bits(rw[lo:hi]), bits(ref_rw[lo:hi])) and np.array_equal(ra[lo:hi], ref_ra[lo:hi]), t


N = 100_000_000
SHAPE = dict(chains=24, section_size=2048, num_splits=1_562_500)


@pytest.fixture(scope="module")
def headline(gpu, oracle):
    w = atg.generate_weights("pareto", N, seed=42, param=1.1)
    t = atg.Table(N, **SHAPE)
    t.set_weights(w)
    t.build()
    t.sync()
    ref = oracle.build(w, want_lists=False, threads=16, **SHAPE)
    tw, ta = t.table()
    assert np.array_equal(bits(tw), bits(ref["row_w"])) and np.array_equal(ta, ref["row_alias"])
    del tw, ta
    return w, t, ref


def test_sectioned_1e9_samples_headline(gpu, oracle, headline):
    torch = gpu
    w, t, ref = headline
    S, rows, seed = 1_000_000_000, 4096, 100
    out = torch.empty(S, dtype=torch.int32, device="cuda")
    t.sample(S, seed=seed, kind="sectioned", section_rows=rows, out=out)
    torch.cuda.synchronize()
    nsec = (N + rows - 1) // rows
    sel = np.unique(np.concatenate([[0, 1, nsec - 2, nsec - 1], np.random.default_rng(1).integers(0, nsec, 300)]))
    exp, first, cnt = oracle.sample_sectioned_subset(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], rows, S,
                                                     seed, sel)
    assert sel[-1] == nsec - 1 and int(first[-1] + cnt[-1]) == S  # the last section ends the run
    pos = 0
    for s, f, c in zip(sel.tolist(), first.tolist(), cnt.tolist()):
        got = out[f:f + c].cpu().numpy()
        assert np.array_equal(got, exp[pos:pos + c]), f"section {s}"
        pos += c
    log(f"{sel.size} sections, {pos} samples compared")


def test_rejection_and_basic_headline(gpu, oracle, headline):
    torch = gpu
    w, t, ref = headline
    t.rejection_build()
    A = t.rejection_array()
    W = ref["stats"]["W"]
    assert t.rejection_weight_sum() == W and t.stats()["W"] == W
    _, A_ref = oracle.rejection_build(w, W)
    assert A.size == A_ref.size and np.array_equal(A, A_ref)
    count = 10_000_000
    out = torch.empty(count, dtype=torch.int32, device="cuda")
    t.rejection_sample(count, seed=3, out=out)
    exp, _ = oracle.sample_rejection(A_ref, w, W, 3, count)
    np.testing.assert_array_equal(out.cpu().numpy(), exp)
    t.sample(count, seed=4, kind="basic", out=out)
    exp = oracle.sample_basic(ref["row_w"], ref["row_alias"], ref["stats"]["W_N"], 4, count)
    np.testing.assert_array_equal(out.cpu().numpy(), exp)
    # 1e9 samples by sections of A, a few hundred whole sections (the first
    # and the last among them) against the oracle
    del out
    S, slots = 1_000_000_000, 4096
    out = torch.empty(S, dtype=torch.int32, device="cuda")
    t.rejection_sample_sectioned(S, seed=21, section_slots=slots, out=out)
    torch.cuda.synchronize()
    nsec = (A_ref.size + slots - 1) // slots
    sel = np.unique(np.concatenate([[0, nsec - 1], np.random.default_rng(2).integers(0, nsec, 300)]))
    exp, first, cnt = oracle.sample_rejection_sectioned(A_ref, w, W, slots, S, 21, sections=sel)
    assert int(first[-1] + cnt[-1]) == S
    for s, f, c in zip(sel.tolist(), first.tolist(), cnt.tolist()):
        assert np.array_equal(out[f:f + c].cpu().numpy(), exp[f:f + c]), f"section {s}"
    log(f"sectioned rejection: {sel.size} sections, {int(cnt.sum())} samples compared")


def test_learned_inverse_cdf_fill_heavy_mix_1e8(gpu, oracle):
    """N = 1e8 with 1% of the items holding 90% of the mass, the rejection
    array filled through the 4096-segment inverse-CDF model fitted on the
    build's own inclusive counts, through the p-ary and the plain bisection
    fills: A (items and acceptance bits) equals the oracle's for each, the
    model equals the oracle's brute-force fit, and no prediction falls back."""
    n, K = N, 4096
    w = atg.generate_weights("heavy_mix", n, seed=42)
    t = atg.Table(n)
    t.set_weights(w)
    ref = None
    for fill in ("learned", "pary", "bisect"):
        t.rejection_build(fill=fill, knots=K)
        W = t.rejection_weight_sum()
        if ref is None:
            incl, A_ref = oracle.rejection_build(w, W)
            x = (float(n) * w) / W
            first = np.concatenate([[0], incl[:-1]]).astype(np.int64)
            acc_ref = np.ones(A_ref.size)
            acc_ref[first] = x - np.floor(x)
            del x, first
            ref = W
            q, e = t.rejection_model(K)
            q_ref, e_ref = oracle.rejection_model(incl, K)
            assert np.array_equal(q, q_ref) and np.array_equal(e, e_ref)
            assert t.rejection_fill_fallbacks() == 0
            log(f"|A| = {A_ref.size}, model bounds: median {int(np.median(e))}, max {int(e.max())}")
        assert W == ref
        A = t.rejection_array()
        assert A.size == A_ref.size and np.array_equal(A, A_ref), fill
        del A
        assert np.array_equal(bits(t.rejection_accept()), bits(acc_ref)), fill
        log(f"{fill} fill: A equal to the oracle's")
    t.close()
