"""GPU build parity: every stage of the gfx950 build against the CPU oracle,
bit for bit (W, leftover lists, prefix sums, splits, table rows)."""
import numpy as np
import pytest

import atg

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def gpu_build(w, dtype="f64", **kw):
    t = atg.Table(len(w), dtype=dtype, **kw)
    t.set_weights(w)
    t.build()
    t.sync()
    return t


def assert_same_build(t, ref, n, check_lists=True):
    st = t.stats()
    rs = ref["stats"]
    assert bits([st["W"]])[0] == bits([rs["W"]])[0], (st["W"], rs["W"])
    assert bits([st["W_N"]])[0] == bits([rs["W_N"]])[0]
    assert (st["numL"], st["numH"], st["p"], st["degenerate"]) == (rs["numL"], rs["numH"], rs["p"], rs["degenerate"])
    if st["monotone"] >= 0:  # computed for the learned split
        assert st["monotone"] == rs["monotone"]
    if check_lists:
        lh = t.lh()
        for k in ("L_item", "H_item"):
            np.testing.assert_array_equal(lh[k], ref[k], err_msg=k)
        for k in ("L_w", "L_pref", "H_w", "H_pref"):
            mism = np.nonzero(bits(lh[k]) != bits(ref[k]))[0]
            assert mism.size == 0, f"{k}: {mism.size} mismatches, first {mism[:5]} gpu={lh[k][mism[:3]]} ref={ref[k][mism[:3]]}"
        if rs["p"] > 0:
            i, j, s = t.splits()
            np.testing.assert_array_equal(i, ref["split_i"])
            np.testing.assert_array_equal(j, ref["split_j"])
            assert np.array_equal(bits(s), bits(ref["split_spill"]))
    tw, ta = t.table()
    bad = np.nonzero((bits(tw) != bits(ref["row_w"])) | (ta != ref["row_alias"]))[0]
    assert bad.size == 0, f"{bad.size} table rows differ, first {bad[:5]}: gpu {tw[bad[:3]]} {ta[bad[:3]]} ref {ref['row_w'][bad[:3]]} {ref['row_alias'][bad[:3]]}"


CASES = [
    # n, dist, psa_plus, chains, section, p
    (1, "unit_uniform", True, 16, 0, 0),
    (2, "unit_uniform", True, 16, 0, 0),
    (3, "pareto", False, 16, 0, 0),
    (10, "unit_uniform", True, 16, 0, 0),
    (11, "pareto", True, 16, 0, 0),
    (100, "pareto", True, 16, 0, 0),
    (1000, "unit_uniform", False, 16, 0, 0),
    (1000, "pareto", True, 16, 0, 0),
    (10007, "pareto", True, 16, 0, 0),
    (10007, "exponential", False, 16, 0, 1400),
    (65536, "pareto", True, 64, 2048, 0),
    (100003, "sine", True, 16, 0, 1400),
    (100003, "powerlaw1_shuffled", True, 8, 2000, 0),
    (505711, "one_light", True, 16, 0, 16576),
    (1 << 20, "unit_uniform", True, 16, 0, 0),
    (1 << 20, "pareto", True, 16, 0, 0),
    (1 << 20, "pareto", False, 16, 0, 1400),
    (1 << 20, "heavy_mix", True, 32, 2048, 0),
    # chain counts that are not powers of two (the leftover rank search
    # starts at the largest power of two below the count)
    (100003, "pareto", True, 1, 1024, 0),
    (200003, "exponential", True, 3, 1500, 0),
    (300007, "pareto", True, 17, 2048, 0),
]


@pytest.mark.parametrize("n,dist,psa,chains,section,p", CASES)
def test_build_matches_oracle(gpu, oracle, n, dist, psa, chains, section, p):
    w = atg.generate_weights(dist, n, seed=42, param=1.1 if dist == "pareto" else 0.0)
    # p = 0 asks for the reference's AUTO value: 320 splits per compute unit.
    p_req = p if p > 0 else 320 * gpu.cuda.get_device_properties(0).multi_processor_count
    t = gpu_build(w, psa_plus=psa, chains=chains, section_size=section, num_splits=p_req)
    ref = oracle.build(w, psa_plus=psa, chains=chains, section_size=section, num_splits=p_req)
    assert_same_build(t, ref, n)


@pytest.mark.parametrize("split", ["learned", "pary"])
@pytest.mark.parametrize("n,dist,p,segments", [
    (1 << 20, "pareto", 65536, 4096), (1 << 20, "heavy_mix", 100000, 4096), (3_000_000, "unit_uniform", 300000, 1024),
    (100003, "sine", 10000, 4096), (100003, "pareto", 7, 4096), (2_000_003, "zipf_shuffled", 200000, 4096)])
def test_learned_split_bit_identical(gpu, oracle, n, dist, p, segments, split):
    """The learned index and the P-ary search return the bisection's splits:
    identical tables."""
    w = atg.generate_weights(dist, n, seed=5, param=1.1 if dist == "pareto" else 0.0)
    plain = gpu_build(w, num_splits=p)
    learned = gpu_build(w, num_splits=p, split=split, learned_segments=segments)
    ref = oracle.build(w, num_splits=p)
    assert_same_build(learned, ref, n)
    for a, b in zip(plain.splits(), learned.splits()):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    st = learned.stats()
    assert st["monotone"] == 1
    assert st["learned_fallbacks"] == 0, st


@pytest.mark.parametrize("layout", ["aos", "soa"])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_layout_and_dtype(gpu, oracle, layout, dtype):
    n = 200003
    w = atg.generate_weights("pareto", n, seed=7, param=1.1)
    if dtype == "f32":
        w = w.astype(np.float32)
    t = gpu_build(w, dtype=dtype, layout=layout)
    ref = oracle.build(w.astype(np.float64), num_splits=81920)
    assert_same_build(t, ref, n)


def test_device_weights_and_rebuild(gpu, oracle):
    torch = gpu
    n = 300000
    w = atg.generate_weights("unit_uniform", n, seed=3)
    dw = torch.from_numpy(w).cuda()
    t = atg.Table(n)
    t.set_weights(dw)
    for _ in range(3):  # repeated builds on the same handle are identical
        t.build()
    t.sync()
    assert_same_build(t, oracle.build(w, num_splits=81920), n)


@pytest.mark.parametrize("case", ["all_equal", "zeros_and_ones", "one_heavy", "best_worst", "integers"])
def test_edge_inputs(gpu, oracle, case):
    n = 50000
    if case == "all_equal":
        w = np.full(n, 3.25)
    elif case == "zeros_and_ones":
        w = np.zeros(n)
        w[::3] = 1.0
    elif case == "one_heavy":
        w = atg.generate_weights("one_heavy", n)
    elif case == "best_worst":
        w = atg.generate_weights("best_worst", n)
    else:
        w = (np.arange(n) % 16).astype(np.float64)
    t = gpu_build(w)
    ref = oracle.build(w, num_splits=81920)
    assert_same_build(t, ref, n)
    v = oracle.verify(w, ref["stats"]["W"], ref["stats"]["W_N"], ref["row_w"], ref["row_alias"])
    assert v["reference_ok"]


@pytest.mark.parametrize("bad", [np.nan, np.inf, -1.0])
def test_invalid_weights_rejected(gpu, bad):
    w = np.ones(1000)
    w[17] = bad
    t = atg.Table(1000)
    t.set_weights(w)
    t.build()
    with pytest.raises(atg.AtgError) as e:
        t.sync()
    assert e.value.status == 1


def test_all_zero_weights_rejected(gpu):
    t = atg.Table(100)
    t.set_weights(np.zeros(100))
    t.build()
    with pytest.raises(atg.AtgError):
        t.sync()


def test_autotune_shape_returns_valid_shape(gpu, oracle, tmp_path, monkeypatch):
    """The shape search picks a shape from its grid; that shape builds the
    oracle's table; a second call with the cache file returns the cached
    choice."""
    monkeypatch.setenv("ATG_TUNE_CACHE", str(tmp_path / "tune.txt"))
    n = 400_000
    w = atg.generate_weights("pareto", n, seed=3, param=1.1)
    shape, ms = atg.autotune_shape(w, reps=1)
    assert ms > 0
    assert shape["section_size"] in (1024, 1536, 2048) and shape["chains"] in (12, 16, 24, 32)
    assert 2 <= shape["num_splits"] <= n // 10
    t = gpu_build(w, chains=shape["chains"], section_size=shape["section_size"], num_splits=shape["num_splits"])
    ref = oracle.build(w, chains=shape["chains"], section_size=shape["section_size"], num_splits=shape["num_splits"])
    assert_same_build(t, ref, n)
    shape2, ms2 = atg.autotune_shape(w, reps=1)
    assert ms2 < 0 and shape2 == shape


@pytest.mark.parametrize("dist", ["pareto", "heavy_mix"])
def test_autotune_leaves_table_unchanged(gpu, oracle, monkeypatch, dist):
    """The run-time tuner varies only knobs that keep the rows: two uncached
    calls at one explicit shape build byte-identical tables, equal to the
    oracle's table for that shape, whatever knobs win."""
    monkeypatch.delenv("ATG_TUNE_CACHE", raising=False)
    n = 600_000
    w = atg.generate_weights(dist, n, seed=8, param=1.1 if dist == "pareto" else 0.9)
    shape = dict(chains=24, section_size=2048, num_splits=n // 64)
    ref = oracle.build(w, want_lists=False, **shape)
    for rep in range(2):
        knobs, ms = atg.autotune(w, reps=1, **shape)
        assert ms > 0 and knobs["split"] in ("bisect", "learned", "pary")
        t = gpu_build(w, split=knobs["split"], learned_segments=knobs["learned_segments"],
                      interleave=knobs["interleave"], **shape)
        tw, ta = t.table()
        assert np.array_equal(bits(tw), bits(ref["row_w"])) and np.array_equal(ta, ref["row_alias"]), (rep, knobs)


def test_split_ratio_mode(gpu, oracle):
    """num_splits < 0 asks for one split per -num_splits items."""
    n = 300_007
    w = atg.generate_weights("exponential", n, seed=4)
    t = gpu_build(w, num_splits=-38)
    ref = oracle.build(w, num_splits=n // 38)
    assert_same_build(t, ref, n)


@pytest.mark.parametrize("n,dist,split,layout,p", [
    (1 << 20, "pareto", "bisect", "aos", 1 << 14),
    (1 << 20, "heavy_mix", "learned", "aos", 1 << 14),
    (1 << 20, "pareto", "pary", "aos", 1 << 14),
    (300_000, "exponential", "bisect", "soa", 0),
    (5000, "pareto", "bisect", "aos", 0),  # too few splits to chunk: sequential
])
def test_interleaved_split_pack_identical(gpu, oracle, n, dist, split, layout, p):
    """Split and pack interleaved on two streams (the reference's
    performSplitAndPackInterleaved) build the oracle's table, rebuild after
    other device work, and match the sequential build."""
    torch = gpu
    w = atg.generate_weights(dist, n, seed=13, param=1.1 if dist == "pareto" else 0.9)
    kw = dict(chains=24, section_size=2048, num_splits=p, split=split, layout=layout)
    ref = oracle.build(w, chains=24, section_size=2048, num_splits=p or 320 * torch.cuda.get_device_properties(0).multi_processor_count,
                       want_lists=False)
    t = atg.Table(n, interleave=True, **kw)
    t.set_weights(w)
    for rep in range(3):
        x = torch.randn(1 << 18, device="cuda")
        (x * 3).sum().item()
        t.build()
        t.sync()
        tw, ta = t.table()
        assert np.array_equal(bits(tw), bits(ref["row_w"])) and np.array_equal(ta, ref["row_alias"]), rep
    assert t.stats()["p"] == ref["stats"]["p"]


@pytest.mark.parametrize("corrupt", ["row_weight", "alias"])
def test_post_build_rejects_corrupted_table(gpu, oracle, corrupt):
    """atg_post_build (postBuild + verifyTableCorrectness, AliasTable.cu:24-65)
    passes the built table and rejects the same rows with one row weight or
    one alias changed, loaded with atg_set_table."""
    import test_oracle as pins
    n = 20000
    w = atg.generate_weights("sine", n)
    t = gpu_build(w, num_splits=1000)
    rel, ok = t.post_build()
    assert ok and rel < 1e-9
    ref = oracle.build(w, num_splits=1000, want_lists=False)
    fn = pins.corrupt_row_weight if corrupt == "row_weight" else pins.corrupt_alias
    rw, ra = fn(ref)
    t.set_table(rw, ra, t.stats()["W"])
    rel, ok = t.post_build()
    assert not ok and rel > 1e-3


@pytest.mark.parametrize("pack", ["basic", "no_weights"])
@pytest.mark.parametrize("n,dist,psa,p,split,layout,dtype", [
    (100003, "pareto", False, 1400, "bisect", "aos", "f64"),
    (100003, "sine", False, 0, "learned", "soa", "f64"),
    (1 << 20, "heavy_mix", False, 16384, "bisect", "aos", "f32"),
    (505711, "one_light", False, 16576, "bisect", "aos", "f64"),
    (65536, "pareto", True, 0, "bisect", "aos", "f64"),
    (11, "unit_uniform", False, 0, "bisect", "soa", "f64"),
])
def test_pack_methods_identical(gpu, oracle, pack, n, dist, psa, p, split, layout, dtype):
    """The pack variants of Variant::pack (PackMethodBasic, PackMethodNoWeights)
    write the oracle's rows, as the default LDS-staged pack does; the
    NO_WEIGHTS pack refuses PSA+ (whose heavy residues are not input
    weights)."""
    w = atg.generate_weights(dist, n, seed=17, param=1.1 if dist == "pareto" else 0.9)
    if dtype == "f32":
        w = w.astype(np.float32)
    if pack == "no_weights" and psa:
        with pytest.raises(atg.AtgError) as e:
            atg.Table(n, psa_plus=True, pack=pack)
        assert e.value.status == 1
        return
    kw = dict(psa_plus=psa, num_splits=p, split=split, layout=layout)
    t = gpu_build(w, dtype=dtype, pack=pack, **kw)
    ref = oracle.build(w.astype(np.float64), psa_plus=psa,
                       num_splits=p or 320 * gpu.cuda.get_device_properties(0).multi_processor_count,
                       want_lists=False)
    assert_same_build(t, ref, n, check_lists=False)
    t.build()  # a rebuild on the same handle
    t.sync()
    assert_same_build(t, ref, n, check_lists=False)
    t.close()


@pytest.mark.parametrize("n,dist,psa,p,layout,dtype", [
    (1 << 20, "pareto", True, 2000, "aos", "f64"),
    (1 << 20, "pareto", False, 1000, "soa", "f64"),
    (1_000_003, "heavy_mix", True, 300, "aos", "f32"),
    (300_007, "unit_uniform", True, 100, "soa", "f64"),
    (200_003, "sine", False, 2, "aos", "f64"),
])
def test_long_chain_pack_matches_oracle(gpu, oracle, n, dist, psa, p, layout, dtype):
    """Chains of more than 256 items (n / p > 256) run the windowed pack: each
    lane steps through windows of its lists staged in LDS.  Rows, lists and
    splits equal the oracle's, on a fresh build and on a rebuild."""
    w = atg.generate_weights(dist, n, seed=29, param=1.1 if dist == "pareto" else 0.9)
    if dtype == "f32":
        w = w.astype(np.float32)
    t = gpu_build(w, dtype=dtype, psa_plus=psa, num_splits=p, layout=layout)
    ref = oracle.build(w.astype(np.float64), psa_plus=psa, num_splits=p)
    assert_same_build(t, ref, n)
    t.build()
    t.sync()
    assert_same_build(t, ref, n)
    t.close()
