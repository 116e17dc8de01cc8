"""Batched build (one launch per phase over all tables) against the oracle and
against single-table builds of the same weights: bit-identical rows and W
for every table; batch sampling against the oracle stream."""
import numpy as np
import pytest

import atg

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def make_batch(sizes, dist="exponential", seed=11, dtype=np.float64):
    offsets = np.zeros(len(sizes) + 1, np.int64)
    offsets[1:] = np.cumsum(sizes)
    w = np.concatenate([atg.generate_weights(dist, n, seed=seed + t, param=1.1 if dist == "pareto" else 0.0)
                        for t, n in enumerate(sizes)]).astype(dtype)
    return offsets, w


SIZES = [1, 2, 3, 10, 16, 17, 100, 999, 1000, 4097, 10007, 65536, 100003, 31, 12345, 2, 250000]


@pytest.mark.parametrize("dist,psa,split", [
    ("exponential", True, "bisect"),
    ("pareto", True, "learned"),
    ("unit_uniform", False, "bisect"),
    ("heavy_mix", True, "bisect"),
])
def test_batch_matches_oracle_and_single(gpu, oracle, dist, psa, split):
    offsets, w = make_batch(SIZES, dist)
    b = atg.Batch(offsets, psa_plus=psa, split=split)
    b.set_weights(w)
    b.build()
    b.sync()
    rw, ra, W = b.tables()
    p_auto = 320 * gpu.cuda.get_device_properties(0).multi_processor_count
    for t, n in enumerate(SIZES):
        lo, hi = offsets[t], offsets[t + 1]
        ref = oracle.build(w[lo:hi], psa_plus=psa, num_splits=p_auto)
        assert bits([W[t]])[0] == bits([ref["stats"]["W"]])[0], (t, n)
        bad = np.nonzero((bits(rw[lo:hi]) != bits(ref["row_w"])) | (ra[lo:hi] != ref["row_alias"]))[0]
        assert bad.size == 0, f"table {t} (n={n}): {bad.size} rows differ, first {bad[:5]}"
        if t in (5, 11, 16):
            single = atg.Table(n, psa_plus=psa, split=split)
            single.set_weights(w[lo:hi])
            single.build()
            single.sync()
            sw, sa = single.table()
            assert np.array_equal(bits(sw), bits(rw[lo:hi])) and np.array_equal(sa, ra[lo:hi])


def test_batch_f32_weights(gpu, oracle):
    sizes = [1024, 5000, 70000, 1 << 17]
    offsets, w = make_batch(sizes, "zipf_shuffled", dtype=np.float32)
    b = atg.Batch(offsets, dtype="f32", split="learned")
    b.set_weights(w)
    b.build()
    b.sync()
    rw, ra, W = b.tables()
    p_auto = 320 * gpu.cuda.get_device_properties(0).multi_processor_count
    for t in range(len(sizes)):
        lo, hi = offsets[t], offsets[t + 1]
        ref = oracle.build(w[lo:hi].astype(np.float64), num_splits=p_auto)
        assert bits([W[t]])[0] == bits([ref["stats"]["W"]])[0]
        assert np.array_equal(bits(rw[lo:hi]), bits(ref["row_w"])) and np.array_equal(ra[lo:hi], ref["row_alias"])


def test_batch_sampling_matches_oracle(gpu, oracle):
    torch = gpu
    sizes = [16, 300, 4096, 20000, 77]
    offsets, w = make_batch(sizes, "pareto")
    b = atg.Batch(offsets)
    b.set_weights(w)
    b.build()
    b.sync()
    rw, ra, W = b.tables()
    per = 1000
    out = torch.empty(len(sizes) * per, dtype=torch.int32, device="cuda")
    b.sample(per, seed=9, out=out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for t, n in enumerate(sizes):
        lo, hi = offsets[t], offsets[t + 1]
        exp = oracle.sample_batch_table(rw[lo:hi], ra[lo:hi], W[t] / n, 9, per, first=t * per)
        np.testing.assert_array_equal(got[t * per:(t + 1) * per], exp)


def test_batch_rebuilds_identical(gpu):
    """Repeated builds, with other device work and synchronisation between
    them, rebuild the same tables."""
    torch = gpu
    offsets, w = make_batch([5000, 123, 40000, 64], "exponential")
    b = atg.Batch(offsets)
    b.set_weights(w)
    b.build()
    b.sync()
    ref = b.tables()
    for rep in range(4):
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()
        b.build()
        b.sync()
        got = b.tables()
        assert np.array_equal(bits(got[0]), bits(ref[0])) and np.array_equal(got[1], ref[1]), rep
    lat = b.latencies_ns()
    assert np.all(lat > 0) and np.all(np.isfinite(lat))


def test_batch_device_weights(gpu, oracle):
    torch = gpu
    sizes = [3000, 17, 90000]
    offsets, w = make_batch(sizes, "sine")
    dw = torch.from_numpy(w).cuda()
    b = atg.Batch(offsets)
    b.set_weights(dw)
    b.build()
    b.sync()
    rw, ra, _ = b.tables()
    p_auto = 320 * gpu.cuda.get_device_properties(0).multi_processor_count
    for t in range(len(sizes)):
        lo, hi = offsets[t], offsets[t + 1]
        ref = oracle.build(w[lo:hi], num_splits=p_auto)
        assert np.array_equal(bits(rw[lo:hi]), bits(ref["row_w"])) and np.array_equal(ra[lo:hi], ref["row_alias"])


def test_batch_bad_table_reported(gpu):
    offsets, w = make_batch([100, 200, 300], "exponential")
    w[150] = np.nan
    b = atg.Batch(offsets)
    b.set_weights(w)
    b.build()
    with pytest.raises(atg.AtgError) as e:
        b.sync()
    assert "table 1" in str(e.value)


def test_batch_graph_replays_match_direct_launches(gpu, oracle):
    """The build captured in a hipGraph (atg_batch_use_graph), replayed several
    times with other device work and synchronisation between the replays and
    with new weights, gives the tables of direct launches every time."""
    torch = gpu
    sizes = [5000, 123, 40000, 64, 1, 70000]
    offsets, w = make_batch(sizes, "pareto")
    direct = atg.Batch(offsets)
    direct.set_weights(w)
    direct.build()
    direct.sync()
    ref = direct.tables()
    b = atg.Batch(offsets)
    b.set_weights(w)
    b.use_graph(True)
    for rep in range(4):
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()
        b.build()
        b.sync()
        got = b.tables()
        assert np.array_equal(bits(got[0]), bits(ref[0])) and np.array_equal(got[1], ref[1]), rep
        assert np.array_equal(bits(got[2]), bits(ref[2])), rep
    assert np.all(b.latencies_ns() > 0)
    # New weights: the same graph builds the new tables.
    _, w2 = make_batch(sizes, "exponential", seed=99)
    b.set_weights(w2)
    b.build()
    b.sync()
    rw, ra, W = b.tables()
    p_auto = 320 * torch.cuda.get_device_properties(0).multi_processor_count
    for t in range(len(sizes)):
        lo, hi = offsets[t], offsets[t + 1]
        r = oracle.build(w2[lo:hi], num_splits=p_auto)
        assert np.array_equal(bits(rw[lo:hi]), bits(r["row_w"])) and np.array_equal(ra[lo:hi], r["row_alias"]), t
    # Back to direct launches.
    b.use_graph(False)
    b.set_weights(w)
    b.build()
    b.sync()
    got = b.tables()
    assert np.array_equal(bits(got[0]), bits(ref[0])) and np.array_equal(got[1], ref[1])


def test_batch_groups_identical_tables(gpu, oracle):
    """Tables dealt into launch groups (largest first, each to the lightest
    group) build the same rows as one group, directly and from a replayed
    graph; every group holds tables; a table's latency is its group's."""
    sizes = [int(x) for x in np.random.default_rng(3).integers(16, 70000, size=200)]
    offsets, w = make_batch(sizes, "exponential", seed=21)
    kw = dict(section_size=2048, chains=32, num_splits=-38)
    one = atg.Batch(offsets, **kw)
    one.set_weights(w)
    one.build()
    one.sync()
    rw1, ra1, W1 = one.tables()
    for group_items, graph in ((offsets[-1] // 5, False), (offsets[-1] // 3, True)):
        b = atg.Batch(offsets, group_items=group_items, **kw)
        g, ng = b.groups()
        assert ng == -(-int(offsets[-1]) // group_items) and set(g.tolist()) == set(range(ng))
        b.set_weights(w)
        if graph:
            b.use_graph(True)
        for _ in range(2):
            b.build()
            b.sync()
            rw, ra, W = b.tables()
            assert np.array_equal(bits(rw), bits(rw1)) and np.array_equal(ra, ra1) and np.array_equal(bits(W), bits(W1))
        for t in (0, 57, 199):
            tw, ta, tW = b.table(t)
            lo, hi = offsets[t], offsets[t + 1]
            assert np.array_equal(bits(tw), bits(rw1[lo:hi])) and np.array_equal(ta, ra1[lo:hi]) and tW == W1[t]
        assert np.all(b.latencies_ns() > 0)
    ref_rw, ref_ra, ref_W = oracle.build_batch(w, offsets, **kw)
    assert np.array_equal(bits(rw1), bits(ref_rw)) and np.array_equal(ra1, ref_ra) and np.array_equal(bits(W1), bits(ref_W))


def test_batch_autotune_keeps_tables(gpu, oracle):
    """The batch tuner varies the split method and the launch groups only:
    the tables after tuning equal the oracle's."""
    sizes = [int(x) for x in np.random.default_rng(4).integers(1000, 300000, size=300)]
    offsets, w = make_batch(sizes, "pareto", seed=31)
    kw = dict(section_size=2048, chains=32, num_splits=-38)
    b = atg.Batch(offsets, **kw)
    b.set_weights(w)
    res = b.autotune(reps=1, makespan_slack=0.1)
    assert res["split"] in ("bisect", "learned") and res["makespan_ms"] > 0 and res["p99_latency_ms"] > 0
    b.build()
    b.sync()
    rw, ra, W = b.tables()
    ref_rw, ref_ra, ref_W = oracle.build_batch(w, offsets, **kw)
    assert np.array_equal(bits(rw), bits(ref_rw)) and np.array_equal(ra, ref_ra) and np.array_equal(bits(W), bits(ref_W))
