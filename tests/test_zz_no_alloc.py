"""atg_build_noalloc (a build with every scratch buffer in a file-scope device
array, for callers that cannot allocate) against atg_build of the same
weights and options through the C ABI: identical rows and W, on calls in a
row with the second input smaller than the first, AoS and SoA, FP64 and
FP32, bisection and the learned split; shapes beyond the scratch array and
unsupported options are rejected."""
import numpy as np
import pytest

import atg

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def aos_split(rows, n):
    r = rows.cpu().numpy().view(np.uint8).reshape(n, 16)
    return r[:, :8].copy().view(np.float64).ravel(), r[:, 8:12].copy().view(np.int32).ravel()


def check_against_build(torch, n, dist, layout, dtype, split, seed, **shape):
    w = atg.generate_weights(dist, n, seed=seed, param=1.1 if dist == "pareto" else 0.9)
    if dtype == "f32":
        w = w.astype(np.float32)
    dw = torch.from_numpy(w).cuda()
    t = atg.Table(n, dtype=dtype, layout=layout, split=split, **shape)
    t.set_weights(dw)
    t.build()
    t.sync()
    ref_w, ref_a = t.table()
    W_ref = t.stats()["W"]
    t.close()
    W = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    err = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    if layout == "aos":
        rows = torch.full((2 * n,), float("nan"), dtype=torch.float64, device="cuda")
        atg.build_noalloc(dw, rows, W_out=W, error_out=err, dtype=dtype, layout=layout, split=split, **shape)
        torch.cuda.synchronize()
        got_w, got_a = aos_split(rows, n)
    else:
        rows = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        al = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        atg.build_noalloc(dw, rows, al, W_out=W, error_out=err, dtype=dtype, layout=layout, split=split, **shape)
        torch.cuda.synchronize()
        got_w, got_a = rows.cpu().numpy(), al.cpu().numpy()
    assert int(err.item()) == 0
    assert bits([W.item()])[0] == bits([W_ref])[0]
    bad = np.nonzero((bits(got_w) != bits(ref_w)) | (got_a != ref_a))[0]
    assert bad.size == 0, f"n={n}: {bad.size} rows differ from atg_build, first {bad[:5]}"


@pytest.mark.parametrize("layout,dtype,split", [("aos", "f64", "bisect"), ("soa", "f32", "learned"),
                                                ("aos", "f64", "learned"), ("aos", "f64", "pary")])
def test_noalloc_equals_build_two_calls(gpu, layout, dtype, split):
    torch = gpu
    # Two calls in a row: the second with a smaller input, so stale scratch
    # contents of the first call are present when the second runs.
    check_against_build(torch, 1 << 20, "pareto", layout, dtype, split, 3, chains=24, section_size=2048,
                        num_splits=(1 << 20) // 64)
    check_against_build(torch, 100_003, "heavy_mix", layout, dtype, split, 4)


def test_noalloc_small_tables(gpu):
    torch = gpu
    for n in (1, 2, 11, 1000):
        check_against_build(torch, n, "exponential", "aos", "f64", "bisect", n)


def test_noalloc_rejects_beyond_scratch_and_unsupported(gpu):
    torch = gpu
    n = 4 << 20
    dw = torch.ones(n, dtype=torch.float64, device="cuda")
    rows = torch.empty(2 * n, dtype=torch.float64, device="cuda")
    with pytest.raises(atg.AtgError) as e:
        atg.build_noalloc(dw, rows)
    assert e.value.status == 1 and "scratch" in str(e.value)
    small = torch.ones(1000, dtype=torch.float64, device="cuda")
    o = atg.default_opts()
    o.interleave = 1
    import ctypes as C
    st = atg.lib().atg_build_noalloc(small.data_ptr(), 0, 1000, 0, C.byref(o), rows.data_ptr(), None, None, None,
                                     None)
    assert st == 5  # ATG_ENOTSUP


@pytest.mark.parametrize("pack,psa", [("basic", True), ("no_weights", False)])
def test_noalloc_pack_methods(gpu, pack, psa):
    torch = gpu
    check_against_build(torch, 100_003, "pareto", "aos", "f64", "bisect", 5, pack=pack, psa_plus=psa)
    check_against_build(torch, 5000, "sine", "soa", "f32", "learned", 6, pack=pack, psa_plus=psa)


def test_noalloc_long_chains(gpu):
    """Chains of more than 256 items (the windowed pack, static LDS only):
    two calls in a row, the second smaller."""
    torch = gpu
    check_against_build(torch, 1 << 20, "pareto", "aos", "f64", "bisect", 8, num_splits=1000)
    check_against_build(torch, 300_007, "unit_uniform", "soa", "f32", "learned", 9, num_splits=100)
