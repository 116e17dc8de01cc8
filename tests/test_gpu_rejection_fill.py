"""The fills of the rejection array (RejectionSampling::fillA): the direct
run-writing fill, the reference's bisection (RejectionSamplingBinarySearch.cu:
12-33), the wave-wide p-ary narrowing (RejectionSamplingPArySearch.cu:12-123)
and the learned inverse-CDF index.  Every fill must write the oracle's A
(items and acceptance bits); the learned fill's model must equal the oracle's
brute-force fit, and a model with wrong bounds may change the fallback count
but never A."""
import numpy as np
import pytest

import atg

pytestmark = pytest.mark.gpu

FILLS = ["direct", "bisect", "pary", "learned"]


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def outlier_sine(n, weight0):
    # RejectionSampling::benchmarkBuild's weights (RejectionSampling.cu:58-60).
    w = atg.generate_weights("sine", n)
    w[0] = weight0
    return w


def expected_accept(w, W, incl):
    """1.0 everywhere but at the first slot of each run, frac(N w / W) there."""
    n = len(w)
    acc = np.ones(int(incl[-1]))
    x = (float(n) * w.astype(np.float64)) / W
    first = np.concatenate([[0], incl[:-1]]).astype(np.int64)
    acc[first] = x - np.floor(x)
    return acc


CASES = {
    "pareto": lambda: atg.generate_weights("pareto", 100003, seed=11, param=1.1),
    "outlier": lambda: outlier_sine(20000, 2e7),
    "uniform": lambda: np.full(4096, 0.25),
    "heavy_mix": lambda: atg.generate_weights("heavy_mix", 300001, seed=4),
    "one": lambda: np.array([2.5]),
    "two": lambda: np.array([1.0, 3.0]),
    "one_heavy_f32": lambda: atg.generate_weights("one_heavy", 70001, seed=13).astype(np.float32),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("fill", FILLS)
def test_fill_matches_oracle(gpu, oracle, case, fill):
    w = CASES[case]()
    n = len(w)
    t = atg.Table(n, dtype="f32" if w.dtype == np.float32 else "f64")
    t.set_weights(w)
    knots = 64 if n > 1000 else 8
    t.rejection_build(fill=fill, knots=knots)
    W = t.rejection_weight_sum()
    w64 = w.astype(np.float64)
    incl, A = oracle.rejection_build(w64, W)
    assert t.rejection_size() == len(A)
    np.testing.assert_array_equal(t.rejection_array(), A)
    assert np.array_equal(bits(t.rejection_accept()), bits(expected_accept(w64, W, incl)))
    if fill == "learned":
        q, e = t.rejection_model(knots)
        q_ref, e_ref = oracle.rejection_model(incl, knots)
        np.testing.assert_array_equal(q, q_ref)
        np.testing.assert_array_equal(e, e_ref)
        assert t.rejection_fill_fallbacks() == 0
    # fillA alone again: the same A (the direct fill's task counter restarts)
    t.rejection_fill()
    np.testing.assert_array_equal(t.rejection_array(), A)
    t.close()


@pytest.mark.parametrize("knots", [1, 4096, 1 << 16])
def test_learned_fill_knot_counts(gpu, oracle, knots):
    w = atg.generate_weights("pareto", 250007, seed=3, param=1.1)
    t = atg.Table(len(w))
    t.set_weights(w)
    t.rejection_build(fill="learned", knots=knots)
    W = t.rejection_weight_sum()
    incl, A = oracle.rejection_build(w, W)
    np.testing.assert_array_equal(t.rejection_array(), A)
    q, e = t.rejection_model(knots)
    q_ref, e_ref = oracle.rejection_model(incl, knots)
    np.testing.assert_array_equal(q, q_ref)
    np.testing.assert_array_equal(e, e_ref)
    assert t.rejection_fill_fallbacks() == 0


def test_learned_fill_wrong_model_falls_back(gpu, oracle):
    n, K = 100003, 64
    w = atg.generate_weights("pareto", n, seed=7, param=1.1)
    t = atg.Table(n)
    t.set_weights(w)
    t.rejection_build(fill="learned", knots=K)
    A = t.rejection_array()
    acc = bits(t.rejection_accept())
    q, e = t.rejection_model(K)
    assert t.rejection_fill_fallbacks() == 0
    # Bounds of zero: the windows shrink to the prediction; most waves miss.
    t.rejection_set_model(q, np.zeros(K, np.uint32))
    t.rejection_fill()
    f1 = t.rejection_fill_fallbacks()
    assert f1 > 0
    np.testing.assert_array_equal(t.rejection_array(), A)
    assert np.array_equal(bits(t.rejection_accept()), acc)
    # Knots shifted away from the items of their slots, bounds kept.
    t.rejection_set_model(np.minimum(q + 5000, n - 1), e)
    t.rejection_fill()
    f2 = t.rejection_fill_fallbacks()
    assert f2 > f1
    np.testing.assert_array_equal(t.rejection_array(), A)
    # The fitted model again: no new fallbacks.
    t.rejection_set_model(q, e)
    t.rejection_fill()
    assert t.rejection_fill_fallbacks() == f2
    np.testing.assert_array_equal(t.rejection_array(), A)
    # A model the fill cannot use safely is refused.
    bad = q.copy()
    bad[3], bad[4] = bad[4] + 1, bad[3]
    with pytest.raises(atg.AtgError):
        t.rejection_set_model(bad, e)
    with pytest.raises(atg.AtgError):
        t.rejection_set_model(np.append(q[:-1], n), e)
    with pytest.raises(atg.AtgError):
        t.rejection_set_model(q[:-1], e[:-1])
    t.close()


def test_fill_arguments_rejected(gpu):
    t = atg.Table(1000)
    t.set_weights(np.ones(1000))
    with pytest.raises(atg.AtgError):
        t.rejection_build(fill="sweep")
    with pytest.raises(atg.AtgError) as ex:
        t.rejection_build(fill="learned", knots=(1 << 22) + 1)
    assert ex.value.status == 1
    with pytest.raises(atg.AtgError) as ex:
        t.rejection_model(64)  # no learned fill built yet
    assert ex.value.status == 4
    t.rejection_build(fill="bisect")
    with pytest.raises(atg.AtgError):
        t.rejection_model(64)
    t.close()


def test_fill_then_sampling_matches_oracle(gpu, oracle):
    """Samples drawn from a learned-filled A equal the oracle's (the sampler
    reads the items and acceptances the fill wrote)."""
    torch = gpu
    n = 50021
    w = atg.generate_weights("heavy_mix", n, seed=8)
    t = atg.Table(n)
    t.set_weights(w)
    t.rejection_build(fill="learned", knots=256)
    W = t.rejection_weight_sum()
    _, A = oracle.rejection_build(w, W)
    count = 400_000
    out = torch.empty(count, dtype=torch.int32, device="cuda")
    t.rejection_sample(count, seed=5, out=out)
    exp, _ = oracle.sample_rejection(A, w, W, 5, count)
    np.testing.assert_array_equal(out.cpu().numpy(), exp)
    t.rejection_sample_sectioned(count, seed=6, section_slots=1024, out=out)
    exp = oracle.sample_rejection_sectioned(A, w, W, 1024, count, 6)
    np.testing.assert_array_equal(out.cpu().numpy(), exp)
    t.close()
