"""Table files on the GPU path: a built table saved with atg_save_table reads
back bit-identically with its W and build shape, and a table loaded into a
fresh handle with atg_set_table samples exactly as the original."""
import numpy as np
import pytest

import atg

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("layout", ["aos", "soa"])
def test_save_read_and_sample_loaded(gpu, tmp_path, layout):
    torch = gpu
    n = 200_003
    w = atg.generate_weights("pareto", n, seed=5, param=1.1)
    t = atg.Table(n, layout=layout, chains=24, section_size=2048, num_splits=n // 38)
    t.set_weights(w)
    t.build()
    t.sync()
    st = t.stats()
    rw, ra = t.table()
    path = tmp_path / "pareto.atg"
    t.save(path, seed=5)
    info, fw, fa = atg.read_table_file(path)
    assert np.array_equal(bits(fw), bits(rw)) and np.array_equal(fa, ra)
    assert info["n"] == n and info["W"] == st["W"] and info["num_splits"] == st["p"]
    assert info["chains"] == 24 and info["section_size"] == 2048 and info["seed"] == 5
    assert info["layout"] == (1 if layout == "soa" else 0)

    u = atg.Table(n, layout=layout)
    u.set_table(fw, fa, info["W"])
    gw, ga = u.table()
    assert np.array_equal(bits(gw), bits(rw)) and np.array_equal(ga, ra)
    for kind in ("basic", "sectioned"):
        o1 = torch.empty(1 << 20, dtype=torch.int32, device="cuda")
        o2 = torch.empty_like(o1)
        t.sample(o1.numel(), seed=11, kind=kind, out=o1)
        u.sample(o2.numel(), seed=11, kind=kind, out=o2)
        torch.cuda.synchronize()
        assert torch.equal(o1, o2), kind


def test_set_table_from_device_and_rebuild(gpu):
    torch = gpu
    n = 50_000
    w = atg.generate_weights("exponential", n, seed=8)
    t = atg.Table(n)
    t.set_weights(w)
    t.build()
    t.sync()
    rw, ra = t.table()
    u = atg.Table(n)
    dw = torch.from_numpy(rw).cuda()
    da = torch.from_numpy(ra).cuda()
    u.set_table(dw, da, t.stats()["W"])
    gw, ga = u.table()
    assert np.array_equal(bits(gw), bits(rw)) and np.array_equal(ga, ra)
    # A build from weights replaces the loaded rows.
    u.set_weights(w)
    u.build()
    u.sync()
    gw, ga = u.table()
    assert np.array_equal(bits(gw), bits(rw)) and np.array_equal(ga, ra)


def test_set_table_rejects_bad_W(gpu):
    u = atg.Table(10)
    with pytest.raises(atg.AtgError):
        u.set_table(np.ones(10), np.zeros(10, np.int32), 0.0)
    with pytest.raises(atg.AtgError):
        u.sample(10, seed=1, kind="basic", out=gpu.empty(10, dtype=gpu.int32, device="cuda"))
