"""The versioned table file (atg_table_file_write / atg_table_file_read):
round trip of the rows and the header, header-only reads, and the errors a
damaged or foreign file must raise.  Host-only: no GPU."""
import numpy as np
import pytest

import atg
from atg._lib import ATG_EINVAL, ATG_EIO


def table(n, seed=3):
    rng = np.random.default_rng(seed)
    w = rng.random(n) * 2.0
    a = rng.integers(0, n, n, dtype=np.int32)
    return w, a


INFO = {"W": 12.5, "layout": 1, "source_dtype": 0, "psa_plus": 1, "chains": 24, "split_method": 1,
        "section_size": 2048, "num_splits": 1234, "seed": 42}


def write(path, n=1000, **over):
    w, a = table(n)
    info = dict(INFO, n=n, **over)
    atg.write_table_file(path, info, w, a)
    return info, w, a


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 65537])
def test_round_trip(tmp_path, n):
    p = tmp_path / "t.atg"
    info, w, a = write(p, n)
    got, w2, a2 = atg.read_table_file(p)
    assert np.array_equal(w.view(np.uint64), w2.view(np.uint64)) and np.array_equal(a, a2)
    for k, v in info.items():
        assert got[k] == v, k
    assert got["checksum"] != 0
    assert p.stat().st_size == 128 + 12 * n


def test_header_only(tmp_path):
    p = tmp_path / "t.atg"
    info, _, _ = write(p, 5000)
    got, w, a = atg.read_table_file(p, rows=False)
    assert w is None and a is None and got["n"] == 5000 and got["W"] == info["W"]


def test_checksum_detects_a_changed_row(tmp_path):
    p = tmp_path / "t.atg"
    write(p, 1000)
    raw = bytearray(p.read_bytes())
    raw[128 + 8 * 500 + 3] ^= 0x10  # one bit of row 500's weight
    p.write_bytes(bytes(raw))
    with pytest.raises(atg.AtgError) as e:
        atg.read_table_file(p)
    assert e.value.status == ATG_EIO and "checksum" in str(e.value)
    atg.read_table_file(p, rows=False)  # the header is intact


def test_checksum_detects_a_changed_alias(tmp_path):
    p = tmp_path / "t.atg"
    write(p, 1000)
    raw = bytearray(p.read_bytes())
    raw[128 + 8 * 1000 + 4 * 999] ^= 0x01
    p.write_bytes(bytes(raw))
    with pytest.raises(atg.AtgError) as e:
        atg.read_table_file(p)
    assert e.value.status == ATG_EIO


def test_header_checksum(tmp_path):
    p = tmp_path / "t.atg"
    write(p, 100)
    raw = bytearray(p.read_bytes())
    raw[24] ^= 0x01  # inside W
    p.write_bytes(bytes(raw))
    with pytest.raises(atg.AtgError) as e:
        atg.read_table_file(p, rows=False)
    assert e.value.status == ATG_EIO and "header" in str(e.value)


def test_foreign_and_truncated_files(tmp_path):
    p = tmp_path / "x.bin"
    p.write_bytes(b"not a table" * 20)
    with pytest.raises(atg.AtgError) as e:
        atg.read_table_file(p)
    assert e.value.status == ATG_EIO
    q = tmp_path / "t.atg"
    write(q, 1000)
    q.write_bytes(q.read_bytes()[:128 + 8 * 1000 + 100])
    with pytest.raises(atg.AtgError) as e:
        atg.read_table_file(q)
    assert e.value.status == ATG_EIO and "truncated" in str(e.value)
    with pytest.raises(atg.AtgError) as e:
        atg.read_table_file(tmp_path / "missing.atg")
    assert e.value.status == ATG_EIO


def test_capacity_and_arguments(tmp_path):
    import ctypes as C
    p = tmp_path / "t.atg"
    write(p, 100)
    fi = atg._lib.TableFileInfo()
    w = np.empty(50, np.float64)
    a = np.empty(50, np.int32)
    rc = atg.lib().atg_table_file_read(str(p).encode(), C.byref(fi), w.ctypes.data, a.ctypes.data, 50)
    assert rc == ATG_EINVAL and b"capacity" in atg.lib().atg_last_error()
    with pytest.raises(atg.AtgError):
        atg.write_table_file(p, dict(INFO, n=0), np.empty(0), np.empty(0, np.int32))
