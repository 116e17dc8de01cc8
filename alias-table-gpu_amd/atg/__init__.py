"""Python view of the atg C ABI (include/atg.h) for tests and the benchmark.

`Table` wraps one alias table (AliasTableSplitGpu + its samplers and the
rejection structure of the reference); device buffers are torch tensors or raw
device pointers.  Every call goes through libatg.so.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import ATG_EINVAL as _EINVAL
from ._lib import (ATG_AOS, ATG_F32, ATG_F64, ATG_MEM_DEVICE, ATG_MEM_HOST, ATG_SAMPLE_BASIC,
                   ATG_SAMPLE_SECTIONED, ATG_SOA, ATG_SPLIT_BISECT, ATG_SPLIT_LEARNED, DIST, AtgError,
                   BuildOpts, Stats, TableFileInfo, check, lib)

__all__ = ["Table", "Batch", "autotune", "autotune_shape", "generate_weights", "generate_batch", "default_opts", "lib", "AtgError", "DIST",
           "write_table_file", "read_table_file"]


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return C.c_void_p(a)
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


_TORCH_DTYPES = {"f64": "torch.float64", "f32": "torch.float32", "i32": "torch.int32", "u32": "torch.uint32",
                 "i64": "torch.int64"}


def _is_device(x):
    return hasattr(x, "is_cuda") and bool(x.is_cuda)


def _device_buffer(x, kind, numel, name, exact=False):
    """Validates a device buffer handed to a kernel: a CUDA tensor, contiguous,
    of element type `kind` ("f64", "f32", "i32", "u32") and at least (exactly,
    with exact=True) `numel` elements.  None passes through."""
    if x is None:
        return None
    if not _is_device(x):
        raise AtgError(_EINVAL, f"{name}: a device (torch CUDA) tensor is required, got {type(x).__name__}")
    if not x.is_contiguous():
        raise AtgError(_EINVAL, f"{name}: tensor must be contiguous")
    want = _TORCH_DTYPES[kind]
    got = str(x.dtype)
    # hist may be int32 or uint32 (same 4-byte counters)
    if got != want and not (kind == "u32" and got == _TORCH_DTYPES["i32"]):
        raise AtgError(_EINVAL, f"{name}: dtype {got}, expected {want}")
    n = x.numel()
    if (exact and n != numel) or n < numel:
        raise AtgError(_EINVAL, f"{name}: {n} elements, {'exactly ' if exact else 'at least '}{numel} required")
    return x


def _split_id(split):
    if split not in _lib.SPLIT_METHODS:
        raise AtgError(_EINVAL, f"unknown split method {split!r}; one of {sorted(_lib.SPLIT_METHODS)}")
    return _lib.SPLIT_METHODS[split]


def _pack_id(pack):
    if pack not in _lib.PACK_METHODS:
        raise AtgError(_EINVAL, f"unknown pack method {pack!r}; one of {sorted(_lib.PACK_METHODS)}")
    return _lib.PACK_METHODS[pack]


def _split_name(method):
    return {v: k for k, v in _lib.SPLIT_METHODS.items()}[int(method)]


def _stream(s):
    if s is None:
        return None
    if hasattr(s, "cuda_stream"):
        return C.c_void_p(s.cuda_stream)
    return C.c_void_p(int(s))


def default_opts():
    o = BuildOpts()
    lib().atg_default_opts(C.byref(o))
    return o


def generate_weights(dist, n, seed=42, param=0.0):
    """Host FP64 weights (atg_generate_weights)."""
    w = np.empty(int(n), dtype=np.float64)
    d = DIST[dist] if isinstance(dist, str) else int(dist)
    check(lib().atg_generate_weights(d, int(n), int(seed), float(param), _ptr(w)))
    return w


def generate_batch(dist, ntables, lo, hi, seed=42, param=0.0, dtype="f64"):
    """Offsets (ntables + 1) and concatenated host weights of a batch of tables
    with sizes log-uniform over [lo, hi] (atg_generate_batch)."""
    d = DIST[dist] if isinstance(dist, str) else int(dist)
    offsets = np.empty(int(ntables) + 1, np.int64)
    check(lib().atg_generate_batch(d, int(ntables), int(lo), int(hi), int(seed), float(param), _ptr(offsets),
                                   None, ATG_F64))
    w = np.empty(int(offsets[-1]), np.float32 if dtype == "f32" else np.float64)
    check(lib().atg_generate_batch(d, int(ntables), int(lo), int(hi), int(seed), float(param), _ptr(offsets),
                                   _ptr(w), ATG_F32 if dtype == "f32" else ATG_F64))
    return offsets, w


def _tune_args(w, dtype):
    if _is_device(w):
        _device_buffer(w, dtype, w.numel(), "autotune: weights")
        return _ptr(w), ATG_MEM_DEVICE, w.numel(), w
    a = np.ascontiguousarray(w, dtype=np.float32 if dtype == "f32" else np.float64)
    return _ptr(a), ATG_MEM_HOST, a.size, a


def _shape_opts(psa_plus, chains, section_size, num_splits, split, learned_segments):
    o = default_opts()
    o.psa_plus = 1 if psa_plus else 0
    o.chains = int(chains)
    o.section_size = int(section_size)
    o.num_splits = int(num_splits)
    o.split_method = _split_id(split)
    o.learned_segments = int(learned_segments)
    return o


def autotune(w, dtype="f64", layout="aos", psa_plus=True, chains=16, section_size=0, num_splits=0, reps=3):
    """Knobs that leave the table unchanged (split method, learned knots,
    split/pack interleave) chosen by atg_autotune for these weights at the
    given build shape, and the median build time in ms (negative when it came
    from the ATG_TUNE_CACHE file)."""
    o = _shape_opts(psa_plus, chains, section_size, num_splits, "bisect", 4096)
    best = BuildOpts()
    ms = C.c_double(0.0)
    ptr, mem, n, _keep = _tune_args(w, dtype)
    check(lib().atg_autotune(ptr, mem, int(n), ATG_F32 if dtype == "f32" else ATG_F64,
                             ATG_SOA if layout == "soa" else ATG_AOS, C.byref(o), int(reps), C.byref(best),
                             C.byref(ms)))
    return {"split": _split_name(best.split_method),
            "learned_segments": int(best.learned_segments), "interleave": bool(best.interleave)}, ms.value


def autotune_shape(w, dtype="f64", layout="aos", psa_plus=True, split="bisect", learned_segments=4096, reps=3):
    """Build shape (section_size, chains, num_splits) chosen by
    atg_autotune_shape; the shape changes the table, so it is an explicit
    input of Table(), and the median build time in ms."""
    o = _shape_opts(psa_plus, 16, 0, 0, split, learned_segments)
    best = BuildOpts()
    ms = C.c_double(0.0)
    ptr, mem, n, _keep = _tune_args(w, dtype)
    check(lib().atg_autotune_shape(ptr, mem, int(n), ATG_F32 if dtype == "f32" else ATG_F64,
                                   ATG_SOA if layout == "soa" else ATG_AOS, C.byref(o), int(reps), C.byref(best),
                                   C.byref(ms)))
    return {"section_size": int(best.section_size), "chains": int(best.chains), "num_splits": int(best.num_splits),
            "split": split, "psa_plus": psa_plus}, ms.value


def build_noalloc(w, rows, aliases=None, W_out=None, error_out=None, dtype="f64", layout="aos", psa_plus=True,
                  chains=16, section_size=0, num_splits=0, split="bisect", learned_segments=4096, stream=None,
                  pack="chunked"):
    """Builds the table of the device weights `w` into caller-owned device
    buffers with no allocation (atg_build_noalloc): AoS rows as a float64
    tensor of 2 n elements (16-byte rows), or SoA row weights (float64, n) and
    `aliases` (int32, n).  W_out: float64 device tensor of 1 element or None;
    error_out: int32 device tensor of 1 element or None."""
    n = w.numel() if _is_device(w) else 0
    _device_buffer(w, dtype, n, "build_noalloc: weights", exact=True)
    if layout == "aos":
        _device_buffer(rows, "f64", 2 * n, "build_noalloc: rows", exact=True)
    else:
        _device_buffer(rows, "f64", n, "build_noalloc: weights out", exact=True)
        if aliases is None:
            raise AtgError(_EINVAL, "build_noalloc: SoA needs aliases")
        _device_buffer(aliases, "i32", n, "build_noalloc: aliases", exact=True)
    _device_buffer(W_out, "f64", 1, "build_noalloc: W_out")
    _device_buffer(error_out, "i32", 1, "build_noalloc: error_out")
    o = _shape_opts(psa_plus, chains, section_size, num_splits, split, learned_segments)
    o.pack_method = _pack_id(pack)
    check(lib().atg_build_noalloc(_ptr(w), ATG_F32 if dtype == "f32" else ATG_F64, int(n),
                                  ATG_SOA if layout == "soa" else ATG_AOS, C.byref(o), _ptr(rows), _ptr(aliases),
                                  _ptr(W_out), _ptr(error_out), _stream(stream)))


def write_table_file(path, info, w, alias):
    """Writes a table file (atg_table_file_write): info is a dict with n, W and
    optionally the build shape (layout, source_dtype, psa_plus, chains,
    split_method, section_size, num_splits) and seed."""
    fi = TableFileInfo()
    for k, _ in TableFileInfo._fields_:
        if k in info and k not in ("checksum", "reserved"):
            setattr(fi, k, info[k])
    wa = np.ascontiguousarray(w, dtype=np.float64)
    aa = np.ascontiguousarray(alias, dtype=np.int32)
    assert wa.size == aa.size == fi.n
    check(lib().atg_table_file_write(str(path).encode(), C.byref(fi), _ptr(wa), _ptr(aa)))


def read_table_file(path, rows=True):
    """(info dict, row weights, aliases) of a table file; rows=False reads the header only."""
    fi = TableFileInfo()
    check(lib().atg_table_file_read(str(path).encode(), C.byref(fi), None, None, 0))
    if not rows:
        return fi.as_dict(), None, None
    w = np.empty(fi.n, np.float64)
    a = np.empty(fi.n, np.int32)
    check(lib().atg_table_file_read(str(path).encode(), C.byref(fi), _ptr(w), _ptr(a), fi.n))
    return fi.as_dict(), w, a


class Table:
    """One alias table on the GPU (AliasTableSplitGpu<TableStorage> of the reference)."""

    def __init__(self, n, dtype="f64", layout="aos", psa_plus=True, chains=16, section_size=0, num_splits=0,
                 split="bisect", learned_segments=4096, record_phases=False, interleave=False, pack="chunked"):
        o = default_opts()
        o.interleave = 1 if interleave else 0
        o.pack_method = _pack_id(pack)
        o.psa_plus = 1 if psa_plus else 0
        o.chains = int(chains)
        o.section_size = int(section_size)
        o.num_splits = int(num_splits)
        o.split_method = _split_id(split)
        o.learned_segments = int(learned_segments)
        o.record_phases = 1 if record_phases else 0
        self.n = int(n)
        self.dtype = dtype
        self.layout = layout
        h = C.c_void_p()
        check(lib().atg_create(C.byref(h), self.n, ATG_F32 if dtype == "f32" else ATG_F64,
                               ATG_SOA if layout == "soa" else ATG_AOS, C.byref(o)))
        self.h = h
        self._keep = None

    def close(self):
        if self.h:
            lib().atg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- construction ---------------------------------------------------
    def set_weights(self, w, stream=None):
        if _is_device(w):
            _device_buffer(w, self.dtype, self.n, "set_weights: weights", exact=True)
            self._keep = w
            check(lib().atg_set_weights(self.h, _ptr(w), ATG_MEM_DEVICE, _stream(stream)))
        else:
            want = np.float32 if self.dtype == "f32" else np.float64
            a = np.ascontiguousarray(w, dtype=want)
            if a.size != self.n:
                raise AtgError(_EINVAL, f"set_weights: {a.size} weights for a table of {self.n}")
            check(lib().atg_set_weights(self.h, _ptr(a), ATG_MEM_HOST, _stream(stream)))

    def build(self, stream=None):
        check(lib().atg_build(self.h, _stream(stream)))

    PHASES = ("sum", "partition", "split", "pack")
    _PHASE_IDS = {"sum": 0, "partition": 1, "split": 2, "pack": 3, "partition_local": 4, "partition_place": 5}

    def build_phase(self, phase, stream=None):
        """One phase of build(): 'sum', 'partition' (PSA+ and prefix sums), 'split' or 'pack';
        'partition' also runs as its two passes 'partition_local' then 'partition_place'."""
        if phase not in self._PHASE_IDS:
            raise AtgError(_EINVAL, f"build_phase: unknown phase {phase!r}")
        check(lib().atg_build_phase(self.h, self._PHASE_IDS[phase], _stream(stream)))

    def sync(self, stream=None):
        check(lib().atg_sync(self.h, _stream(stream)))

    def stats(self):
        st = Stats()
        check(lib().atg_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def post_build(self, verify=True):
        rel = C.c_double(0.0)
        ok = C.c_int32(0)
        check(lib().atg_post_build(self.h, 1 if verify else 0, C.byref(rel), C.byref(ok)))
        return rel.value, bool(ok.value)

    def table(self):
        w = np.empty(self.n, dtype=np.float64)
        a = np.empty(self.n, dtype=np.int32)
        check(lib().atg_get_table(self.h, _ptr(w), _ptr(a), ATG_MEM_HOST, None))
        return w, a

    def save(self, path, seed=0):
        """Writes the built table to a table file (atg_save_table)."""
        check(lib().atg_save_table(self.h, str(path).encode(), int(seed)))

    def set_table(self, w, alias, W, stream=None):
        """Loads rows (e.g. from read_table_file) for sampling (atg_set_table)."""
        if _is_device(w) or _is_device(alias):
            _device_buffer(w, "f64", self.n, "set_table: weights", exact=True)
            _device_buffer(alias, "i32", self.n, "set_table: aliases", exact=True)
            self._keep = (w, alias)
            check(lib().atg_set_table(self.h, _ptr(w), _ptr(alias), float(W), ATG_MEM_DEVICE, _stream(stream)))
        else:
            wa = np.ascontiguousarray(w, dtype=np.float64)
            aa = np.ascontiguousarray(alias, dtype=np.int32)
            if not wa.size == aa.size == self.n:
                raise AtgError(_EINVAL, f"set_table: {wa.size} weights and {aa.size} aliases for a table of {self.n}")
            check(lib().atg_set_table(self.h, _ptr(wa), _ptr(aa), float(W), ATG_MEM_HOST, _stream(stream)))

    def device_table_ptr(self):
        p0, p1 = C.c_void_p(), C.c_void_p()
        check(lib().atg_table_device_ptr(self.h, C.byref(p0), C.byref(p1)))
        return p0.value, p1.value

    def splits(self):
        p = self.stats()["p"]
        i = np.empty(p + 1, np.int32)
        j = np.empty(p + 1, np.int32)
        s = np.empty(p + 1, np.float64)
        check(lib().atg_get_splits(self.h, _ptr(i), _ptr(j), _ptr(s), p + 1))
        return i, j, s

    def lh(self):
        st = self.stats()
        nl, nh = st["numL"], st["numH"]
        out = {"L_item": np.empty(nl, np.int32), "L_w": np.empty(nl), "L_pref": np.empty(nl),
               "H_item": np.empty(nh, np.int32), "H_w": np.empty(nh), "H_pref": np.empty(nh)}
        check(lib().atg_get_lh(self.h, *[_ptr(out[k]) for k in
                                         ("L_item", "L_w", "L_pref", "H_item", "H_w", "H_pref")]))
        return out

    # -- sampling ----------------------------------------------------------
    def sample(self, count, seed, kind="sectioned", section_rows=4096, out=None, hist=None, checksum=False,
               stream=None):
        if kind not in ("sectioned", "basic"):
            raise AtgError(_EINVAL, f"sample: unknown kind {kind!r}")
        k = ATG_SAMPLE_SECTIONED if kind == "sectioned" else ATG_SAMPLE_BASIC
        _device_buffer(out, "i32", int(count), "sample: out")
        _device_buffer(hist, "u32", self.n, "sample: hist")
        cs = C.c_uint64(0)
        check(lib().atg_sample(self.h, k, int(count), int(seed), int(section_rows), _ptr(out), _ptr(hist),
                               C.byref(cs) if checksum else None, _stream(stream)))
        return cs.value if checksum else None

    def rejection_build(self, stream=None, fill="direct", knots=0):
        """RejectionSampling::build with the fill of A chosen by name
        (atg_rejection_build_fill): "direct", "bisect", "pary" or "learned"
        (knots = the learned model's segments, 0 = 4096)."""
        if fill not in _lib.FILL_METHODS:
            raise AtgError(_EINVAL, f"rejection_build: unknown fill {fill!r}, expected one of {sorted(_lib.FILL_METHODS)}")
        check(lib().atg_rejection_build_fill(self.h, _lib.FILL_METHODS[fill], int(knots), _stream(stream)))

    def rejection_fill(self, stream=None):
        """RejectionSampling::fillA alone: the last build's fill again (atg_rejection_fill_a)."""
        check(lib().atg_rejection_fill_a(self.h, _stream(stream)))

    def rejection_fill_fallbacks(self):
        v = C.c_uint64(0)
        check(lib().atg_rejection_fill_fallbacks(self.h, C.byref(v)))
        return v.value

    def rejection_model(self, knots):
        """The learned fill's model: knot items (knots + 1) and error bounds (knots)."""
        q = np.empty(int(knots) + 1, np.int32)
        e = np.empty(int(knots), np.uint32)
        check(lib().atg_rejection_get_model(self.h, _ptr(q), _ptr(e), int(knots)))
        return q, e

    def rejection_set_model(self, knot_items, err_bounds):
        q = np.ascontiguousarray(knot_items, dtype=np.int32)
        e = np.ascontiguousarray(err_bounds, dtype=np.uint32)
        if len(q) != len(e) + 1:
            raise AtgError(_EINVAL, f"rejection_set_model: {len(q)} knot items for {len(e)} bounds")
        check(lib().atg_rejection_set_model(self.h, _ptr(q), _ptr(e), len(e)))

    def rejection_accept(self):
        """The acceptance value of every slot of A."""
        sz = self.rejection_size()
        a = np.empty(sz, np.float64)
        check(lib().atg_get_rejection_accept(self.h, _ptr(a), sz))
        return a

    def rejection_size(self):
        v = C.c_uint64(0)
        check(lib().atg_rejection_size(self.h, C.byref(v)))
        return v.value

    def rejection_sample_sectioned(self, count, seed, section_slots=4096, out=None, hist=None, checksum=False,
                                   stream=None):
        """Rejection sampling by LDS-staged sections of A (atg_rejection_sample_sectioned)."""
        _device_buffer(out, "i32", int(count), "rejection_sample_sectioned: out")
        _device_buffer(hist, "u32", self.n, "rejection_sample_sectioned: hist")
        cs = C.c_uint64(0)
        check(lib().atg_rejection_sample_sectioned(self.h, int(count), int(seed), int(section_slots), _ptr(out),
                                                   _ptr(hist), C.byref(cs) if checksum else None, _stream(stream)))
        return cs.value if checksum else None

    def rejection_weight_sum(self):
        """W the rejection structure was built with (atg_rejection_weight_sum)."""
        v = C.c_double(0.0)
        check(lib().atg_rejection_weight_sum(self.h, C.byref(v)))
        return v.value

    def rejection_array(self):
        sz = self.rejection_size()
        A = np.empty(sz, np.int32)
        check(lib().atg_get_rejection_array(self.h, _ptr(A), sz))
        return A

    def rejection_sample(self, count, seed, out=None, hist=None, checksum=False, stream=None):
        _device_buffer(out, "i32", int(count), "rejection_sample: out")
        _device_buffer(hist, "u32", self.n, "rejection_sample: hist")
        cs = C.c_uint64(0)
        check(lib().atg_rejection_sample(self.h, int(count), int(seed), _ptr(out), _ptr(hist),
                                         C.byref(cs) if checksum else None, _stream(stream)))
        return cs.value if checksum else None


class Batch:
    """Many independent tables built together, one launch per phase (atg_batch_*).

    Table t owns weights[offsets[t]:offsets[t+1]] and is built exactly as a
    `Table` with the same options would build it alone."""

    def __init__(self, offsets, dtype="f64", psa_plus=True, chains=16, section_size=0, num_splits=0,
                 split="bisect", learned_segments=4096, group_items=0):
        o = default_opts()
        o.batch_group_items = int(group_items)
        o.psa_plus = 1 if psa_plus else 0
        o.chains = int(chains)
        o.section_size = int(section_size)
        o.num_splits = int(num_splits)
        o.split_method = _split_id(split)
        o.learned_segments = int(learned_segments)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.ntables = self.offsets.size - 1
        self.total = int(self.offsets[-1])
        self.dtype = dtype
        h = C.c_void_p()
        check(lib().atg_batch_create(C.byref(h), _ptr(self.offsets), self.ntables,
                                     ATG_F32 if dtype == "f32" else ATG_F64, C.byref(o)))
        self.h = h
        self._keep = None

    def close(self):
        if self.h:
            lib().atg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, w, stream=None):
        if _is_device(w):
            _device_buffer(w, self.dtype, self.total, "Batch.set_weights: weights", exact=True)
            self._keep = w
            check(lib().atg_batch_set_weights(self.h, _ptr(w), ATG_MEM_DEVICE, _stream(stream)))
        else:
            want = np.float32 if self.dtype == "f32" else np.float64
            a = np.ascontiguousarray(w, dtype=want)
            if a.size != self.total:
                raise AtgError(_EINVAL, f"Batch.set_weights: {a.size} weights for {self.total} items")
            check(lib().atg_batch_set_weights(self.h, _ptr(a), ATG_MEM_HOST, _stream(stream)))

    def build(self, stream=None):
        check(lib().atg_batch_build(self.h, _stream(stream)))

    def use_graph(self, enable=True):
        """Captures the build into a hipGraph replayed by build() (atg_batch_use_graph)."""
        check(lib().atg_batch_use_graph(self.h, 1 if enable else 0))

    def sync(self, stream=None):
        check(lib().atg_batch_sync(self.h, _stream(stream)))

    def tables(self):
        """(row weights, aliases) of all tables concatenated, and W per table."""
        w = np.empty(self.total, np.float64)
        a = np.empty(self.total, np.int32)
        W = np.empty(self.ntables, np.float64)
        check(lib().atg_batch_get_tables(self.h, _ptr(w), _ptr(a), _ptr(W)))
        return w, a, W

    def table(self, t):
        """(row weights, aliases, W) of table t alone (atg_batch_get_table)."""
        n = int(self.offsets[t + 1] - self.offsets[t])
        w = np.empty(n, np.float64)
        a = np.empty(n, np.int32)
        W = C.c_double(0.0)
        check(lib().atg_batch_get_table(self.h, int(t), _ptr(w), _ptr(a), C.byref(W)))
        return w, a, W.value

    def groups(self):
        """Group of every table (the launch groups of batch_group_items)."""
        g = np.empty(self.ntables, np.int32)
        ng = C.c_int64(0)
        check(lib().atg_batch_groups(self.h, _ptr(g), C.byref(ng)))
        return g, ng.value

    def sample(self, per_table, seed, out, stream=None):
        """per_table samples of every table into the device buffer `out`
        (int32, ntables * per_table; item indices local to each table)."""
        if out is None:
            raise AtgError(_EINVAL, "Batch.sample: out is required")
        _device_buffer(out, "i32", self.ntables * int(per_table), "Batch.sample: out")
        check(lib().atg_batch_sample(self.h, int(per_table), int(seed), _ptr(out), _stream(stream)))

    def autotune(self, reps=3, makespan_slack=0.05, tune_split=True):
        """Tunes the launch groups and (tune_split) the split method in place
        (atg_batch_autotune); returns the chosen split and group size, the
        median makespan and the 99th-percentile table latency in ms."""
        o = BuildOpts()
        ms, p99 = C.c_double(0.0), C.c_double(0.0)
        check(lib().atg_batch_autotune(self.h, int(reps), 1 if tune_split else 0, float(makespan_slack), C.byref(o),
                                       C.byref(ms), C.byref(p99)))
        return {"split": _split_name(o.split_method),
                "group_items": int(o.batch_group_items), "makespan_ms": ms.value, "p99_latency_ms": p99.value}

    def latencies_ns(self):
        v = np.empty(self.ntables, np.float64)
        check(lib().atg_batch_latencies(self.h, _ptr(v)))
        return v
