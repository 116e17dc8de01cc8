"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).  Test-only."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")


def _load():
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    return C.CDLL(LIB_PATH)


L = _load()
P = C.c_void_p


class Opts(C.Structure):
    _fields_ = [("psa_plus", C.c_int32), ("chains", C.c_int32), ("section_size", C.c_int64),
                ("num_splits", C.c_int64), ("threads", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("W", C.c_double), ("W_N", C.c_double), ("num_sections", C.c_int64), ("section_len", C.c_int64),
                ("numL", C.c_int64), ("numH", C.c_int64), ("R", C.c_int64), ("p", C.c_int64),
                ("wToSplit", C.c_double), ("nToSplit", C.c_int64), ("split_fallbacks", C.c_int64),
                ("degenerate", C.c_int32), ("monotone", C.c_int32), ("max_dL", C.c_double), ("min_dH", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


L.oracle_philox4x32_10.argtypes = [P, P, P]
L.oracle_section_layout.argtypes = [C.c_int64, C.c_int64, P, P]
L.oracle_section_sum.restype = C.c_double
L.oracle_section_sum.argtypes = [P, C.c_int64]
L.oracle_tree64_reduce.restype = C.c_double
L.oracle_tree64_reduce.argtypes = [P, C.c_int64]
L.oracle_weight_sum.restype = C.c_double
L.oracle_weight_sum.argtypes = [P, C.c_int64, C.c_int64, P]
L.oracle_local_scan.restype = C.c_double
L.oracle_local_scan.argtypes = [P, C.c_int64, P]
L.oracle_build.restype = C.c_int
L.oracle_build.argtypes = [P, C.c_int64, C.POINTER(Opts)] + [P] * 11 + [C.POINTER(Stats)]
L.oracle_split_one.restype = C.c_int
L.oracle_split_one.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double, P, C.c_int64, P, P, C.c_int64, P, P, P]
L.oracle_verify_table.argtypes = [P, C.c_int64, C.c_double, C.c_double, P, P, P, P, P, P]
L.oracle_sample_basic.argtypes = [P, P, C.c_int64, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, P]
L.oracle_sample_batch_table.argtypes = [P, P, C.c_int64, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, P]
L.oracle_binomial.restype = C.c_uint64
L.oracle_binomial.argtypes = [C.c_uint64, C.c_double, C.c_uint64, C.c_uint64]
L.oracle_section_counts.argtypes = [C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, P]
L.oracle_sample_sectioned.argtypes = [P, P, C.c_int64, C.c_double, C.c_int64, C.c_uint64, C.c_uint64, P]
L.oracle_rejection_build.restype = C.c_int
L.oracle_rejection_build.argtypes = [P, C.c_int64, C.c_double, P, P, P]
L.oracle_sample_rejection.argtypes = [P, C.c_uint64, P, C.c_int64, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, P, P]
L.oracle_ref_split_build.argtypes = [P, C.c_int64, C.c_int64, P, P, P]
L.oracle_ref_sweep_build.argtypes = [P, C.c_int64, P, P, P]
L.oracle_ref_stack_build.argtypes = [P, C.c_int64, P, P, P]


L.oracle_build_batch.argtypes = [P, C.c_int32, P, C.c_int64, P, C.c_int64, P, P, P, P, C.c_int32]
L.oracle_build_batch.restype = C.c_int
L.oracle_check_batch.argtypes = [P, C.c_int32, P, C.c_int64, P, P, P, P, P, P, P, C.c_int32]
L.oracle_sample_sectioned_subset.argtypes = [P, P, C.c_int64, C.c_double, C.c_int64, C.c_uint64, C.c_uint64, P,
                                             C.c_int64, P, P, P]
L.oracle_set_threads.argtypes = [C.c_int32]


def ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.oracle_philox4x32_10(c, k, o)
    return list(o)


def section_layout(n, param):
    g, l = C.c_int64(), C.c_int64()
    L.oracle_section_layout(int(n), int(param), C.byref(g), C.byref(l))
    return g.value, l.value


def weight_sum(w, section_param):
    w = np.ascontiguousarray(w, dtype=np.float64)
    G, _ = section_layout(len(w), section_param)
    sums = np.empty(G)
    W = L.oracle_weight_sum(ptr(w), len(w), int(section_param), ptr(sums))
    return W, sums


def local_scan(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    ex = np.empty(len(v))
    agg = L.oracle_local_scan(ptr(v), len(v), ptr(ex))
    return ex, agg


def build(w, psa_plus=True, chains=16, section_size=None, num_splits=81920, threads=8, want_lists=True):
    """Canonical build; section_size None = the reference's min(1000, N/10)."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    n = len(w)
    if section_size is None or section_size <= 0:
        section_size = max(1, min(1000, n // 10))
    o = Opts(1 if psa_plus else 0, int(chains), int(section_size), int(num_splits), int(threads))
    st = Stats()
    rw = np.zeros(n)
    ra = np.full(n, -1, np.int32)
    res = {"row_w": rw, "row_alias": ra}
    lists = [None] * 9
    if want_lists:
        cap = n + 2
        pcap = max(2, min(int(num_splits), n // 10)) + 1
        lists = [np.zeros(cap, np.int32), np.zeros(cap), np.zeros(cap), np.zeros(cap, np.int32), np.zeros(cap),
                 np.zeros(cap), np.zeros(pcap, np.int32), np.zeros(pcap, np.int32), np.zeros(pcap)]
    r = L.oracle_build(ptr(w), n, C.byref(o), ptr(rw), ptr(ra), *[ptr(x) for x in lists], C.byref(st))
    if r != 0:
        raise ValueError("oracle_build rejected the input")
    s = st.as_dict()
    res["stats"] = s
    if want_lists:
        nl, nh, p = s["numL"], s["numH"], s["p"]
        for k, a, m in zip(["L_item", "L_w", "L_pref", "H_item", "H_w", "H_pref"], lists[:6], [nl] * 3 + [nh] * 3):
            res[k] = a[:m]
        if p > 0:
            res["split_i"], res["split_j"], res["split_spill"] = lists[6][:p + 1], lists[7][:p + 1], lists[8][:p + 1]
    return res


def verify(w, W, W_N, row_w, row_alias):
    w = np.ascontiguousarray(w, dtype=np.float64)
    rel, ab, oor, ok = C.c_double(), C.c_double(), C.c_int64(), C.c_int32()
    L.oracle_verify_table(ptr(w), len(w), W, W_N, ptr(np.ascontiguousarray(row_w)),
                          ptr(np.ascontiguousarray(row_alias, dtype=np.int32)), C.byref(rel), C.byref(ab),
                          C.byref(oor), C.byref(ok))
    return {"max_rel": rel.value, "max_abs": ab.value, "alias_out_of_range": oor.value, "reference_ok": bool(ok.value)}


def split_one(k, p, nToSplit, wToSplit, Lp, Hp, Hw):
    i, j, s = C.c_int32(), C.c_int32(), C.c_double()
    found = L.oracle_split_one(int(k), int(p), int(nToSplit), float(wToSplit), ptr(Lp), len(Lp), ptr(Hp), ptr(Hw),
                               len(Hp), C.byref(i), C.byref(j), C.byref(s))
    return bool(found), i.value, j.value, s.value


def sample_basic(row_w, row_alias, W_N, seed, count, first=0):
    out = np.empty(int(count), np.int32)
    L.oracle_sample_basic(ptr(row_w), ptr(row_alias), len(row_w), W_N, int(seed), int(first), int(count), ptr(out))
    return out


def sample_batch_table(row_w, row_alias, W_N, seed, count, first):
    """Samples first .. first + count - 1 of a batch, drawn from one table."""
    out = np.empty(int(count), np.int32)
    L.oracle_sample_batch_table(ptr(row_w), ptr(row_alias), len(row_w), W_N, int(seed), int(first), int(count),
                                ptr(out))
    return out


def binomial(n, q, seed, node):
    return L.oracle_binomial(int(n), float(q), int(seed), int(node))


def section_counts(n, section_rows, total, seed):
    nsec = (n + section_rows - 1) // section_rows
    c = np.empty(nsec, np.uint64)
    L.oracle_section_counts(int(n), int(section_rows), int(total), int(seed), ptr(c))
    return c


def sample_sectioned(row_w, row_alias, W_N, section_rows, total, seed):
    out = np.empty(int(total), np.int32)
    L.oracle_sample_sectioned(ptr(row_w), ptr(row_alias), len(row_w), W_N, int(section_rows), int(total), int(seed),
                              ptr(out))
    return out


def rejection_build(w, W):
    w = np.ascontiguousarray(w, dtype=np.float64)
    incl = np.empty(len(w), np.uint64)
    sz = C.c_uint64()
    L.oracle_rejection_build(ptr(w), len(w), W, ptr(incl), None, C.byref(sz))
    A = np.empty(sz.value, np.int32)
    L.oracle_rejection_build(ptr(w), len(w), W, ptr(incl), ptr(A), C.byref(sz))
    return incl, A


def sample_rejection(A, w, W, seed, count, first=0):
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.empty(int(count), np.int32)
    tr = C.c_uint64()
    L.oracle_sample_rejection(ptr(A), len(A), ptr(w), len(w), W, int(seed), int(first), int(count), ptr(out),
                              C.byref(tr))
    return out, tr.value


def ref_split_build(w, p):
    n = len(w)
    wt = np.zeros(n + 2)
    wt[:n] = w
    rw = np.zeros(n)
    ra = np.full(n, -1, np.int32)
    W = C.c_double()
    L.oracle_ref_split_build(ptr(wt), n, int(p), C.byref(W), ptr(rw), ptr(ra))
    return W.value, rw, ra


def ref_sweep_build(w):
    n = len(w)
    wt = np.zeros(n + 2)
    wt[:n] = w
    rw = np.zeros(n)
    ra = np.full(n, -1, np.int32)
    W = C.c_double()
    L.oracle_ref_sweep_build(ptr(wt), n, C.byref(W), ptr(rw), ptr(ra))
    return W.value, rw, ra


def ref_stack_build(w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    n = len(w)
    rw = np.zeros(n)
    ra = np.full(n, -1, np.int32)
    W = C.c_double()
    L.oracle_ref_stack_build(ptr(w), n, C.byref(W), ptr(rw), ptr(ra))
    return W.value, rw, ra


def _batch_opts(psa_plus, chains, section_size, num_splits):
    return Opts(1 if psa_plus else 0, int(chains), int(section_size), int(num_splits), 1)


def build_batch(w, offsets, which=None, psa_plus=True, chains=16, section_size=0, num_splits=0, threads=16):
    """Rows (concatenated; rows of tables not in `which` stay 0 / -1) and W of
    the tables of a batch, each built as oracle.build would build it alone
    with the options the engine's batch derives per table (section_size 0:
    min(1000, n/10); num_splits < 0: one split per -num_splits items)."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    f32 = w.dtype == np.float32
    w = np.ascontiguousarray(w)
    total = int(offsets[-1])
    rw = np.zeros(total)
    ra = np.full(total, -1, np.int32)
    W = np.zeros(len(offsets) - 1)
    sel = None if which is None else np.ascontiguousarray(which, dtype=np.int64)
    o = _batch_opts(psa_plus, chains, section_size, num_splits)
    r = L.oracle_build_batch(ptr(w), 1 if f32 else 0, ptr(offsets), len(offsets) - 1, ptr(sel),
                             0 if sel is None else len(sel), C.byref(o), ptr(rw), ptr(ra), ptr(W), int(threads))
    if r != 0:
        raise ValueError("oracle_build_batch rejected a table")
    return rw, ra, W


def check_batch(w, offsets, row_w, row_alias, W_table, psa_plus=True, chains=16, section_size=0, num_splits=0,
                threads=16):
    """Per table: the canonical W and the reference's table check
    (AliasTable.cu:24-65, double precision): (W_canonical, ok, max_abs)."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    f32 = w.dtype == np.float32
    nt = len(offsets) - 1
    Wc = np.zeros(nt)
    ok = np.zeros(nt, np.int32)
    mabs = np.zeros(nt)
    o = _batch_opts(psa_plus, chains, section_size, num_splits)
    L.oracle_check_batch(ptr(np.ascontiguousarray(w)), 1 if f32 else 0, ptr(offsets), nt, C.byref(o),
                         ptr(np.ascontiguousarray(row_w)), ptr(np.ascontiguousarray(row_alias, dtype=np.int32)),
                         ptr(np.ascontiguousarray(W_table, dtype=np.float64)), ptr(Wc), ptr(ok), ptr(mabs),
                         int(threads))
    return Wc, ok.astype(bool), mabs


def sample_sectioned_subset(row_w, row_alias, W_N, section_rows, total, seed, sections):
    """Samples of the given sections of a sectioned run (concatenated), the
    first output position and the count of every given section."""
    sections = np.ascontiguousarray(sections, dtype=np.int64)
    n = len(row_w)
    counts = section_counts(n, section_rows, total, seed)
    out = np.empty(int(counts[sections].sum()), np.int32)
    first = np.empty(len(sections), np.uint64)
    cnt = np.empty(len(sections), np.uint64)
    L.oracle_sample_sectioned_subset(ptr(np.ascontiguousarray(row_w)), ptr(np.ascontiguousarray(row_alias)), n,
                                     float(W_N), int(section_rows), int(total), int(seed), ptr(sections),
                                     len(sections), ptr(out), ptr(first), ptr(cnt))
    return out, first, cnt


def set_threads(threads):
    L.oracle_set_threads(int(threads))


L.oracle_rejection_section_prefix.argtypes = [P, C.c_uint64, P, C.c_int64, C.c_double, C.c_int64, P]
L.oracle_rejection_section_counts.argtypes = [P, C.c_int64, C.c_uint64, C.c_uint64, P]
L.oracle_sample_rejection_sectioned.argtypes = [P, C.c_uint64, P, C.c_int64, C.c_double, C.c_int64, C.c_uint64,
                                                C.c_uint64, P, C.c_int64, P, P, P]


def rejection_section_prefix(A, w, W, section_slots):
    """Canonical exclusive prefix (nsec + 1 values) of the sections'
    acceptance masses."""
    A = np.ascontiguousarray(A, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    nsec = (len(A) + section_slots - 1) // section_slots
    pre = np.empty(nsec + 1)
    L.oracle_rejection_section_prefix(ptr(A), len(A), ptr(w), len(w), float(W), int(section_slots), ptr(pre))
    return pre


def rejection_section_counts(pre, total, seed):
    pre = np.ascontiguousarray(pre, dtype=np.float64)
    c = np.empty(len(pre) - 1, np.uint64)
    L.oracle_rejection_section_counts(ptr(pre), len(pre) - 1, int(total), int(seed), ptr(c))
    return c


def sample_rejection_sectioned(A, w, W, section_slots, total, seed, sections=None):
    """Sectioned rejection samples: all of them (sections None), or the
    samples of the given sections at their positions of the full run (other
    positions untouched), with each given section's first position and count."""
    A = np.ascontiguousarray(A, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.empty(int(total), np.int32)
    sel = None if sections is None else np.ascontiguousarray(sections, dtype=np.int64)
    k = 0 if sel is None else len(sel)
    first = np.empty(max(k, 1), np.uint64)
    cnt = np.empty(max(k, 1), np.uint64)
    L.oracle_sample_rejection_sectioned(ptr(A), len(A), ptr(w), len(w), float(W), int(section_slots), int(total),
                                        int(seed), ptr(sel), k, ptr(out), ptr(first) if sel is not None else None,
                                        ptr(cnt) if sel is not None else None)
    return (out, first[:k], cnt[:k]) if sel is not None else out


L.oracle_split_inverse.argtypes = [C.c_int64, C.c_int64, C.c_double, P, C.c_int64, P, P, C.c_int64, P, P, P]


def split_inverse(p, nToSplit, wToSplit, Lp, Hp, Hw):
    """Splits by SplitMethodInverse's brute force (every heavy position
    against every split); i = j = -1 where no position satisfies FOUND."""
    si = np.empty(p + 1, np.int32)
    sj = np.empty(p + 1, np.int32)
    ss = np.empty(p + 1)
    Lp, Hp, Hw = (np.ascontiguousarray(x, dtype=np.float64) for x in (Lp, Hp, Hw))
    L.oracle_split_inverse(int(p), int(nToSplit), float(wToSplit), ptr(Lp), len(Lp), ptr(Hp), ptr(Hw), len(Hp),
                           ptr(si), ptr(sj), ptr(ss))
    return si, sj, ss


L.oracle_rejection_fill_bisect.argtypes = [P, C.c_int64, C.c_uint64, P]
L.oracle_rejection_model.argtypes = [P, C.c_int64, C.c_uint64, C.c_int32, P, P]


def rejection_fill_bisect(incl):
    """A by the reference's per-slot bisection over the inclusive counts."""
    incl = np.ascontiguousarray(incl, dtype=np.uint64)
    A = np.empty(int(incl[-1]), np.int32)
    L.oracle_rejection_fill_bisect(ptr(incl), len(incl), len(A), ptr(A))
    return A


def rejection_model(incl, knots):
    """Learned inverse-CDF model over the inclusive counts: knot items (knots
    + 1) and per-segment error bounds (knots), by brute force over every slot."""
    incl = np.ascontiguousarray(incl, dtype=np.uint64)
    q = np.empty(knots + 1, np.int32)
    e = np.empty(knots, np.uint32)
    L.oracle_rejection_model(ptr(incl), len(incl), int(incl[-1]), int(knots), ptr(q), ptr(e))
    return q, e
