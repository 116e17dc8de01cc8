"""ctypes binding of libatg.so (the C ABI in include/atg.h).

Loading fails loudly when the library has not been built: there is no
fallback implementation of any kernel.
"""
import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libatg.so")

ATG_OK, ATG_EINVAL, ATG_ENOMEM, ATG_EDEVICE, ATG_ESTATE, ATG_ENOTSUP, ATG_ETIMEOUT, ATG_EIO = range(8)
STATUS_NAMES = {0: "ATG_OK", 1: "ATG_EINVAL", 2: "ATG_ENOMEM", 3: "ATG_EDEVICE", 4: "ATG_ESTATE",
                5: "ATG_ENOTSUP", 6: "ATG_ETIMEOUT", 7: "ATG_EIO"}
ATG_F64, ATG_F32 = 0, 1
ATG_AOS, ATG_SOA = 0, 1
ATG_MEM_HOST, ATG_MEM_DEVICE = 0, 1
ATG_SPLIT_BISECT, ATG_SPLIT_LEARNED, ATG_SPLIT_PARY = 0, 1, 2
SPLIT_METHODS = {"bisect": ATG_SPLIT_BISECT, "learned": ATG_SPLIT_LEARNED, "pary": ATG_SPLIT_PARY}
ATG_SAMPLE_BASIC, ATG_SAMPLE_SECTIONED = 0, 1
ATG_PACK_CHUNKED, ATG_PACK_BASIC, ATG_PACK_NO_WEIGHTS = 0, 1, 2
PACK_METHODS = {"chunked": ATG_PACK_CHUNKED, "basic": ATG_PACK_BASIC, "no_weights": ATG_PACK_NO_WEIGHTS}
ATG_FILL_DIRECT, ATG_FILL_BISECT, ATG_FILL_PARY, ATG_FILL_LEARNED = 0, 1, 2, 3
FILL_METHODS = {"direct": ATG_FILL_DIRECT, "bisect": ATG_FILL_BISECT, "pary": ATG_FILL_PARY,
                "learned": ATG_FILL_LEARNED}

DIST = {
    "sine": 1, "best_worst": 2, "uniform": 3, "ramp": 4, "powerlaw1": 5, "powerlaw2": 6,
    "one_heavy": 7, "one_light": 8, "powerlaw1_shuffled": 9, "powerlaw05_shuffled": 10,
    "unit_uniform": 100, "pareto": 101, "exponential": 102, "heavy_mix": 103, "zipf_shuffled": 104,
}


class BuildOpts(C.Structure):
    _fields_ = [("psa_plus", C.c_int32), ("chains", C.c_int32), ("section_size", C.c_int64),
                ("num_splits", C.c_int64), ("split_method", C.c_int32), ("learned_segments", C.c_int32),
                ("record_phases", C.c_int32), ("interleave", C.c_int32), ("batch_group_items", C.c_int64),
                ("pack_method", C.c_int32), ("reserved0", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("W", C.c_double), ("W_N", C.c_double), ("n", C.c_int64), ("numL", C.c_int64),
                ("numH", C.c_int64), ("R", C.c_int64), ("p", C.c_int64), ("num_sections", C.c_int64),
                ("section_len", C.c_int64), ("split_fallbacks", C.c_int64), ("learned_fallbacks", C.c_int64),
                ("degenerate", C.c_int32), ("monotone", C.c_int32), ("phase_ms", C.c_double * 5)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "phase_ms"}
        d["phase_ms"] = list(self.phase_ms)
        return d


class TableFileInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("W", C.c_double), ("layout", C.c_int32), ("source_dtype", C.c_int32),
                ("psa_plus", C.c_int32), ("chains", C.c_int32), ("split_method", C.c_int32),
                ("reserved", C.c_int32), ("section_size", C.c_int64), ("num_splits", C.c_int64),
                ("seed", C.c_uint64), ("checksum", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


# (name, restype, argtypes) for every entry point of include/atg.h
_P, _I32, _I64, _U64, _D = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
SIGNATURES = [
    ("atg_default_opts", None, [C.POINTER(BuildOpts)]),
    ("atg_create", _I32, [C.POINTER(_P), _I64, _I32, _I32, C.POINTER(BuildOpts)]),
    ("atg_destroy", None, [_P]),
    ("atg_set_weights", _I32, [_P, _P, _I32, _P]),
    ("atg_build", _I32, [_P, _P]),
    ("atg_build_phase", _I32, [_P, _I32, _P]),
    ("atg_sync", _I32, [_P, _P]),
    ("atg_post_build", _I32, [_P, _I32, C.POINTER(_D), C.POINTER(_I32)]),
    ("atg_get_stats", _I32, [_P, C.POINTER(Stats)]),
    ("atg_get_table", _I32, [_P, _P, _P, _I32, _P]),
    ("atg_table_device_ptr", _I32, [_P, C.POINTER(_P), C.POINTER(_P)]),
    ("atg_get_splits", _I32, [_P, _P, _P, _P, _I64]),
    ("atg_get_lh", _I32, [_P, _P, _P, _P, _P, _P, _P]),
    ("atg_sample", _I32, [_P, _I32, _U64, _U64, _I64, _P, _P, C.POINTER(_U64), _P]),
    ("atg_rejection_build", _I32, [_P, _P]),
    ("atg_rejection_build_fill", _I32, [_P, _I32, _I32, _P]),
    ("atg_rejection_fill_a", _I32, [_P, _P]),
    ("atg_rejection_fill_fallbacks", _I32, [_P, C.POINTER(_U64)]),
    ("atg_rejection_get_model", _I32, [_P, _P, _P, _I32]),
    ("atg_rejection_set_model", _I32, [_P, _P, _P, _I32]),
    ("atg_get_rejection_accept", _I32, [_P, _P, _U64]),
    ("atg_rejection_size", _I32, [_P, C.POINTER(_U64)]),
    ("atg_rejection_weight_sum", _I32, [_P, C.POINTER(_D)]),
    ("atg_get_rejection_array", _I32, [_P, _P, _U64]),
    ("atg_rejection_sample", _I32, [_P, _U64, _U64, _P, _P, C.POINTER(_U64), _P]),
    ("atg_rejection_sample_sectioned", _I32, [_P, _U64, _U64, _I64, _P, _P, C.POINTER(_U64), _P]),
    ("atg_generate_weights", _I32, [_I32, _I64, _U64, _D, _P]),
    ("atg_generate_batch", _I32, [_I32, _I64, _I64, _I64, _U64, _D, _P, _P, _I32]),
    ("atg_autotune", _I32, [_P, _I32, _I64, _I32, _I32, C.POINTER(BuildOpts), _I32, C.POINTER(BuildOpts), C.POINTER(_D)]),
    ("atg_autotune_shape", _I32, [_P, _I32, _I64, _I32, _I32, C.POINTER(BuildOpts), _I32, C.POINTER(BuildOpts),
                                  C.POINTER(_D)]),
    ("atg_batch_autotune", _I32, [_P, _I32, _I32, _D, C.POINTER(BuildOpts), C.POINTER(_D), C.POINTER(_D)]),
    ("atg_batch_create", _I32, [C.POINTER(_P), _P, _I64, _I32, C.POINTER(BuildOpts)]),
    ("atg_batch_destroy", None, [_P]),
    ("atg_batch_set_weights", _I32, [_P, _P, _I32, _P]),
    ("atg_batch_build", _I32, [_P, _P]),
    ("atg_batch_sync", _I32, [_P, _P]),
    ("atg_batch_get_tables", _I32, [_P, _P, _P, _P]),
    ("atg_batch_get_table", _I32, [_P, _I64, _P, _P, C.POINTER(_D)]),
    ("atg_batch_groups", _I32, [_P, _P, C.POINTER(_I64)]),
    ("atg_batch_sample", _I32, [_P, _U64, _U64, _P, _P]),
    ("atg_batch_latencies", _I32, [_P, _P]),
    ("atg_batch_use_graph", _I32, [_P, _I32]),
    ("atg_build_noalloc", _I32, [_P, _I32, _I64, _I32, C.POINTER(BuildOpts), _P, _P, _P, _P, _P]),
    ("atg_table_file_write", _I32, [C.c_char_p, C.POINTER(TableFileInfo), _P, _P]),
    ("atg_table_file_read", _I32, [C.c_char_p, C.POINTER(TableFileInfo), _P, _P, _I64]),
    ("atg_save_table", _I32, [_P, C.c_char_p, _U64]),
    ("atg_set_table", _I32, [_P, _P, _P, _D, _I32, _P]),
    ("atg_last_error", C.c_char_p, []),
    ("atg_version", C.c_char_p, []),
    ("atg_device_count", _I32, []),
]

_lib = None


class AtgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def lib():
    """The loaded libatg.so; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libatg.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != ATG_OK:
        raise AtgError(status, lib().atg_last_error().decode())
    return status
