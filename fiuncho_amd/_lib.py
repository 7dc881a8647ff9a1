"""ctypes binding of libfiuncho_amd.so (the C ABI in include/fiuncho_amd.h).

The library is built in-tree (``fiuncho_amd/libfiuncho_amd.so``) by
``__graft_entry__.build()`` / ``make -C fiuncho_amd/csrc``.  There is no CPU
fallback: if the library is missing, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfiuncho_amd.so")

FIU_OK = 0
FIU_ERR_INVALID = 1
FIU_ERR_DEVICE = 2
FIU_ERR_NOMEM = 3
FIU_ERR_UNSUPPORTED = 4
FIU_ERR_IO = 5
FIU_ERR_REJECTED = 6
FIU_ERR_CANCELLED = 7
FIU_ERR_PENDING = 8

ALGO_AUTO, ALGO_DIRECT, ALGO_MARGINAL, ALGO_MFMA = 0, 1, 2, 3

# Every symbol include/fiuncho_amd.h declares (checked by the CPU tests).
EXPORTS = (
    "fiu_last_error", "fiu_version", "fiu_set_device",
    "fiu_dataset_create", "fiu_dataset_read", "fiu_dataset_from_genotypes", "fiu_read_files", "fiu_dataset_info", "fiu_dataset_words",
    "fiu_dataset_destroy", "fiu_config_init", "fiu_search", "fiu_search_static", "fiu_search_static_bytes",
    "fiu_ct_size",
    "fiu_combine_batch", "fiu_contingency_batch", "fiu_mi_batch",
    "fiu_batch_config_init", "fiu_search_batch", "fiu_server_create", "fiu_server_submit",
    "fiu_server_wait", "fiu_server_query", "fiu_server_destroy", "fiu_set_tuning", "fiu_get_tuning",
    "fiu_synth_spec_init", "fiu_synth_generate", "fiu_dataset_synth",
    "fiu_result_str", "fiu_result_serialize", "fiu_result_deserialize", "fiu_write_results",
)


class FiuError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"fiuncho_amd error {status}: {message}")
        self.status = status


class Config(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("algorithm", C.c_int32),
                ("chunk_combos", C.c_uint64), ("candidate_cap", C.c_uint32),
                ("flags", C.c_uint32), ("stream", C.c_void_p), ("variant", C.c_int32),
                ("reserved", C.c_int32)]


class TuneEntry(C.Structure):
    _fields_ = [("order", C.c_uint16), ("batch_unit", C.c_uint16), ("snps_lo", C.c_uint32),
                ("snps_hi", C.c_uint32), ("samples_lo", C.c_uint32), ("samples_hi", C.c_uint32),
                ("algorithm", C.c_int32), ("variant", C.c_int32), ("ms", C.c_float)]


FLAG_NO_THRESHOLD_SAMPLE = 1
FIU_GENO_READER_FLUSH = 1
FLAG_MFMA_INT8 = 2


class SearchStats(C.Structure):
    _fields_ = [("combinations", C.c_uint64), ("launches", C.c_uint64),
                ("redo_chunks", C.c_uint64), ("scan_ms", C.c_double),
                ("total_ms", C.c_double), ("algorithm", C.c_int32), ("variant", C.c_int32)]


class BatchConfig(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("max_active", C.c_uint32), ("memory_budget", C.c_uint64),
                ("step_ms", C.c_double), ("candidate_cap", C.c_uint32), ("flags", C.c_uint32)]


class Scan(C.Structure):
    _fields_ = [("dataset", C.c_void_p), ("order", C.c_uint16), ("reserved", C.c_uint16),
                ("dist_step", C.c_int32), ("dist_offset", C.c_int32), ("outputs", C.c_uint32),
                ("out_combos", C.c_void_p), ("out_mi", C.c_void_p), ("out_count", C.c_uint32),
                ("status", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("units", C.c_uint64), ("redo_ranges", C.c_uint64),
                ("device_ms", C.c_double), ("total_ms", C.c_double)]


class Request(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("order", C.c_uint16), ("reserved", C.c_uint16),
                ("dataset", C.c_void_p), ("words", C.c_void_p),
                ("snps", C.c_uint64), ("cases_words", C.c_uint64), ("ctrls_words", C.c_uint64),
                ("cases", C.c_uint32), ("ctrls", C.c_uint32),
                ("dist_step", C.c_int32), ("dist_offset", C.c_int32),
                ("outputs", C.c_uint32), ("reserved2", C.c_uint32)]


class RequestStats(C.Structure):
    _fields_ = [("submit_ms", C.c_double), ("admit_ms", C.c_double), ("start_ms", C.c_double),
                ("done_ms", C.c_double), ("latency_ms", C.c_double), ("combinations", C.c_uint64),
                ("steps", C.c_uint64), ("algorithm", C.c_int32), ("reserved", C.c_int32)]


class ServerInfo(C.Structure):
    _fields_ = [("queued", C.c_uint64), ("active", C.c_uint64), ("completed", C.c_uint64),
                ("memory_budget", C.c_uint64), ("memory_used", C.c_uint64), ("steps", C.c_uint64),
                ("device_ms", C.c_double)]


class SynthSpec(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("cases", C.c_uint32), ("ctrls", C.c_uint32),
                ("planted_order", C.c_uint32), ("snps", C.c_uint64),
                ("maf_lo", C.c_double), ("maf_hi", C.c_double),
                ("low_frac", C.c_double), ("low_lo", C.c_double), ("low_hi", C.c_double),
                ("seed", C.c_uint64), ("planted", C.c_int64 * 8),
                ("penetrance_seed", C.c_uint64), ("pen_lo", C.c_double), ("pen_hi", C.c_double),
                ("threads", C.c_int32), ("reserved", C.c_int32)]


_lib = None


def lib():
    """Load the library (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FiuError(FIU_ERR_DEVICE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, sz, u32, i32, u16 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_int32, C.c_uint16
    L.fiu_last_error.restype = C.c_char_p
    L.fiu_version.restype = C.c_char_p
    L.fiu_set_device.argtypes = [C.c_int]
    L.fiu_dataset_create.argtypes = [vp, sz, sz, sz, u32, u32, C.POINTER(vp)]
    L.fiu_dataset_read.argtypes = [C.POINTER(C.c_char_p), sz, C.POINTER(vp)]
    L.fiu_dataset_from_genotypes.argtypes = [vp, sz, sz, vp, u32, C.POINTER(vp)]
    L.fiu_read_files.argtypes = [C.POINTER(C.c_char_p), sz, vp, sz, C.POINTER(sz), C.POINTER(u32),
                                 C.POINTER(u32), C.POINTER(sz), C.POINTER(sz)]
    L.fiu_dataset_info.argtypes = [vp, C.POINTER(sz), C.POINTER(u32), C.POINTER(u32),
                                   C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]
    L.fiu_dataset_words.argtypes = [vp, vp]
    L.fiu_dataset_destroy.argtypes = [vp]
    L.fiu_config_init.argtypes = [C.POINTER(Config)]
    L.fiu_config_init.restype = None
    L.fiu_search.argtypes = [vp, u16, i32, i32, u32, C.POINTER(Config), vp, vp,
                             C.POINTER(u32), C.POINTER(SearchStats)]
    L.fiu_search_static.argtypes = L.fiu_search.argtypes
    L.fiu_search_static_bytes.restype = sz
    L.fiu_search_static_bytes.argtypes = []
    L.fiu_ct_size.restype = sz
    L.fiu_ct_size.argtypes = [u16]
    L.fiu_combine_batch.argtypes = [vp, vp, u16, sz, vp]
    L.fiu_contingency_batch.argtypes = [vp, vp, u16, sz, vp, vp]
    L.fiu_mi_batch.argtypes = [vp, sz, u16, u32, u32, vp]
    L.fiu_batch_config_init.argtypes = [C.POINTER(BatchConfig)]
    L.fiu_batch_config_init.restype = None
    L.fiu_search_batch.argtypes = [C.POINTER(Scan), sz, C.POINTER(BatchConfig), C.POINTER(BatchStats)]
    L.fiu_server_create.argtypes = [C.POINTER(BatchConfig), C.POINTER(vp)]
    L.fiu_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.fiu_server_wait.argtypes = [vp, C.c_uint64, C.c_double, vp, vp, C.POINTER(u32), C.POINTER(RequestStats)]
    L.fiu_server_query.argtypes = [vp, C.POINTER(ServerInfo)]
    L.fiu_server_destroy.argtypes = [vp]
    L.fiu_set_tuning.argtypes = [C.POINTER(TuneEntry), sz]
    L.fiu_get_tuning.argtypes = [C.POINTER(TuneEntry), sz, C.POINTER(sz)]
    L.fiu_synth_spec_init.argtypes = [C.POINTER(SynthSpec)]
    L.fiu_synth_spec_init.restype = None
    L.fiu_synth_generate.argtypes = [C.POINTER(SynthSpec), vp, vp, C.POINTER(sz), C.POINTER(sz)]
    L.fiu_dataset_synth.argtypes = [C.POINTER(SynthSpec), C.POINTER(vp)]
    L.fiu_result_str.restype = sz
    L.fiu_result_str.argtypes = [vp, u16, C.c_float, C.c_char_p, sz]
    L.fiu_result_serialize.restype = sz
    L.fiu_result_serialize.argtypes = [vp, u16, C.c_float, vp, sz]
    L.fiu_result_deserialize.argtypes = [vp, sz, vp, u16, C.POINTER(u16), C.POINTER(C.c_float), C.POINTER(sz)]
    L.fiu_write_results.argtypes = [C.c_char_p, vp, vp, u32, u16]
    _lib = L
    return L


def check(status):
    if status != FIU_OK:
        raise FiuError(status, lib().fiu_last_error().decode())
