"""ctypes binding of libkgpu.so (include/kgpu.h).

There is no fallback: if the shared library is missing or does not load,
every entry point raises.  Build it with ``python -m kgpu.build``.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# KGPU_LIB: another build of the same sources (the host-sanitized one)
LIB_PATH = os.environ.get("KGPU_LIB") or os.path.join(HERE, "libkgpu.so")

OK = 0
ERR_UNSUPPORTED, ERR_OVERFLOW, ERR_OUT_OF_RANGE, ERR_MEMORY, ERR_INTERNAL, ERR_DEVICE = 1, 2, 3, 4, 5, 6

T_BOOL, T_INT8, T_INT16, T_INT32, T_INT64 = 1, 2, 3, 4, 5
T_UINT8, T_UINT16, T_UINT32, T_UINT64 = 6, 7, 8, 9
T_FLOAT32, T_FLOAT64, T_DATE32, T_TIMESTAMP, T_DECIMAL64, T_DECIMAL128 = 10, 11, 12, 13, 14, 15
COL_DEVICE = 1

AGG_COUNT_STAR, AGG_COUNT, AGG_SUM, AGG_MIN, AGG_MAX, AGG_AVG = 1, 2, 3, 4, 5, 6
AGG_ANSI = 1
AGG_MERGE_AUTO, AGG_MERGE_WINDOWS, AGG_MERGE_BY_KEY = 0, 1, 2
JOIN_INNER, JOIN_LEFT_OUTER, JOIN_LEFT_SEMI, JOIN_LEFT_ANTI = 0, 1, 2, 3
JOIN_NULL_AWARE_ANTI, JOIN_EXISTENCE, JOIN_RIGHT_SEMI, JOIN_RIGHT_ANTI = 4, 5, 6, 7
JOIN_RIGHT_OUTER, JOIN_FULL_OUTER = 8, 9
# residual filter ops (postfix)
(X_PROBE, X_BUILD, X_CONST, X_ADD, X_SUB, X_MUL, X_MOD, X_EQ, X_NE, X_LT, X_LE, X_GT, X_GE,
 X_AND, X_OR, X_NOT, X_IS_NULL) = range(1, 18)

WIDTH = {T_BOOL: 0, T_INT8: 1, T_UINT8: 1, T_INT16: 2, T_UINT16: 2, T_INT32: 4, T_UINT32: 4,
         T_FLOAT32: 4, T_DATE32: 4, T_INT64: 8, T_UINT64: 8, T_FLOAT64: 8, T_TIMESTAMP: 8,
         T_DECIMAL64: 8, T_DECIMAL128: 16}


class KgError(RuntimeError):
    """A failed libkgpu call: ``code`` and the Spark error class in the text."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class Column(C.Structure):
    _fields_ = [("type", C.c_int32), ("precision", C.c_int32), ("scale", C.c_int32),
                ("flags", C.c_uint32), ("length", C.c_int64), ("offset", C.c_int64),
                ("values", C.c_void_p), ("validity", C.c_void_p)]


class FilterOptions(C.Structure):
    """kg_filter_options (KeyFilterOptions)."""
    _fields_ = [("max_values", C.c_uint64), ("exact_max_values", C.c_uint64),
                ("bloom", C.c_int32), ("side", C.c_int32), ("bloom_max_items", C.c_uint64),
                ("bloom_max_bytes", C.c_uint64), ("bloom_fpp", C.c_double),
                ("bloom_max_fpp", C.c_double), ("bloom_bytes", C.c_uint64)]


class KeyFilterInfo(C.Structure):
    _fields_ = [("form", C.c_int32), ("has_range", C.c_int32), ("members", C.c_int64),
                ("values", C.c_int64), ("lo", C.c_uint64 * 2), ("hi", C.c_uint64 * 2),
                ("flo", C.c_double), ("fhi", C.c_double), ("nulls_pass", C.c_int32),
                ("empty", C.c_int32), ("keys", C.c_uint64), ("bloom_bytes", C.c_uint64),
                ("row_cost_ns", C.c_double)]


class FrameHeader(C.Structure):
    """kg_frame_header (frame.rs FrameHeader)."""
    _fields_ = [("id", C.c_uint64), ("stage_attempt", C.c_uint32), ("map", C.c_uint32),
                ("attempt", C.c_uint32), ("sequence", C.c_uint32), ("reduce", C.c_uint32),
                ("kind", C.c_uint32)]


class ArrowSchema(C.Structure):
    pass


ArrowSchema._fields_ = [("format", C.c_char_p), ("name", C.c_char_p), ("metadata", C.c_char_p),
                        ("flags", C.c_int64), ("n_children", C.c_int64),
                        ("children", C.POINTER(C.POINTER(ArrowSchema))),
                        ("dictionary", C.POINTER(ArrowSchema)),
                        ("release", C.CFUNCTYPE(None, C.POINTER(ArrowSchema))),
                        ("private_data", C.c_void_p)]


class ArrowArray(C.Structure):
    pass


ArrowArray._fields_ = [("length", C.c_int64), ("null_count", C.c_int64), ("offset", C.c_int64),
                       ("n_buffers", C.c_int64), ("n_children", C.c_int64),
                       ("buffers", C.POINTER(C.c_void_p)),
                       ("children", C.POINTER(C.POINTER(ArrowArray))),
                       ("dictionary", C.POINTER(ArrowArray)),
                       ("release", C.CFUNCTYPE(None, C.POINTER(ArrowArray))),
                       ("private_data", C.c_void_p)]


class ArrowDeviceArray(C.Structure):
    _fields_ = [("array", ArrowArray), ("device_id", C.c_int64), ("device_type", C.c_int32),
                ("sync_event", C.c_void_p), ("reserved", C.c_int64 * 3)]


ARROW_DEVICE_CPU, ARROW_DEVICE_ROCM, ARROW_DEVICE_ROCM_HOST = 1, 10, 11


class GraceJoinStats(C.Structure):
    _fields_ = [("spilled", C.c_int32), ("max_depth", C.c_uint32),
                ("partitions_joined", C.c_uint64), ("repartitions", C.c_uint64),
                ("reversed", C.c_uint64), ("spill_bytes", C.c_uint64)]


class AggSpec(C.Structure):
    _fields_ = [("func", C.c_int32), ("input", C.c_int32)]


class Expr(C.Structure):
    """kg_expr: one op of a residual filter program."""
    _fields_ = [("op", C.c_int32), ("arg", C.c_int32), ("imm", C.c_int64)]


class JoinSpec(C.Structure):
    _fields_ = [("join_type", C.c_int32), ("num_keys", C.c_int32),
                ("probe_keys", C.c_int32 * 4), ("build_keys", C.c_int32 * 4),
                ("null_safe", C.c_uint8 * 4), ("no_direct_index", C.c_uint8)]


_lib = None

_SIGS = {
    # name: (restype, argtypes)
    "kg_last_error": (C.c_char_p, []),
    "kg_abi_version": (C.c_int, []),
    "kg_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.c_void_p]),
    "kg_ctx_destroy": (None, [C.c_void_p]),
    "kg_ctx_hbm_in_use": (C.c_uint64, [C.c_void_p]),
    "kg_dev_alloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "kg_dev_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "kg_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "kg_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "kg_batch_num_rows": (C.c_int64, [C.c_void_p]),
    "kg_batch_num_columns": (C.c_int32, [C.c_void_p]),
    "kg_batch_column": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "kg_batch_column_to_host": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "kg_batch_release": (None, [C.c_void_p]),
    "kg_hash_columns": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p,
                                  C.c_uint32]),
    "kg_partition": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_uint32,
                               C.c_void_p, C.c_uint32]),
    "kg_repartition": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_int32,
                                 C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "kg_agg_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p]),
    "kg_agg_sink": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "kg_agg_finalize": (C.c_int, [C.c_void_p]),
    "kg_agg_result_type": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "kg_agg_result": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_agg_set_bypass": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64]),
    "kg_agg_set_merge": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "kg_agg_partial_next": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_agg_partial_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_agg_destroy": (None, [C.c_void_p]),
    "kg_bloom_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "kg_bloom_bytes_for": (C.c_uint64, [C.c_uint64, C.c_double]),
    "kg_spark_bloom_bits_for": (C.c_uint64, [C.c_uint64, C.c_double]),
    "kg_spark_bloom_hashes_for": (C.c_int32, [C.c_uint64, C.c_uint64]),
    "kg_spark_bloom_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_void_p]),
    "kg_bloom_insert": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "kg_bloom_expected_fpp": (C.c_double, [C.c_void_p, C.c_uint64]),
    "kg_bloom_export": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_bloom_probe": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "kg_bloom_probe_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                         C.c_void_p]),
    "kg_bloom_destroy": (None, [C.c_void_p]),
    "kg_filter_default_options": (None, [C.c_void_p]),
    "kg_key_filter_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_key_filter_add": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "kg_key_filter_merge_builder": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_key_filter_finish": (C.c_int, [C.c_void_p]),
    "kg_key_filter_info_get": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_key_filter_members": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_key_filter_evaluate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                         C.c_void_p]),
    "kg_key_filter_encode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_key_filter_decode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "kg_key_filter_merge": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_key_filter_destroy": (None, [C.c_void_p]),
    "kg_row_stage_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]),
    "kg_row_stage_set": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]),
    "kg_row_stage_start_partition": (C.c_int, [C.c_void_p]),
    "kg_row_stage_apply": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p,
                                     C.c_void_p]),
    "kg_row_stage_stats": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "kg_row_stage_destroy": (None, [C.c_void_p]),
    "kg_column_from_arrow": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "kg_column_from_arrow_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_batch_export_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_batch_export_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_split_partitions": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p,
                                      C.c_uint32, C.c_void_p, C.c_void_p]),
    "kg_grace_join_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "kg_grace_join_set_filter": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "kg_grace_join_build_sink": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    "kg_grace_join_build_finalize": (C.c_int, [C.c_void_p]),
    "kg_grace_join_probe_push": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    "kg_grace_join_finish": (C.c_int, [C.c_void_p]),
    "kg_grace_join_next": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "kg_grace_join_stats_get": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_grace_join_destroy": (None, [C.c_void_p]),
    "kg_frame_end": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_crc32c": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "kg_exchange_frames": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                     C.c_void_p]),
    "kg_select_range": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64,
                                  C.c_void_p, C.c_void_p]),
    "kg_gather": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                            C.c_void_p]),
    "kg_join_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "kg_join_build_sink": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]),
    "kg_join_build_finalize": (C.c_int, [C.c_void_p]),
    "kg_join_build_rows": (C.c_int64, [C.c_void_p]),
    "kg_join_probe": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p]),
    "kg_join_probe_selected": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64,
                                         C.c_void_p, C.c_int64, C.c_void_p]),
    "kg_join_destroy": (None, [C.c_void_p]),
    "kg_join_set_filter": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "kg_join_local_create": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_join_local_destroy": (None, [C.c_void_p]),
    "kg_join_probe_push": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p,
                                     C.c_int64]),
    "kg_join_probe_next": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "kg_join_tail_next": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p,
                                    C.c_void_p]),
    "kg_joinagg_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p,
                                    C.c_uint32]),
    "kg_joinagg_probe": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p]),
    "kg_joinagg_result": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_gen_join_build": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_int64,
                                    C.c_void_p, C.c_void_p]),
    "kg_gen_join_probe": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_int64,
                                    C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_zipf_cdf": (C.c_int, [C.c_uint64, C.c_double, C.c_void_p]),
    "kg_zipf_cdf_device": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.c_void_p]),
    "kg_gen_agg": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_void_p,
                             C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p]),
    "kg_gen_zipf_keys": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int64, C.c_int64,
                                   C.c_void_p, C.c_uint64, C.c_void_p]),
    "kg_prof_enable": (None, [C.c_int]),
    "kg_prof_reset": (None, []),
    "kg_prof_get": (C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p]),
    "kg_ctx_mem_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_tune_get": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
    "kg_tune_set": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
    "kg_autotune_joinagg": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "kg_tune_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "kg_tune_load": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p]),
    "kg_calib_random_read": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_int64, C.c_int32,
                                       C.c_void_p]),
    "kg_calib_atomic_add": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int64, C.c_int32, C.c_void_p]),
}


def exported_symbols():
    """Every function name include/kgpu.h declares (parsed from the header)."""
    import re
    hdr = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "kgpu.h")
    text = open(hdr).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kg_[a-z0-9_]+)\s*\(", text)))


def lib():
    """The loaded library (raises OSError when it is missing: no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built (python -m kgpu.build)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue  # reported by the export test; calling it raises AttributeError
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != OK:
        raise KgError(rc, lib().kg_last_error().decode())


_SIGS.update({
    "kg_grace_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_void_p]),
    "kg_grace_build_sink": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "kg_grace_probe_sink": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "kg_grace_finish": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kg_grace_destroy": (None, [C.c_void_p]),
})


class ServeRequest(C.Structure):
    """kg_serve_request (include/kgpu.h)."""
    _fields_ = [("id", C.c_int64), ("build_keys", C.c_void_p), ("build_rows", C.c_int64),
                ("table", C.c_int32), ("_pad", C.c_int32), ("probe_keys", C.c_void_p),
                ("probe_valid", C.c_void_p), ("probe_pay", C.c_void_p),
                ("probe_rows", C.c_int64)]


_SIGS.update({
    "kg_server_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_int32, C.c_void_p]),
    "kg_server_add_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "kg_server_submit": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kg_server_progress": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
    "kg_server_take_result": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
    "kg_server_pending": (C.c_int64, [C.c_void_p]),
    "kg_server_destroy": (None, [C.c_void_p]),
})

_SIGS.update({
    "kg_join_partition": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_uint32,
                                    C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint32]),
})

FORM_RANGE, FORM_SMALL, FORM_EXACT, FORM_BLOOM, FORM_DENSE = 0, 1, 2, 3, 4
SIDE_BOTH, SIDE_UPPER, SIDE_LOWER = 0, 1, 2
FRAME_HEADER_BYTES = 44
FRAME_DATA, FRAME_END, FRAME_VALUE = 1, 2, 3
