"""ctypes binding of the CPU oracle (oracle/libkgo.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- as the checker of the HIP
path, never as the thing measured as the product or shipped.  Each wrapper
names the reference file:line the C function restates (see oracle/kgo.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# KGO_LIB: another build of the same sources (the sanitized one)
_LIB_PATH = os.environ.get("KGO_LIB") or os.path.join(_HERE, "libkgo.so")

T_BOOL, T_INT8, T_INT16, T_INT32, T_INT64 = 1, 2, 3, 4, 5
T_UINT8, T_UINT16, T_UINT32, T_UINT64 = 6, 7, 8, 9
T_FLOAT32, T_FLOAT64, T_DATE32, T_TIMESTAMP, T_DECIMAL64, T_DECIMAL128 = 10, 11, 12, 13, 14, 15

AGG_COUNT_STAR, AGG_COUNT, AGG_SUM, AGG_MIN, AGG_MAX, AGG_AVG = 1, 2, 3, 4, 5, 6
JOIN_INNER, JOIN_LEFT_OUTER, JOIN_LEFT_SEMI, JOIN_LEFT_ANTI = 0, 1, 2, 3
JOIN_NULL_AWARE_ANTI, JOIN_EXISTENCE, JOIN_RIGHT_SEMI, JOIN_RIGHT_ANTI = 4, 5, 6, 7
JOIN_RIGHT_OUTER, JOIN_FULL_OUTER = 8, 9
# residual filter program ops (postfix)
(X_PROBE, X_BUILD, X_CONST, X_ADD, X_SUB, X_MUL, X_MOD, X_EQ, X_NE, X_LT, X_LE, X_GT, X_GE,
 X_AND, X_OR, X_NOT, X_IS_NULL) = range(1, 18)

_NP_OF = {
    T_INT8: np.int8, T_INT16: np.int16, T_INT32: np.int32, T_INT64: np.int64,
    T_UINT8: np.uint8, T_UINT16: np.uint16, T_UINT32: np.uint32, T_UINT64: np.uint64,
    T_FLOAT32: np.float32, T_FLOAT64: np.float64, T_DATE32: np.int32, T_TIMESTAMP: np.int64,
    T_DECIMAL64: np.int64,
}


def build() -> str:
    """Compile the oracle with its Makefile (gcc) and return the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class _Col(C.Structure):
    _fields_ = [("type", C.c_int32), ("precision", C.c_int32), ("scale", C.c_int32),
                ("_pad", C.c_int32), ("length", C.c_int64), ("offset", C.c_int64),
                ("values", C.c_void_p), ("validity", C.c_void_p)]


class _Out(C.Structure):
    _fields_ = [("type", C.c_int32), ("precision", C.c_int32), ("scale", C.c_int32),
                ("_pad", C.c_int32), ("i64", C.c_void_p), ("hi", C.c_void_p),
                ("f64", C.c_void_p), ("valid", C.c_void_p)]


class _AggSpec(C.Structure):
    _fields_ = [("func", C.c_int32), ("input", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        u64, u32, i64, vp = C.c_uint64, C.c_uint32, C.c_int64, C.c_void_p
        L.kgo_xxh64.restype = u64
        L.kgo_xxh64.argtypes = [vp, C.c_size_t, u64]
        L.kgo_xxh64_fixed.restype = u64
        L.kgo_xxh64_fixed.argtypes = [u64, u64, C.c_int]
        L.kgo_fold_column.restype = u64
        L.kgo_fold_column.argtypes = [u64, u64]
        L.kgo_fold.restype = u64
        L.kgo_fold.argtypes = [u64, u64]
        L.kgo_key_group.restype = u32
        L.kgo_key_group.argtypes = [u64]
        L.kgo_radix.restype = u32
        L.kgo_radix.argtypes = [u64, u32]
        L.kgo_partition_of.restype = u32
        L.kgo_partition_of.argtypes = [u64, u32]
        L.kgo_key_group_range.argtypes = [u32, u32, vp, vp]
        L.kgo_hash_columns.restype = C.c_int
        L.kgo_hash_columns.argtypes = [C.c_int, vp, i64, vp]
        L.kgo_route.restype = C.c_int
        L.kgo_route.argtypes = [C.c_int, vp, i64, u32, vp]
        L.kgo_split.argtypes = [vp, i64, u32, vp, vp]
        L.kgo_row_hash.restype = C.c_int
        L.kgo_row_hash.argtypes = [C.c_int, vp, i64, vp]
        L.kgo_finalize.restype = u64
        L.kgo_finalize.argtypes = [u64, u64]
        L.kgo_join_partition.restype = u32
        L.kgo_join_partition.argtypes = [u64, u64, u32, u32]
        L.kgo_sbbf_bytes_for.restype = C.c_size_t
        L.kgo_sbbf_bytes_for.argtypes = [u64, C.c_double]
        L.kgo_sbbf_blocks_for_bytes.restype = C.c_size_t
        L.kgo_sbbf_blocks_for_bytes.argtypes = [C.c_size_t]
        L.kgo_sbbf_insert.argtypes = [vp, C.c_size_t, u64]
        L.kgo_sbbf_check.restype = C.c_int
        L.kgo_sbbf_check.argtypes = [vp, C.c_size_t, u64]
        L.kgo_murmur3_int.restype = C.c_int32
        L.kgo_murmur3_int.argtypes = [C.c_int32, C.c_int32]
        L.kgo_murmur3_long.restype = C.c_int32
        L.kgo_murmur3_long.argtypes = [i64, C.c_int32]
        L.kgo_spark_xxhash64.restype = i64
        L.kgo_spark_xxhash64.argtypes = [i64, C.c_int]
        L.kgo_spark_xxhash64_batch.restype = None
        L.kgo_spark_xxhash64_batch.argtypes = [vp, i64, C.c_int, vp]
        L.kgo_sbbf_check_batch.restype = None
        L.kgo_sbbf_check_batch.argtypes = [vp, C.c_size_t, vp, C.c_int64, vp]
        L.kgo_sbbf_expected_fpp.restype = C.c_double
        L.kgo_sbbf_expected_fpp.argtypes = [C.c_size_t, u64]
        L.kgo_sbbf_build.restype = C.c_int
        L.kgo_sbbf_build.argtypes = [vp, i64, C.c_size_t, C.c_double, C.c_size_t, C.c_double,
                                     vp, vp]
        L.kgo_mix64.restype = u64
        L.kgo_mix64.argtypes = [u64]
        L.kgo_rand.restype = u64
        L.kgo_rand.argtypes = [u64, u64, u64]
        L.kgo_below.restype = u64
        L.kgo_below.argtypes = [u64, u64]
        L.kgo_unit.restype = C.c_double
        L.kgo_unit.argtypes = [u64]
        L.kgo_perm.restype = u64
        L.kgo_perm.argtypes = [u64, u64, u64]
        L.kgo_zipf_cdf.argtypes = [u64, C.c_double, vp]
        L.kgo_zipf_rank.restype = u64
        L.kgo_zipf_rank.argtypes = [vp, u64, C.c_double]
        L.kgo_gen_join_build.argtypes = [u64, i64, i64, i64, vp, vp]
        L.kgo_gen_join_probe.argtypes = [u64, i64, i64, i64, u32, vp, vp, vp]
        L.kgo_gen_agg.argtypes = [u64, i64, i64, vp, u64, vp, vp, vp, vp, vp]
        L.kgo_gen_zipf_keys.argtypes = [u64, u64, i64, i64, vp, u64, vp]
        L.kgo_agg_new.restype = vp
        L.kgo_agg_new.argtypes = [vp, C.c_int, vp, C.c_int, vp, C.c_int]
        L.kgo_agg_update.restype = C.c_int
        L.kgo_agg_update.argtypes = [vp, vp, vp, i64]
        L.kgo_agg_merge.restype = C.c_int
        L.kgo_agg_merge.argtypes = [vp, vp]
        L.kgo_agg_groups.restype = i64
        L.kgo_agg_groups.argtypes = [vp]
        L.kgo_agg_finish.restype = C.c_int
        L.kgo_agg_finish.argtypes = [vp, vp, vp, vp]
        L.kgo_agg_result_type.argtypes = [vp, C.c_int, vp, vp, vp]
        L.kgo_agg_free.argtypes = [vp]
        L.kgo_last_error.restype = C.c_char_p
        L.kgo_avg_fast.restype = C.c_int
        L.kgo_avg_fast.argtypes = [i64, i64, C.c_int, i64, C.c_int, C.c_int, vp, vp]
        L.kgo_join.restype = i64
        L.kgo_join.argtypes = [C.c_int, C.c_int, vp, vp, vp, i64, i64, C.c_int, i64, vp, vp]
        L.kgo_join_ex.restype = i64
        L.kgo_join_ex.argtypes = [C.c_int, C.c_int, vp, vp, vp, i64, i64, C.c_int, C.c_int, vp,
                                  vp, vp, i64, vp, vp]
        L.kgo_agg_set_exact_f64.argtypes = [vp, C.c_int]
        L.kgo_pagg_new.restype = vp
        L.kgo_pagg_new.argtypes = [vp, C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, C.c_int]
        L.kgo_pagg_update.restype = C.c_int
        L.kgo_pagg_update.argtypes = [vp, vp, vp, i64]
        L.kgo_pagg_finalize.restype = C.c_int
        L.kgo_pagg_finalize.argtypes = [vp]
        L.kgo_pagg_groups.restype = i64
        L.kgo_pagg_groups.argtypes = [vp]
        L.kgo_pagg_finish.restype = C.c_int
        L.kgo_pagg_finish.argtypes = [vp, vp, vp, vp]
        L.kgo_pagg_result_type.argtypes = [vp, C.c_int, vp, vp, vp]
        L.kgo_pagg_free.argtypes = [vp]
        L.kgo_grace_factorized.restype = C.c_int
        L.kgo_grace_factorized.argtypes = [vp, vp, i64, vp, vp, i64, u32, C.c_int, vp, vp, vp]
        f64 = C.c_double
        L.kgo_filter_default_opts.restype = None
        L.kgo_filter_default_opts.argtypes = [vp]
        L.kgo_kf_widens.restype = C.c_int
        L.kgo_kf_widens.argtypes = [C.c_int] * 6
        L.kgo_kf_new.restype = vp
        L.kgo_kf_new.argtypes = [C.c_int, C.c_int, C.c_int, vp]
        L.kgo_kf_free.restype = None
        L.kgo_kf_free.argtypes = [vp]
        L.kgo_kf_add.restype = C.c_int
        L.kgo_kf_add.argtypes = [vp, vp, vp, i64]
        L.kgo_kf_merge_builder.restype = C.c_int
        L.kgo_kf_merge_builder.argtypes = [vp, vp]
        L.kgo_kf_finish.restype = C.c_int
        L.kgo_kf_finish.argtypes = [vp]
        L.kgo_kf_get_info.restype = C.c_int
        L.kgo_kf_get_info.argtypes = [vp, vp]
        L.kgo_kf_members.restype = i64
        L.kgo_kf_members.argtypes = [vp, vp]
        L.kgo_kf_bloom_words.restype = i64
        L.kgo_kf_bloom_words.argtypes = [vp, vp]
        L.kgo_kf_row_cost.restype = f64
        L.kgo_kf_row_cost.argtypes = [vp]
        L.kgo_kf_evaluate.restype = i64
        L.kgo_kf_evaluate.argtypes = [vp, vp, i64, vp]
        L.kgo_kf_encode.restype = i64
        L.kgo_kf_encode.argtypes = [vp, vp, i64]
        L.kgo_kf_decode.restype = vp
        L.kgo_kf_decode.argtypes = [vp, i64, vp]
        L.kgo_kf_merge.restype = vp
        L.kgo_kf_merge.argtypes = [vp, vp, vp, vp]
        L.kgo_sbbf_pass_rate.restype = f64
        L.kgo_sbbf_pass_rate.argtypes = [vp, C.c_size_t]
        L.kgo_rs_new.restype = vp
        L.kgo_rs_new.argtypes = [u64, f64]
        L.kgo_rs_free.restype = None
        L.kgo_rs_free.argtypes = [vp]
        L.kgo_rs_set.restype = C.c_int
        L.kgo_rs_set.argtypes = [vp, C.c_int, vp, C.c_int]
        L.kgo_rs_start_partition.restype = None
        L.kgo_rs_start_partition.argtypes = [vp]
        L.kgo_rs_stats.restype = None
        L.kgo_rs_stats.argtypes = [vp, C.c_int, vp, vp, vp]
        L.kgo_rs_apply.restype = i64
        L.kgo_rs_apply.argtypes = [vp, C.c_int, vp, i64, vp]
        L.kgo_crc32c.restype = u32
        L.kgo_crc32c.argtypes = [vp, u64]
        L.kgo_frame_header.restype = None
        L.kgo_frame_header.argtypes = [vp, vp, u64, vp]
        L.kgo_frame_decode.restype = C.c_int
        L.kgo_frame_decode.argtypes = [vp, u64, vp, vp, vp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class Col:
    """A host column: numpy values (+ optional bool validity), Arrow layout."""

    def __init__(self, ctype, values, valid=None, precision=0, scale=0):
        self.type = ctype
        self.precision = precision
        self.scale = scale
        if ctype == T_BOOL:
            v = np.asarray(values, dtype=bool)
            self.length = len(v)
            self.values = np.packbits(v, bitorder="little")
        elif ctype == T_DECIMAL128:
            v = np.ascontiguousarray(values)
            if v.dtype != np.uint64 or v.ndim != 2:
                v = dec128_words([int(x) for x in values])
            self.length = v.shape[0]
            self.values = v
        else:
            self.values = np.ascontiguousarray(values, dtype=_NP_OF[ctype])
            self.length = len(self.values)
        self.valid = None if valid is None else np.asarray(valid, dtype=bool)
        self.bitmap = (None if self.valid is None
                       else np.packbits(self.valid, bitorder="little"))

    def c(self) -> _Col:
        return _Col(self.type, self.precision, self.scale, 0, self.length, 0,
                    _ptr(self.values), _ptr(self.bitmap))


def dec128_words(ints) -> np.ndarray:
    out = np.zeros((len(ints), 2), dtype=np.uint64)
    for i, x in enumerate(ints):
        u = x & ((1 << 128) - 1)
        out[i, 0] = u & 0xFFFFFFFFFFFFFFFF
        out[i, 1] = u >> 64
    return out


def _cols(cols):
    arr = (_Col * max(1, len(cols)))()
    for i, c in enumerate(cols):
        arr[i] = c.c()
    return arr


def error() -> str:
    return lib().kgo_last_error().decode()


# --- hashing -------------------------------------------------------------
def xxh64(data: bytes, seed: int = 0) -> int:
    return lib().kgo_xxh64(data, len(data), seed)


def xxh64_fixed(value: int, width: int) -> int:
    u = value & ((1 << 128) - 1)
    return lib().kgo_xxh64_fixed(u & 0xFFFFFFFFFFFFFFFF, u >> 64, width)


def hash_columns(cols) -> np.ndarray:
    n = cols[0].length
    out = np.zeros(n, dtype=np.uint64)
    ca = _cols(cols)
    if lib().kgo_hash_columns(len(cols), C.addressof(ca), n, _ptr(out)) != 0:
        raise RuntimeError(error())
    return out


def route(cols, partitions: int) -> np.ndarray:
    n = cols[0].length if cols else 0
    out = np.zeros(n, dtype=np.uint32)
    ca = _cols(cols)
    if lib().kgo_route(len(cols), C.addressof(ca), n, partitions, _ptr(out)) != 0:
        raise RuntimeError(error())
    return out


def split(pids: np.ndarray, partitions: int):
    pids = np.ascontiguousarray(pids, dtype=np.uint32)
    offs = np.zeros(partitions + 1, dtype=np.int64)
    rows = np.zeros(len(pids), dtype=np.int64)
    lib().kgo_split(_ptr(pids), len(pids), partitions, _ptr(offs), _ptr(rows))
    return offs, rows


def row_hash(cols) -> np.ndarray:
    n = cols[0].length
    out = np.zeros(n, dtype=np.uint64)
    ca = _cols(cols)
    lib().kgo_row_hash(len(cols), C.addressof(ca), n, _ptr(out))
    return out


def key_group_range(p: int, n: int):
    s, e = C.c_uint32(), C.c_uint32()
    lib().kgo_key_group_range(p, n, C.byref(s), C.byref(e))
    return s.value, e.value


# --- SBBF ----------------------------------------------------------------
def sbbf_build(hashes: np.ndarray, nbytes: int = 0, fpp: float = 0.01,
               max_bytes: int = 8 << 20, max_fpp: float = 0.05):
    hashes = np.ascontiguousarray(hashes, dtype=np.uint64)
    b = nbytes or lib().kgo_sbbf_bytes_for(len(hashes), fpp)
    nb = lib().kgo_sbbf_blocks_for_bytes(min(b, max_bytes))
    words = np.zeros(max(nb, 1) * 8, dtype=np.uint32)
    nbo = C.c_size_t()
    kept = lib().kgo_sbbf_build(_ptr(hashes), len(hashes), nbytes, fpp, max_bytes, max_fpp,
                                _ptr(words), C.byref(nbo))
    return bool(kept), words[: nbo.value * 8]


def sbbf_check(words: np.ndarray, h: int) -> bool:
    words = np.ascontiguousarray(words, dtype=np.uint32)
    return bool(lib().kgo_sbbf_check(_ptr(words), len(words) // 8, h))


def sbbf_check_batch(words: np.ndarray, hashes: np.ndarray) -> np.ndarray:
    words = np.ascontiguousarray(words, dtype=np.uint32)
    hashes = np.ascontiguousarray(hashes, dtype=np.uint64)
    out = np.zeros(len(hashes), dtype=np.uint8)
    lib().kgo_sbbf_check_batch(_ptr(words), len(words) // 8, _ptr(hashes), len(hashes), _ptr(out))
    return out.astype(bool)


def spark_xxhash64(values: np.ndarray, width: int) -> np.ndarray:
    """Spark's xxhash64 (seed 42) of int (width 4) or bigint (width 8) values."""
    v = np.ascontiguousarray(np.asarray(values).astype(np.int64))
    out = np.zeros(len(v), dtype=np.int64)
    lib().kgo_spark_xxhash64_batch(_ptr(v), len(v), width, _ptr(out))
    return out


# --- generators ----------------------------------------------------------
def zipf_cdf(n: int, s: float) -> np.ndarray:
    cdf = np.zeros(n, dtype=np.float64)
    lib().kgo_zipf_cdf(n, s, _ptr(cdf))
    return cdf


def gen_join_build(seed, lo, hi, n_total, with_val=True):
    key = np.zeros(hi - lo, dtype=np.int64)
    val = np.zeros(hi - lo, dtype=np.int64) if with_val else None
    lib().kgo_gen_join_build(seed, lo, hi, n_total, _ptr(key), _ptr(val))
    return key, val


def gen_join_probe(seed, lo, hi, key_range, null_permille=10):
    fk = np.zeros(hi - lo, dtype=np.int64)
    ok = np.zeros(hi - lo, dtype=np.uint8)
    amt = np.zeros(hi - lo, dtype=np.int64)
    lib().kgo_gen_join_probe(seed, lo, hi, key_range, null_permille, _ptr(fk), _ptr(ok), _ptr(amt))
    return fk, ok.astype(bool), amt


def gen_zipf_keys(seed, key_seed, lo, hi, cdf, ndistinct):
    key = np.zeros(hi - lo, dtype=np.int64)
    lib().kgo_gen_zipf_keys(seed, key_seed, lo, hi, _ptr(cdf), ndistinct, _ptr(key))
    return key


def gen_agg(seed, lo, hi, cdf, ndistinct):
    n = hi - lo
    key = np.zeros(n, dtype=np.int64)
    dec = np.zeros(n, dtype=np.int64)
    f = np.zeros(n, dtype=np.float64)
    i32 = np.zeros(n, dtype=np.int32)
    k8 = np.zeros(n, dtype=np.int8)
    lib().kgo_gen_agg(seed, lo, hi, _ptr(cdf), ndistinct, _ptr(key), _ptr(dec), _ptr(f),
                      _ptr(i32), _ptr(k8))
    return key, dec, f, i32, k8


# --- aggregation ---------------------------------------------------------
class Agg:
    """The oracle's hash aggregate (Sink::sink = update, combine = merge,
    finalize = finish)."""

    def __init__(self, key_type, input_types, aggs, ansi=False, exact_f64=False):
        self.key_type = key_type
        kc = _Col(key_type[0], key_type[1], key_type[2], 0, 0, 0, None, None) if key_type else None
        self._kc = kc
        its = (_Col * max(1, len(input_types)))()
        for i, (t, p, s) in enumerate(input_types):
            its[i] = _Col(t, p, s, 0, 0, 0, None, None)
        specs = (_AggSpec * max(1, len(aggs)))()
        for i, (f, inp) in enumerate(aggs):
            specs[i] = _AggSpec(f, inp)
        self.naggs = len(aggs)
        self.h = lib().kgo_agg_new(C.addressof(kc) if kc is not None else None, len(input_types),
                                   C.addressof(its), len(aggs), C.addressof(specs), int(ansi))
        if not self.h:
            raise RuntimeError(error())
        if exact_f64:
            lib().kgo_agg_set_exact_f64(self.h, 1)

    def update(self, key: Col | None, inputs):
        n = key.length if key is not None else (inputs[0].length if inputs else 0)
        kc = key.c() if key is not None else None
        ca = _cols(inputs)
        self._keep = (key, inputs)
        rc = lib().kgo_agg_update(self.h, C.addressof(kc) if kc is not None else None,
                                  C.addressof(ca), n)
        if rc != 0:
            raise ArithmeticError(error())

    def merge(self, other: "Agg"):
        if lib().kgo_agg_merge(self.h, other.h) != 0:
            raise ArithmeticError(error())

    def finish(self, raw=False):
        """Returns (keys, key_valid, [(values, valid, (type, p, s))...]).
        Decimal results are Python ints, or with raw=True an (n, 2) int64
        array of (low, high) words."""
        return _finish(lib().kgo_agg_groups(self.h), self.naggs,
                       lambda i, t, p, s: lib().kgo_agg_result_type(self.h, i, t, p, s),
                       lambda k, kv, outs: lib().kgo_agg_finish(self.h, k, kv, outs), raw)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().kgo_agg_free(self.h)
                self.h = None
        except Exception:
            pass


def _finish(g, naggs, result_type, finish, raw):
    """Shared by Agg and AggPar: allocate the output columns, finish, decode."""
    if True:
        L = lib()
        keys = np.zeros(max(g, 1), dtype=np.int64)
        kv = np.zeros(max(g, 1), dtype=np.uint8)
        outs = (_Out * max(1, naggs))()
        bufs = []
        for i in range(naggs):
            t, p, s = C.c_int32(), C.c_int32(), C.c_int32()
            result_type(i, C.byref(t), C.byref(p), C.byref(s))
            lo = np.zeros(max(g, 1), dtype=np.int64)
            hi = np.zeros(max(g, 1), dtype=np.int64)
            f = np.zeros(max(g, 1), dtype=np.float64)
            v = np.zeros(max(g, 1), dtype=np.uint8)
            outs[i] = _Out(t.value, p.value, s.value, 0, _ptr(lo), _ptr(hi), _ptr(f), _ptr(v))
            bufs.append((lo, hi, f, v, (t.value, p.value, s.value)))
        if finish(_ptr(keys), _ptr(kv), C.addressof(outs)) != 0:
            raise ArithmeticError(error())
        res = []
        for lo, hi, f, v, (t, p, s) in bufs:
            if t in (T_FLOAT64, T_FLOAT32):  # float32 MIN/MAX come back widened
                vals = f[:g].copy()
            elif t == T_DECIMAL64 and raw:
                vals = lo[:g].copy()
            elif t == T_DECIMAL128 and raw:
                vals = np.stack([lo[:g], hi[:g]], axis=1)
            elif t in (T_DECIMAL64, T_DECIMAL128):
                vals = [((int(h) << 64) | (int(l) & 0xFFFFFFFFFFFFFFFF)) for l, h in zip(lo[:g], hi[:g])]
                vals = [x - (1 << 128) if x >= (1 << 127) else x for x in vals]
            else:
                vals = lo[:g].copy()
            res.append((vals, v[:g].astype(bool), (t, p, s)))
        return keys[:g].copy(), kv[:g].astype(bool), res


class AggPar:
    """The aggregate on `threads` threads as the reference's pool runs it:
    thread-local tables in 16 radix partitions, then a parallel merge per
    partition (kgo_pagg_*).  exact_f64: float sums compensated (fsum-grade)."""

    def __init__(self, key_type, input_types, aggs, threads, ansi=False, exact_f64=False):
        kc = _Col(key_type[0], key_type[1], key_type[2], 0, 0, 0, None, None) if key_type else None
        its = (_Col * max(1, len(input_types)))()
        for i, (t, p, s) in enumerate(input_types):
            its[i] = _Col(t, p, s, 0, 0, 0, None, None)
        specs = (_AggSpec * max(1, len(aggs)))()
        for i, (f, inp) in enumerate(aggs):
            specs[i] = _AggSpec(f, inp)
        self.naggs = len(aggs)
        self.h = lib().kgo_pagg_new(C.addressof(kc) if kc is not None else None,
                                    len(input_types), C.addressof(its), len(aggs),
                                    C.addressof(specs), int(ansi), int(exact_f64), int(threads))
        if not self.h:
            raise RuntimeError(error())

    def update(self, key: Col | None, inputs):
        n = key.length if key is not None else (inputs[0].length if inputs else 0)
        kc = key.c() if key is not None else None
        ca = _cols(inputs)
        if lib().kgo_pagg_update(self.h, C.addressof(kc) if kc is not None else None,
                                 C.addressof(ca), n) != 0:
            raise ArithmeticError(error())

    def finalize(self):
        if lib().kgo_pagg_finalize(self.h) != 0:
            raise ArithmeticError(error())

    def finish(self, raw=False):
        self.finalize()
        return _finish(lib().kgo_pagg_groups(self.h), self.naggs,
                       lambda i, t, p, s: lib().kgo_pagg_result_type(self.h, i, t, p, s),
                       lambda k, kv, outs: lib().kgo_pagg_finish(self.h, k, kv, outs), raw)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().kgo_pagg_free(self.h)
                self.h = None
        except Exception:
            pass


class _Expr(C.Structure):
    _fields_ = [("op", C.c_int32), ("arg", C.c_int32), ("imm", C.c_int64)]


def join_ex(join_type, probe_keys, build_keys, null_safe=None, nested=False, filt=None,
            probe_cols=None, build_cols=None, cap=None):
    """All ten join types with an optional residual filter: `filt` is a list
    of (op, arg, imm) postfix ops over probe_cols/build_cols.  Returns
    (probe rows, build rows) as kgo_join_ex emits them."""
    n_p = probe_keys[0].length
    n_b = build_keys[0].length
    if cap is None:
        cap = (n_p + 1) * (n_b + 1) if nested else max(16, 4 * (n_p + n_b))
    op = np.zeros(cap, dtype=np.int64)
    ob = np.zeros(cap, dtype=np.int64)
    pk, bk = _cols(probe_keys), _cols(build_keys)
    ns = None if null_safe is None else np.asarray(null_safe, dtype=np.uint8)
    nf = len(filt) if filt else 0
    fx = (_Expr * max(1, nf))(*[_Expr(*x) for x in (filt or [])])
    pc = _cols(probe_cols or [])
    bc = _cols(build_cols or [])
    n = lib().kgo_join_ex(join_type, len(probe_keys), C.addressof(pk), C.addressof(bk), _ptr(ns),
                          n_p, n_b, int(nested), nf, C.addressof(fx), C.addressof(pc),
                          C.addressof(bc), cap, _ptr(op), _ptr(ob))
    if n < 0:
        return join_ex(join_type, probe_keys, build_keys, null_safe, nested, filt, probe_cols,
                       build_cols, cap * 4)
    return op[:n].copy(), ob[:n].copy()


def join(join_type, probe_keys, build_keys, null_safe=None, nested=False, cap=None):
    n_p = probe_keys[0].length
    n_b = build_keys[0].length
    if cap is None:
        cap = n_p * max(n_b, 1) if nested else max(16, 4 * (n_p + n_b))
    op = np.zeros(cap, dtype=np.int64)
    ob = np.zeros(cap, dtype=np.int64)
    pk, bk = _cols(probe_keys), _cols(build_keys)
    ns = None if null_safe is None else np.asarray(null_safe, dtype=np.uint8)
    n = lib().kgo_join(join_type, len(probe_keys), C.addressof(pk), C.addressof(bk), _ptr(ns),
                       n_p, n_b, int(nested), cap, _ptr(op), _ptr(ob))
    if n < 0:
        return join(join_type, probe_keys, build_keys, null_safe, nested, cap * 4)
    return op[:n].copy(), ob[:n].copy()


def grace_factorized(bkey: Col, bpay: Col, pkey: Col, ppay: Col, parts: int, mode: int = 0):
    """Configuration 4's per-partition (COUNT, SUM, sum valid) factorized per
    key (kgo_grace_factorized); mode 1 takes the sum mod 2^31."""
    counts = np.zeros(parts, dtype=np.int64)
    sums = np.zeros(parts, dtype=np.int64)
    valid = np.zeros(parts, dtype=np.uint8)
    bk, bp, pk, pp = bkey.c(), bpay.c(), pkey.c(), ppay.c()
    if lib().kgo_grace_factorized(C.byref(bk), C.byref(bp), bkey.length, C.byref(pk),
                                  C.byref(pp), pkey.length, parts, mode, _ptr(counts),
                                  _ptr(sums), _ptr(valid)) != 0:
        raise RuntimeError(error())
    return counts, sums, valid.astype(bool)


# --- runtime KeyFilter (kgo_filter.c) ------------------------------------
SIDE_BOTH, SIDE_UPPER, SIDE_LOWER = 0, 1, 2
FORM_RANGE, FORM_SMALL, FORM_EXACT, FORM_BLOOM, FORM_DENSE = 0, 1, 2, 3, 4


class FilterOpts(C.Structure):
    """KeyFilterOptions (runtime_filter/filter.rs:28-107)."""
    _fields_ = [("max_values", C.c_uint64), ("exact_max_values", C.c_uint64),
                ("bloom", C.c_int32), ("side", C.c_int32), ("bloom_max_items", C.c_uint64),
                ("bloom_max_bytes", C.c_uint64), ("bloom_fpp", C.c_double),
                ("bloom_max_fpp", C.c_double), ("bloom_bytes", C.c_uint64)]

    @classmethod
    def default(cls, **kw):
        o = cls()
        lib().kgo_filter_default_opts(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class _KfInfo(C.Structure):
    _fields_ = [("form", C.c_int32), ("has_range", C.c_int32), ("nmembers", C.c_int64),
                ("values", C.c_int64), ("lo_lo", C.c_uint64), ("lo_hi", C.c_uint64),
                ("hi_lo", C.c_uint64), ("hi_hi", C.c_uint64), ("flo", C.c_double),
                ("fhi", C.c_double), ("nulls_pass", C.c_int32), ("empty", C.c_int32),
                ("keys", C.c_uint64), ("bloom_bytes", C.c_uint64)]


def _i128(lo, hi):
    v = (int(hi) << 64) | int(lo)
    return v - (1 << 128) if v >= 1 << 127 else v


class KeyFilter:
    """A KeyFilterBuilder until finish(), then the KeyFilter it published
    (runtime_filter/filter.rs:180-615)."""

    def __init__(self, ctype, precision=0, scale=0, opts=None, _h=None):
        self.opts = opts if opts is not None else FilterOpts.default()
        self.h = _h if _h is not None else lib().kgo_kf_new(ctype, precision, scale,
                                                           C.byref(self.opts))
        if not self.h:
            raise ValueError("unsupported runtime filter key type")

    def add(self, col: Col, include=None):
        inc = None if include is None else np.packbits(np.asarray(include, bool), bitorder="little")
        c = col.c()
        if lib().kgo_kf_add(self.h, C.byref(c), _ptr(inc), col.length) != 0:
            raise ValueError("keys of another type")
        return self

    def merge_builder(self, other: "KeyFilter"):
        if lib().kgo_kf_merge_builder(self.h, other.h) != 0:
            raise ValueError("builder merge refused")
        return self

    def finish(self):
        lib().kgo_kf_finish(self.h)
        return self

    def info(self) -> dict:
        i = _KfInfo()
        lib().kgo_kf_get_info(self.h, C.byref(i))
        return {"form": i.form, "has_range": bool(i.has_range), "nmembers": i.nmembers,
                "values": i.values, "lo": _i128(i.lo_lo, i.lo_hi), "hi": _i128(i.hi_lo, i.hi_hi),
                "flo": i.flo, "fhi": i.fhi, "nulls_pass": bool(i.nulls_pass),
                "empty": bool(i.empty), "keys": i.keys, "bloom_bytes": i.bloom_bytes}

    def members(self):
        n = lib().kgo_kf_members(self.h, None)
        w = np.zeros(2 * max(n, 1), dtype=np.uint64)
        lib().kgo_kf_members(self.h, _ptr(w))
        return [_i128(w[2 * k], w[2 * k + 1]) for k in range(n)]

    def bloom_words(self) -> np.ndarray:
        n = lib().kgo_kf_bloom_words(self.h, None)
        w = np.zeros(max(n, 1), dtype=np.uint32)
        lib().kgo_kf_bloom_words(self.h, _ptr(w))
        return w[:n]

    def row_cost(self) -> float:
        return lib().kgo_kf_row_cost(self.h)

    def evaluate(self, col: Col):
        """Per-row pass flags, or None when the column does not widen."""
        out = np.zeros(max(col.length, 1), dtype=np.uint8)
        c = col.c()
        n = lib().kgo_kf_evaluate(self.h, C.byref(c), col.length, _ptr(out))
        return None if n < 0 else out[:col.length].astype(bool)

    def encode(self) -> bytes:
        n = lib().kgo_kf_encode(self.h, None, 0)
        buf = np.zeros(max(n, 1), dtype=np.uint8)
        lib().kgo_kf_encode(self.h, _ptr(buf), n)
        return buf[:n].tobytes()

    @classmethod
    def decode(cls, data: bytes) -> "KeyFilter":
        err = C.create_string_buffer(128)
        buf = np.frombuffer(data, dtype=np.uint8) if data else np.zeros(1, np.uint8)
        h = lib().kgo_kf_decode(_ptr(buf), len(data), err)
        if not h:
            raise ValueError(err.value.decode())
        return cls(0, _h=h)

    def merge(self, other: "KeyFilter", opts=None) -> "KeyFilter":
        o = opts if opts is not None else self.opts
        err = C.create_string_buffer(128)
        h = lib().kgo_kf_merge(self.h, other.h, C.byref(o), err)
        if not h:
            raise ValueError(err.value.decode())
        return KeyFilter(0, opts=o, _h=h)

    def __del__(self):
        try:
            if self.h:
                lib().kgo_kf_free(self.h)
                self.h = None
        except Exception:
            pass


def sbbf_pass_rate(words: np.ndarray) -> float:
    w = np.ascontiguousarray(words, dtype=np.uint32)
    return lib().kgo_sbbf_pass_rate(_ptr(w), len(w) // 8)


class RowStage:
    """A scan driver's row stage (operators/runtime_filter.rs:240-395)."""

    def __init__(self, min_rows=65_536, max_pass_rate=0.9):
        self.h = lib().kgo_rs_new(min_rows, max_pass_rate)
        self.keep = {}

    def set(self, slot, kf: KeyFilter | None, column):
        self.keep[slot] = kf
        lib().kgo_rs_set(self.h, slot, kf.h if kf is not None else None, column)

    def start_partition(self):
        lib().kgo_rs_start_partition(self.h)

    def apply(self, cols):
        """Pass flags of the batch's rows, or None when nothing is filtered."""
        n = cols[0].length
        out = np.zeros(max(n, 1), dtype=np.uint8)
        arr = _cols(cols)
        r = lib().kgo_rs_apply(self.h, len(cols), arr, n, _ptr(out))
        if r == -2:
            raise ValueError("a runtime filter's column is missing")
        return None if r < 0 else out[:n].astype(bool)

    def stats(self, slot):
        t, d, o = C.c_uint64(), C.c_uint64(), C.c_int()
        lib().kgo_rs_stats(self.h, slot, C.byref(t), C.byref(d), C.byref(o))
        return t.value, d.value, bool(o.value)

    def __del__(self):
        try:
            if self.h:
                lib().kgo_rs_free(self.h)
                self.h = None
        except Exception:
            pass


# --- exchange frames (kgo_frame.c) -----------------------------------------
FRAME_DATA, FRAME_END, FRAME_VALUE = 1, 2, 3
FRAME_ERRORS = {1: "truncated", 2: "magic", 3: "version", 4: "kind", 5: "compression",
                6: "checksum", 7: "body"}


class FrameHeader(C.Structure):
    _fields_ = [("id", C.c_uint64), ("stage_attempt", C.c_uint32), ("map", C.c_uint32),
                ("attempt", C.c_uint32), ("sequence", C.c_uint32), ("reduce", C.c_uint32),
                ("kind", C.c_uint32)]


def crc32c(data: bytes) -> int:
    buf = np.frombuffer(data, dtype=np.uint8) if data else np.zeros(1, np.uint8)
    return lib().kgo_crc32c(_ptr(buf), len(data))


def frame_encode(h: FrameHeader, body: bytes) -> bytes:
    out = np.zeros(44, dtype=np.uint8)
    buf = np.frombuffer(body, dtype=np.uint8) if body else np.zeros(1, np.uint8)
    lib().kgo_frame_header(C.byref(h), _ptr(buf), len(body), _ptr(out))
    return out.tobytes() + body


def frame_decode(frame: bytes):
    """(header, body) of one frame; ValueError naming the refusal otherwise."""
    h = FrameHeader()
    off, ln = C.c_uint64(), C.c_uint64()
    buf = np.frombuffer(frame, dtype=np.uint8) if frame else np.zeros(1, np.uint8)
    r = lib().kgo_frame_decode(_ptr(buf), len(frame), C.byref(h), C.byref(off), C.byref(ln))
    if r:
        raise ValueError(FRAME_ERRORS[r])
    return h, frame[off.value:off.value + ln.value]
