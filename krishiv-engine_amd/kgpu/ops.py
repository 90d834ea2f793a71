"""Host-side mirror of the reference engine's operator interface for the GPU
path: Arrow RecordBatch (or numpy columns) in, RecordBatch out.

  HashAggregate  -- Sink (sink per batch, finalize) + Source (result):
                    crates/krishiv-exec/src/aggregate/sink.rs:222,1459-1585
  HashJoin       -- build Sink + probe Operator sharing one table:
                    crates/krishiv-exec/src/join/mod.rs:1014-1211
  JoinAggregate  -- inner join fused with GROUP BY the build key
  RuntimeBloom   -- crates/krishiv-io/src/runtime_filter/bloom.rs
  route / repartition -- crates/krishiv-exec/src/exchange/mod.rs:159-212,447-501

Every method calls libkgpu through ctypes; nothing is computed here.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (AGG_ANSI, COL_DEVICE, WIDTH, AggSpec, ArrowArray, ArrowDeviceArray, ArrowSchema,
                   Column, Expr, FilterOptions, FrameHeader, GraceJoinStats, JoinSpec,
                   KeyFilterInfo, KgError, check, lib,
                   T_BOOL, T_DECIMAL128, T_DECIMAL64, T_INT8, T_INT16, T_INT32, T_INT64,
                   T_UINT8, T_UINT16, T_UINT32, T_UINT64, T_FLOAT32, T_FLOAT64, T_DATE32,
                   T_TIMESTAMP)

_NP = {T_INT8: np.int8, T_INT16: np.int16, T_INT32: np.int32, T_INT64: np.int64,
       T_UINT8: np.uint8, T_UINT16: np.uint16, T_UINT32: np.uint32, T_UINT64: np.uint64,
       T_FLOAT32: np.float32, T_FLOAT64: np.float64, T_DATE32: np.int32, T_TIMESTAMP: np.int64,
       T_DECIMAL64: np.int64, T_BOOL: np.uint8}


# ---------------------------------------------------------------- context
class Context:
    """One GPU and an HBM budget (kg_ctx)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        check(lib().kg_ctx_create(device, hbm_budget, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().kg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def hbm_in_use(self) -> int:
        return lib().kg_ctx_hbm_in_use(self.h)

    def mem_stats(self) -> dict:
        """The device allocator's counters (kg_ctx_mem_stats)."""
        v = (C.c_uint64 * 6)()
        check(lib().kg_ctx_mem_stats(self.h, v))
        return dict(zip(("in_use", "cached", "hip_mallocs", "hip_frees", "cache_hits",
                         "releases"), list(v)))

    def synchronize(self):
        check(lib().kg_ctx_synchronize(self.h))


class DeviceBuffer:
    """Plain HBM owned by the caller (inputs kept resident)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = max(int(nbytes), 8)
        p = C.c_void_p()
        check(lib().kg_dev_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    @staticmethod
    def from_numpy(ctx: Context, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(ctx, a.nbytes)
        if a.nbytes:
            check(lib().kg_memcpy_h2d(ctx.h, b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_numpy(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().kg_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().kg_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- columns
class Col:
    """A column for the C ABI: host numpy/pyarrow buffers, or device buffers.

    ``values`` may be a numpy array (host), a DeviceBuffer (device), or None.
    ``valid`` a bool numpy array (host, converted to an LSB bitmap), a
    DeviceBuffer holding an LSB bitmap, or None (no NULLs)."""

    def __init__(self, ctype, values, valid=None, precision=0, scale=0, length=None, offset=0):
        self.type, self.precision, self.scale = ctype, precision, scale
        self.offset = offset
        self._keep = []
        self.device = isinstance(values, DeviceBuffer)
        if self.device:
            self.length = int(length)
            self.vptr = values.ptr
            self._keep.append(values)
            if isinstance(valid, DeviceBuffer):
                self.bptr = valid.ptr
                self._keep.append(valid)
            elif valid is None:
                self.bptr = None
            else:
                raise TypeError("a device column takes a device validity bitmap")
            return
        if ctype == T_BOOL:
            v = np.asarray(values, dtype=bool)
            self.length = len(v) if length is None else length
            self.values = np.packbits(v, bitorder="little")
        elif ctype == T_DECIMAL128:
            v = np.asarray(values)
            if v.dtype != np.uint64 or v.ndim != 2:
                ints = [int(x) for x in values]
                v = np.zeros((len(ints), 2), dtype=np.uint64)
                for i, x in enumerate(ints):
                    u = x & ((1 << 128) - 1)
                    v[i, 0] = u & 0xFFFFFFFFFFFFFFFF
                    v[i, 1] = u >> 64
            self.values = np.ascontiguousarray(v)
            self.length = self.values.shape[0] if length is None else length
        else:
            self.values = np.ascontiguousarray(values, dtype=_NP[ctype])
            self.length = len(self.values) if length is None else length
        self.vptr = self.values.ctypes.data if self.values.size else None
        if valid is None:
            self.bitmap = None
        else:
            self.bitmap = np.packbits(np.asarray(valid, dtype=bool), bitorder="little")
        self.bptr = None if self.bitmap is None else self.bitmap.ctypes.data

    @staticmethod
    def from_arrow(arr) -> "Col":
        """Zero-copy view of a pyarrow primitive array's buffers."""
        import pyarrow as pa
        arr = arr.combine_chunks() if isinstance(arr, pa.ChunkedArray) else arr
        t = arr.type
        m = {pa.int8(): T_INT8, pa.int16(): T_INT16, pa.int32(): T_INT32, pa.int64(): T_INT64,
             pa.uint8(): T_UINT8, pa.uint16(): T_UINT16, pa.uint32(): T_UINT32,
             pa.uint64(): T_UINT64, pa.float32(): T_FLOAT32, pa.float64(): T_FLOAT64,
             pa.date32(): T_DATE32, pa.bool_(): T_BOOL}
        prec = scale = 0
        if t in m:
            ctype = m[t]
        elif pa.types.is_decimal(t):
            prec, scale = t.precision, t.scale
            ctype = T_DECIMAL128 if t.bit_width == 128 else T_DECIMAL64
            if t.bit_width not in (64, 128):
                raise KgError(1, "KRISHIV_UNSUPPORTED_FEATURE: decimal width")
        elif pa.types.is_timestamp(t) or pa.types.is_date64(t) or pa.types.is_duration(t):
            ctype = T_TIMESTAMP
        else:
            raise KgError(1, f"KRISHIV_UNSUPPORTED_FEATURE: {t}")
        bufs = arr.buffers()
        c = Col.__new__(Col)
        c.type, c.precision, c.scale = ctype, prec, scale
        c.length = len(arr)
        c.offset = arr.offset
        c.device = False
        c._keep = [arr]
        c.vptr = bufs[1].address if bufs[1] is not None else None
        c.bptr = bufs[0].address if (bufs[0] is not None and arr.null_count > 0) else None
        return c

    def c(self) -> Column:
        return Column(self.type, self.precision, self.scale, COL_DEVICE if self.device else 0,
                      self.length, self.offset, self.vptr, self.bptr)


def _cols(cols):
    arr = (Column * max(1, len(cols)))()
    for i, c in enumerate(cols):
        arr[i] = c.c()
    return arr


def _names(names):
    if names is None:
        return None
    enc = [n.encode() for n in names]
    return (C.c_char_p * len(enc))(*enc)


def arrow_col(array, on_device: bool = False, schema: ArrowSchema | None = None) -> "Col":
    """A Col viewing an Arrow array without a copy (kg_column_from_arrow):
    a pyarrow Array (host), or an exported (ArrowArray, ArrowSchema) /
    (ArrowDeviceArray, ArrowSchema) pair."""
    keep = [array]
    out = Column()
    if isinstance(array, ArrowDeviceArray):
        check(lib().kg_column_from_arrow_device(C.byref(array), C.byref(schema), C.byref(out)))
    elif isinstance(array, ArrowArray):
        check(lib().kg_column_from_arrow(C.byref(array), C.byref(schema), int(on_device),
                                         C.byref(out)))
    else:
        arr, sch = ArrowArray(), ArrowSchema()
        array._export_to_c(C.addressof(arr), C.addressof(sch))
        keep = [array, _Released(arr, sch)]
        check(lib().kg_column_from_arrow(C.byref(arr), C.byref(sch), 0, C.byref(out)))
    c = Col.__new__(Col)
    c.type, c.precision, c.scale = out.type, out.precision, out.scale
    c.length, c.offset, c.device = out.length, out.offset, bool(out.flags & COL_DEVICE)
    c.vptr, c.bptr, c._keep = out.values, out.validity, keep
    return c


class _Released:
    """Calls an exported Arrow array's and schema's release when dropped."""

    def __init__(self, arr, sch):
        self.arr, self.sch = arr, sch

    def __del__(self):
        try:
            if self.arr.release:
                self.arr.release(C.pointer(self.arr))
            if self.sch.release:
                self.sch.release(C.pointer(self.sch))
        except Exception:
            pass


def _type_col(ctype, precision=0, scale=0) -> Column:
    return Column(ctype, precision, scale, 0, 0, 0, None, None)


# ---------------------------------------------------------------- batches
class Batch:
    """A device-resident batch (kg_batch) owned by the library.  It keeps its
    context alive: a batch must be released before its context."""

    def __init__(self, handle, ctx=None):
        self.h = handle
        self.ctx = ctx

    def __del__(self):
        try:
            if self.h:
                lib().kg_batch_release(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def num_rows(self) -> int:
        return lib().kg_batch_num_rows(self.h)

    @property
    def num_columns(self) -> int:
        return lib().kg_batch_num_columns(self.h)

    def device_column(self, i) -> Column:
        c = Column()
        check(lib().kg_batch_column(self.h, i, C.byref(c)))
        return c

    def as_col(self, i) -> Col:
        """Column i as a device Col (no copy) for another operator."""
        d = self.device_column(i)
        c = Col.__new__(Col)
        c.type, c.precision, c.scale = d.type, d.precision, d.scale
        c.length, c.offset, c.device = d.length, 0, True
        c.vptr, c.bptr, c._keep = d.values, d.validity, [self]
        return c

    def to_arrow(self, names=None):
        """A pyarrow RecordBatch copied to host memory (kg_batch_export_host)."""
        import pyarrow as pa
        arr, sch = ArrowArray(), ArrowSchema()
        check(lib().kg_batch_export_host(self.h, _names(names), C.byref(arr), C.byref(sch)))
        return pa.RecordBatch._import_from_c(C.addressof(arr), C.addressof(sch))

    def export_device(self, names=None):
        """(ArrowDeviceArray, ArrowSchema) in HBM without a copy; the batch
        moves into the export (call the array's release to free it)."""
        arr, sch = ArrowDeviceArray(), ArrowSchema()
        check(lib().kg_batch_export_device(self.h, _names(names), C.byref(arr), C.byref(sch)))
        self.h = None
        return arr, sch

    def column(self, i):
        """(values, valid) on the host: numpy values (python ints for
        decimal128) and a bool validity array."""
        d = self.device_column(i)
        n = self.num_rows
        valid = np.ones(n, dtype=np.uint8)
        if d.type == T_DECIMAL128:
            raw = np.zeros((n, 2), dtype=np.uint64)
            check(lib().kg_batch_column_to_host(self.h, i, raw.ctypes.data, valid.ctypes.data))
            vals = [((int(h) << 64) | int(l)) for l, h in raw]
            vals = [x - (1 << 128) if x >= (1 << 127) else x for x in vals]
            return vals, valid.astype(bool)
        dt = _NP[d.type]
        vals = np.zeros(n, dtype=dt)
        check(lib().kg_batch_column_to_host(self.h, i, vals.ctypes.data, valid.ctypes.data))
        if d.type == T_BOOL:
            vals = vals.astype(bool)
        return vals, valid.astype(bool)

    def column_raw(self, i):
        """(values, valid) on the host without decoding: decimal128 as an
        (n, 2) int64 array of (low, high) words, other types as numpy."""
        d = self.device_column(i)
        n = self.num_rows
        valid = np.ones(n, dtype=np.uint8)
        if d.type == T_DECIMAL128:
            vals = np.zeros((n, 2), dtype=np.int64)
        else:
            vals = np.zeros(n, dtype=_NP[d.type])
        check(lib().kg_batch_column_to_host(self.h, i, vals.ctypes.data, valid.ctypes.data))
        return vals, valid.astype(bool)

    def column_type(self, i):
        d = self.device_column(i)
        return d.type, d.precision, d.scale

    def to_arrow(self, names=None):
        """The batch as a pyarrow RecordBatch (host copy)."""
        import pyarrow as pa
        arrays = []
        for i in range(self.num_columns):
            t, p, s = self.column_type(i)
            vals, valid = self.column(i)
            mask = ~valid
            if t == T_DECIMAL128 or t == T_DECIMAL64:
                from decimal import Decimal
                typ = pa.decimal128(max(p, 1), s)
                pyv = [None if m else Decimal(int(v)).scaleb(-s) for v, m in zip(vals, mask)]
                arrays.append(pa.array(pyv, type=typ))
            elif t == T_DATE32:
                arrays.append(pa.array(vals, mask=mask, type=pa.date32()))
            else:
                arrays.append(pa.array(vals, mask=mask))
        names = names or [f"c{i}" for i in range(len(arrays))]
        return pa.RecordBatch.from_arrays(arrays, names=names)


# ---------------------------------------------------------------- exchange
def hash_columns(ctx: Context, cols) -> np.ndarray:
    n = cols[0].length
    out = np.zeros(n, dtype=np.uint64)
    ca = _cols(cols)
    check(lib().kg_hash_columns(ctx.h, len(cols), ca, n, out.ctypes.data, 0))
    return out


def route(ctx: Context, keys, partitions: int) -> np.ndarray:
    n = keys[0].length
    out = np.zeros(n, dtype=np.uint32)
    ca = _cols(keys)
    check(lib().kg_partition(ctx.h, len(keys), ca, n, partitions, out.ctypes.data, 0))
    return out


def join_partition(ctx: Context, keys, radix_bits: int, depth: int, seed: int):
    """(h, partition) of each row: the join's row-layout hash of the key
    columns and its grace partition at `depth` (host numpy arrays)."""
    n = keys[0].length
    h = np.zeros(n, dtype=np.uint64)
    p = np.zeros(n, dtype=np.uint8)
    if n:
        check(lib().kg_join_partition(ctx.h, len(keys), _cols(keys), n, radix_bits, depth,
                                      seed, h.ctypes.data, p.ctypes.data, 0))
    return h, p


def repartition(ctx: Context, cols, key_cols, partitions: int):
    """Rows regrouped by partition (stable): (Batch, offsets[partitions+1])."""
    n = cols[0].length
    ca = _cols(cols)
    kc = (C.c_int32 * max(1, len(key_cols)))(*key_cols)
    offs = np.zeros(partitions + 1, dtype=np.int64)
    h = C.c_void_p()
    check(lib().kg_repartition(ctx.h, len(cols), ca, n, len(key_cols), kc, partitions,
                               offs.ctypes.data, C.byref(h)))
    return Batch(h, ctx), offs


def select_range(ctx: Context, col: Col, lo: int, hi: int, out_rows: DeviceBuffer | None = None):
    """Row ids whose value is in [lo, hi] (NULLs never pass), in order:
    (DeviceBuffer of int32, count)."""
    out = out_rows or DeviceBuffer(ctx, max(col.length, 1) * 4)
    cnt = C.c_int64(0)
    c = col.c()
    check(lib().kg_select_range(ctx.h, C.byref(c), col.length, lo, hi, C.c_void_p(out.ptr),
                                C.byref(cnt)))
    return out, cnt.value


def gather(ctx: Context, cols, rows: DeviceBuffer, n: int) -> Batch:
    """The rows (device int32 ids) of `cols` as a device batch."""
    h = C.c_void_p()
    check(lib().kg_gather(ctx.h, len(cols), _cols(cols), C.c_void_p(rows.ptr), n, C.byref(h)))
    return Batch(h, ctx)


# ---------------------------------------------------------------- aggregation
class HashAggregate:
    """GROUP BY one key column with count/sum/min/max/avg aggregates."""

    def __init__(self, ctx: Context, key_type, input_types, aggs, ansi=False,
                 expected_groups=1 << 16, partial=False, merge=False):
        """partial: emit partial states (PartialAggregate); merge: sink the
        partial states of a first phase (the final phase)."""
        self.ctx = ctx
        kt = _type_col(*key_type) if key_type is not None else None
        its = (Column * max(1, len(input_types)))(*[_type_col(*t) for t in input_types])
        sp = (AggSpec * max(1, len(aggs)))(*[AggSpec(f, i) for f, i in aggs])
        h = C.c_void_p()
        flags = (AGG_ANSI if ansi else 0) | (2 if partial else 0) | (4 if merge else 0)
        check(lib().kg_agg_create(ctx.h, C.byref(kt) if kt is not None else None,
                                  len(input_types), its, len(aggs), sp,
                                  flags, expected_groups, C.byref(h)))
        self.h = h
        self.naggs = len(aggs)
        self.has_key = key_type is not None

    def sink(self, key: Col | None, inputs, nrows=None):
        if nrows is None:
            nrows = key.length if key is not None else inputs[0].length
        kc = key.c() if key is not None else None
        ca = _cols(inputs)
        check(lib().kg_agg_sink(self.h, C.byref(kc) if kc is not None else None, ca, nrows))

    def finalize(self):
        check(lib().kg_agg_finalize(self.h))

    def result_type(self, i):
        c = Column()
        check(lib().kg_agg_result_type(self.h, i, C.byref(c)))
        return c.type, c.precision, c.scale

    def result(self) -> Batch:
        h = C.c_void_p()
        check(lib().kg_agg_result(self.h, C.byref(h)))
        return Batch(h, self.ctx)

    def set_bypass(self, min_rows=100_000, max_group_percent=80, max_table_bytes=16 << 20):
        """BypassPolicy (aggregate/partial.rs:20-50) of a partial aggregate."""
        check(lib().kg_agg_set_bypass(self.h, min_rows, max_group_percent, max_table_bytes))

    def set_merge(self, form=0, size_from_sketch=True):
        """The drain's merge form (0 auto, 1 windows in LDS, 2 by key) and
        whether the table is sized from the sketch of the spilled keys."""
        check(lib().kg_agg_set_merge(self.h, form, 1 if size_from_sketch else 0))

    def partial_batches(self):
        """The partial-state batches flushed or passed through so far."""
        out = []
        while True:
            h, done = C.c_void_p(), C.c_int32()
            check(lib().kg_agg_partial_next(self.h, C.byref(h), C.byref(done)))
            if h.value:
                out.append(Batch(h, self.ctx))
            if done.value:
                return out

    def partial_stats(self):
        """(flushes, bypassed_rows)."""
        f, r = C.c_uint64(), C.c_uint64()
        check(lib().kg_agg_partial_stats(self.h, C.byref(f), C.byref(r)))
        return f.value, r.value

    def close(self):
        if self.h:
            lib().kg_agg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- join
class HashJoin:
    def __init__(self, ctx: Context, join_type, probe_keys, build_keys, null_safe=None,
                 expected_build_rows=1 << 16, direct_index=True):
        self.ctx = ctx
        spec = JoinSpec()
        spec.join_type = join_type
        spec.num_keys = len(probe_keys)
        spec.no_direct_index = 0 if direct_index else 1
        for i, (p, b) in enumerate(zip(probe_keys, build_keys)):
            spec.probe_keys[i] = p
            spec.build_keys[i] = b
            spec.null_safe[i] = int(bool(null_safe[i])) if null_safe else 0
        h = C.c_void_p()
        check(lib().kg_join_create(ctx.h, C.byref(spec), expected_build_rows, C.byref(h)))
        self.h = h

    def build_sink(self, cols, nrows=None):
        nrows = cols[0].length if nrows is None else nrows
        check(lib().kg_join_build_sink(self.h, len(cols), _cols(cols), nrows))

    def build_finalize(self):
        check(lib().kg_join_build_finalize(self.h))

    def build_rows(self) -> int:
        return lib().kg_join_build_rows(self.h)

    def probe(self, cols, nrows=None) -> Batch:
        nrows = cols[0].length if nrows is None else nrows
        h = C.c_void_p()
        check(lib().kg_join_probe(self.h, len(cols), _cols(cols), nrows, C.byref(h)))
        return Batch(h, self.ctx)

    def set_filter(self, program):
        """The residual filter: a list of (op, arg, imm) postfix ops (X_*)."""
        ops = (Expr * max(1, len(program)))(*[Expr(*x) for x in program])
        check(lib().kg_join_set_filter(self.h, len(program), ops))

    def local(self) -> "JoinLocal":
        """A driver's probe state (its own stream), for bounded output."""
        return JoinLocal(self)

    def tail(self, probe_types, max_rows=1 << 62):
        """The build rows the join type returns after every probe (right
        semi/anti/outer, full outer), as a list of batches of at most
        max_rows rows; [] for the other join types."""
        pt = (Column * max(1, len(probe_types)))(*[_type_col(*t) for t in probe_types])
        out = []
        while True:
            h = C.c_void_p()
            done = C.c_int32(0)
            check(lib().kg_join_tail_next(self.h, len(probe_types), pt, max_rows, C.byref(h),
                                          C.byref(done)))
            if h.value:
                out.append(Batch(h, self.ctx))
            if done.value:
                return out

    def probe_selected(self, cols, sel: DeviceBuffer, nsel: int, nrows=None) -> Batch:
        nrows = cols[0].length if nrows is None else nrows
        h = C.c_void_p()
        check(lib().kg_join_probe_selected(self.h, len(cols), _cols(cols), nrows,
                                           C.c_void_p(sel.ptr), nsel, C.byref(h)))
        return Batch(h, self.ctx)

    # join fused with GROUP BY the build key
    def aggregate(self, input_types, aggs, ansi=False):
        its = (Column * max(1, len(input_types)))(*[_type_col(*t) for t in input_types])
        sp = (AggSpec * max(1, len(aggs)))(*[AggSpec(f, i) for f, i in aggs])
        check(lib().kg_joinagg_create(self.h, len(input_types), its, len(aggs), sp,
                                      AGG_ANSI if ansi else 0))

    def probe_aggregate(self, cols, input_cols, nrows=None):
        nrows = cols[0].length if nrows is None else nrows
        ic = (C.c_int32 * max(1, len(input_cols)))(*input_cols)
        check(lib().kg_joinagg_probe(self.h, len(cols), _cols(cols), nrows, ic))

    def aggregate_result(self) -> Batch:
        h = C.c_void_p()
        check(lib().kg_joinagg_result(self.h, C.byref(h)))
        return Batch(h, self.ctx)

    def close(self):
        if self.h:
            lib().kg_join_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GraceJoin:
    """A HashJoin whose build side may pass an HBM budget (kg_grace_join):
    partitions spill to pinned host memory and are joined after finish()."""

    def __init__(self, ctx: Context, join_type, probe_keys, build_keys, table_budget,
                 null_safe=None):
        self.ctx = ctx
        spec = JoinSpec()
        spec.join_type = join_type
        spec.num_keys = len(probe_keys)
        spec.no_direct_index = 0
        for i, (p, b) in enumerate(zip(probe_keys, build_keys)):
            spec.probe_keys[i] = p
            spec.build_keys[i] = b
            spec.null_safe[i] = int(bool(null_safe[i])) if null_safe else 0
        h = C.c_void_p()
        check(lib().kg_grace_join_create(ctx.h, C.byref(spec), table_budget, C.byref(h)))
        self.h = h

    def set_filter(self, program):
        ops = (Expr * max(1, len(program)))(*[Expr(*x) for x in program])
        check(lib().kg_grace_join_set_filter(self.h, len(program), ops))

    def build_sink(self, cols, nrows=None):
        nrows = cols[0].length if nrows is None else nrows
        check(lib().kg_grace_join_build_sink(self.h, len(cols), _cols(cols), nrows))

    def build_finalize(self):
        check(lib().kg_grace_join_build_finalize(self.h))

    def _drain(self, max_rows):
        out = []
        while True:
            h, done = C.c_void_p(), C.c_int32(0)
            check(lib().kg_grace_join_next(self.h, max_rows, C.byref(h), C.byref(done)))
            if h.value:
                out.append(Batch(h, self.ctx))
            if done.value:
                return out

    def probe(self, cols, nrows=None, max_rows=1 << 20):
        """Pushes a probe batch; its output batches (none while spilled)."""
        nrows = cols[0].length if nrows is None else nrows
        check(lib().kg_grace_join_probe_push(self.h, len(cols), _cols(cols), nrows))
        return self._drain(max_rows)

    def finish(self, max_rows=1 << 20):
        """Every probe batch pushed: the remaining output batches."""
        check(lib().kg_grace_join_finish(self.h))
        return self._drain(max_rows)

    def stats(self) -> dict:
        st = GraceJoinStats()
        check(lib().kg_grace_join_stats_get(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in GraceJoinStats._fields_}

    def close(self):
        if self.h:
            lib().kg_grace_join_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class JoinLocal:
    """A driver's probe state over a finalized HashJoin (kg_join_local):
    push a batch, then take its output in batches of at most max_rows."""

    def __init__(self, join: HashJoin):
        self.join = join
        h = C.c_void_p()
        check(lib().kg_join_local_create(join.h, C.byref(h)))
        self.h = h
        self._keep = None

    def push(self, cols, nrows=None, sel: DeviceBuffer | None = None, nsel: int = 0):
        nrows = cols[0].length if nrows is None else nrows
        self._keep = (cols, sel)
        check(lib().kg_join_probe_push(self.h, len(cols), _cols(cols), nrows,
                                       C.c_void_p(sel.ptr) if sel is not None else None, nsel))

    def next(self, max_rows: int):
        """(Batch, done)."""
        h = C.c_void_p()
        done = C.c_int32(0)
        check(lib().kg_join_probe_next(self.h, max_rows, C.byref(h), C.byref(done)))
        return Batch(h, self.join.ctx), bool(done.value)

    def probe_all(self, cols, max_rows: int, nrows=None, sel=None, nsel=0):
        """Every output batch of one probe batch."""
        self.push(cols, nrows, sel, nsel)
        out = []
        while True:
            b, done = self.next(max_rows)
            out.append(b)
            if done:
                return out

    def close(self):
        if self.h:
            lib().kg_join_local_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- bloom
def _i128(lo, hi):
    v = (int(hi) << 64) | int(lo)
    return v - (1 << 128) if v >= 1 << 127 else v


class KeyFilter:
    """A runtime KeyFilterBuilder on the device until finish(), then the
    KeyFilter it published (runtime_filter/filter.rs:180-615)."""

    def __init__(self, ctx: Context, key_type=None, options: FilterOptions | None = None,
                 _handle=None):
        self.ctx = ctx
        if _handle is not None:
            self.h = _handle
            return
        kt = _type_col(*key_type)
        h = C.c_void_p()
        check(lib().kg_key_filter_create(ctx.h, C.byref(kt),
                                         C.byref(options) if options is not None else None,
                                         C.byref(h)))
        self.h = h

    @staticmethod
    def options(**kw) -> FilterOptions:
        o = FilterOptions()
        lib().kg_filter_default_options(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def add(self, key: Col, include: Col | None = None, nrows=None):
        nrows = key.length if nrows is None else nrows
        kc = key.c()
        ic = include.c() if include is not None else None
        check(lib().kg_key_filter_add(self.h, C.byref(kc),
                                      C.byref(ic) if ic is not None else None, nrows))
        return self

    def merge_builder(self, other: "KeyFilter"):
        check(lib().kg_key_filter_merge_builder(self.h, other.h))
        return self

    def finish(self):
        check(lib().kg_key_filter_finish(self.h))
        return self

    def info(self) -> dict:
        i = KeyFilterInfo()
        check(lib().kg_key_filter_info_get(self.h, C.byref(i)))
        return {"form": i.form, "has_range": bool(i.has_range), "nmembers": i.members,
                "values": i.values, "lo": _i128(i.lo[0], i.lo[1]), "hi": _i128(i.hi[0], i.hi[1]),
                "flo": i.flo, "fhi": i.fhi, "nulls_pass": bool(i.nulls_pass),
                "empty": bool(i.empty), "keys": i.keys, "bloom_bytes": i.bloom_bytes,
                "row_cost_ns": i.row_cost_ns}

    def members(self):
        n = C.c_int64(0)
        check(lib().kg_key_filter_members(self.h, None, C.byref(n)))
        w = np.zeros(2 * max(n.value, 1), dtype=np.uint64)
        check(lib().kg_key_filter_members(self.h, C.c_void_p(w.ctypes.data), C.byref(n)))
        return [_i128(w[2 * k], w[2 * k + 1]) for k in range(n.value)]

    def evaluate(self, col: Col, nrows=None, out_rows: DeviceBuffer | None = None):
        """Row ids that may match: (DeviceBuffer of int32 ids, count)."""
        nrows = col.length if nrows is None else nrows
        out = out_rows or DeviceBuffer(self.ctx, max(nrows, 1) * 4)
        cnt = C.c_int64(0)
        cc = col.c()
        check(lib().kg_key_filter_evaluate(self.h, C.byref(cc), nrows, C.c_void_p(out.ptr),
                                           C.byref(cnt)))
        return out, cnt.value

    def encode(self) -> bytes:
        n = C.c_uint64(0)
        check(lib().kg_key_filter_encode(self.h, None, C.byref(n)))
        buf = (C.c_uint8 * max(n.value, 1))()
        check(lib().kg_key_filter_encode(self.h, buf, C.byref(n)))
        return bytes(buf)[:n.value]

    @classmethod
    def decode(cls, ctx: Context, data: bytes) -> "KeyFilter":
        h = C.c_void_p()
        buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data or b"\0")
        check(lib().kg_key_filter_decode(ctx.h, buf, len(data), C.byref(h)))
        return cls(ctx, _handle=h)

    def merge(self, other: "KeyFilter", options: FilterOptions | None = None) -> "KeyFilter":
        h = C.c_void_p()
        check(lib().kg_key_filter_merge(self.h, other.h,
                                        C.byref(options) if options is not None else None,
                                        C.byref(h)))
        return KeyFilter(self.ctx, _handle=h)

    def close(self):
        if self.h:
            lib().kg_key_filter_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RowStage:
    """A scan driver's runtime-filter row stage (operators/runtime_filter.rs)."""

    def __init__(self, ctx: Context, min_rows=65_536, max_pass_percent=90):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().kg_row_stage_create(ctx.h, min_rows, max_pass_percent, C.byref(h)))
        self.h = h
        self.keep = {}

    def set(self, slot, f: KeyFilter | None, column):
        self.keep[slot] = f
        check(lib().kg_row_stage_set(self.h, slot, f.h if f is not None else None, column))

    def start_partition(self):
        check(lib().kg_row_stage_start_partition(self.h))

    def apply(self, cols, nrows=None, out_rows: DeviceBuffer | None = None):
        """(DeviceBuffer of row ids, count), or (None, -1) when nothing is
        filtered."""
        nrows = cols[0].length if nrows is None else nrows
        out = out_rows or DeviceBuffer(self.ctx, max(nrows, 1) * 4)
        cnt = C.c_int64(0)
        check(lib().kg_row_stage_apply(self.h, len(cols), _cols(cols), nrows, C.c_void_p(out.ptr),
                                       C.byref(cnt)))
        return (None, -1) if cnt.value < 0 else (out, cnt.value)

    def stats(self, slot):
        t, d, o = C.c_uint64(), C.c_uint64(), C.c_int32()
        check(lib().kg_row_stage_stats(self.h, slot, C.byref(t), C.byref(d), C.byref(o)))
        return t.value, d.value, bool(o.value)

    def close(self):
        if self.h:
            lib().kg_row_stage_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def exchange_frames(ctx: Context, cols, part_offsets, base: FrameHeader, names=None):
    """The Data frames of one split batch (kg_exchange_frames): (DeviceBuffer
    of the frames back to back, frame offsets (partitions + 1), total bytes)."""
    offs = np.ascontiguousarray(part_offsets, dtype=np.int64)
    parts = len(offs) - 1
    fo = np.zeros(parts + 1, dtype=np.uint64)
    total = C.c_uint64(0)
    nm = None
    if names is not None:
        enc = [n.encode() for n in names]
        nm = (C.c_char_p * len(enc))(*enc)
    cc = _cols(cols)
    check(lib().kg_exchange_frames(ctx.h, len(cols), cc, nm, C.c_void_p(offs.ctypes.data), parts,
                                   C.byref(base), None, 0, C.c_void_p(fo.ctypes.data),
                                   C.byref(total)))
    out = DeviceBuffer(ctx, max(total.value, 8))
    check(lib().kg_exchange_frames(ctx.h, len(cols), cc, nm, C.c_void_p(offs.ctypes.data), parts,
                                   C.byref(base), C.c_void_p(out.ptr), out.nbytes,
                                   C.c_void_p(fo.ctypes.data), C.byref(total)))
    return out, fo, total.value


def frame_end(h: FrameHeader) -> bytes:
    buf = (C.c_uint8 * 44)()
    check(lib().kg_frame_end(C.byref(h), buf))
    return bytes(buf)


def crc32c(ctx: Context, data: DeviceBuffer, nbytes: int) -> int:
    out = C.c_uint32(0)
    check(lib().kg_crc32c(ctx.h, C.c_void_p(data.ptr), nbytes, C.byref(out)))
    return out.value


class RuntimeBloom:
    """The runtime filter's split-block Bloom filter (Parquet layout), or with
    spark_bits > 0 Spark's BloomFilterImpl layout of that many bits and
    spark_hashes hash functions (items: Spark's xxhash64 of the key)."""

    def __init__(self, ctx: Context, nbytes: int = 0, spark_bits: int = 0, spark_hashes: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        if spark_bits:
            check(lib().kg_spark_bloom_create(ctx.h, spark_bits, spark_hashes, C.byref(h)))
        else:
            check(lib().kg_bloom_create(ctx.h, nbytes, C.byref(h)))
        self.h = h

    @staticmethod
    def bytes_for(items: int, fpp: float) -> int:
        return lib().kg_bloom_bytes_for(items, fpp)

    @staticmethod
    def spark_bits_for(items: int, fpp: float) -> int:
        return lib().kg_spark_bloom_bits_for(items, fpp)

    @staticmethod
    def spark_hashes_for(items: int, bits: int) -> int:
        return lib().kg_spark_bloom_hashes_for(items, bits)

    def insert(self, key: Col, select: Col | None = None, nrows=None):
        nrows = key.length if nrows is None else nrows
        kc = key.c()
        sc = select.c() if select is not None else None
        check(lib().kg_bloom_insert(self.h, C.byref(kc), C.byref(sc) if sc is not None else None,
                                    nrows))

    def expected_fpp(self, items: int) -> float:
        return lib().kg_bloom_expected_fpp(self.h, items)

    def to_bytes(self) -> bytes:
        n = C.c_uint64(0)
        check(lib().kg_bloom_export(self.h, None, C.byref(n)))
        buf = (C.c_uint8 * n.value)()
        check(lib().kg_bloom_export(self.h, buf, C.byref(n)))
        return bytes(buf)

    def probe(self, key: Col, out_rows: DeviceBuffer | None = None, nrows=None):
        """Row ids that may match: (DeviceBuffer of int32 ids, count)."""
        nrows = key.length if nrows is None else nrows
        out = out_rows or DeviceBuffer(self.ctx, max(nrows, 1) * 4)
        cnt = C.c_int64(0)
        kc = key.c()
        check(lib().kg_bloom_probe(self.h, C.byref(kc), nrows, C.c_void_p(out.ptr), C.byref(cnt)))
        return out, cnt.value

    def close(self):
        if self.h:
            lib().kg_bloom_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- profiling
def prof_enable(on=True):
    lib().kg_prof_enable(int(on))


def prof_reset():
    lib().kg_prof_reset()


def prof_get(family: str):
    ms = C.c_double(0)
    n = C.c_int64(0)
    check(lib().kg_prof_get(family.encode(), C.byref(ms), C.byref(n)))
    return ms.value, n.value


class TuneConfig(C.Structure):
    """kg_tune_config: the join / fused-aggregate shape of one build size class."""
    _fields_ = [("load_factor_pct", C.c_int32), ("scatter_rows", C.c_int32),
                ("merge_lanes", C.c_int32), ("_pad", C.c_int32)]

    def as_dict(self) -> dict:
        return {"load_factor_pct": self.load_factor_pct, "scatter_rows": self.scatter_rows,
                "merge_lanes": self.merge_lanes}


def tune_get(ctx, build_rows: int) -> dict:
    c = TuneConfig()
    check(lib().kg_tune_get(ctx.h, build_rows, C.byref(c)))
    return c.as_dict()


def tune_set(ctx, build_rows: int, load_factor_pct: int = None, scatter_rows: int = 0,
             merge_lanes: int = 1024):
    """Forces the shape of build_rows' size class; with nothing but the
    rows given the class returns to its untuned default."""
    if load_factor_pct is None:
        load_factor_pct = 0 if (scatter_rows, merge_lanes) == (0, 1024) else 50
    c = TuneConfig(load_factor_pct, scatter_rows, merge_lanes, 0)
    check(lib().kg_tune_set(ctx.h, build_rows, C.byref(c)))


def tune_save(ctx, path: str):
    check(lib().kg_tune_save(ctx.h, path.encode()))


def tune_load(ctx, path: str) -> int:
    n = C.c_int32(0)
    check(lib().kg_tune_load(ctx.h, path.encode(), C.byref(n)))
    return n.value


def autotune_joinagg(ctx, build_rows: int, probe_rows: int):
    """Times every candidate shape for this build size class and keeps the
    fastest (kg_autotune_joinagg); returns (shape, ms per build + probe,
    candidates timed)."""
    c = TuneConfig()
    ms = C.c_double(0)
    n = C.c_int32(0)
    check(lib().kg_autotune_joinagg(ctx.h, build_rows, probe_rows, C.byref(c), C.byref(ms),
                                    C.byref(n)))
    return c.as_dict(), ms.value, n.value


def calib_random_read(ctx, table_bytes: int, read_bytes: int, ops: int, reps: int = 3) -> float:
    """Random reads per second from a device table of table_bytes (the
    ceiling a probe phase over a table of that size is measured against)."""
    ms = C.c_double(0)
    check(lib().kg_calib_random_read(ctx.h, table_bytes, read_bytes, ops, reps, C.byref(ms)))
    return ops / (ms.value * 1e-3)


def calib_atomic_add(ctx, addresses: int, ops: int, reps: int = 3) -> float:
    """No-return 64-bit global atomic adds per second over `addresses` words."""
    ms = C.c_double(0)
    check(lib().kg_calib_atomic_add(ctx.h, addresses, ops, reps, C.byref(ms)))
    return ops / (ms.value * 1e-3)


# ---------------------------------------------------------------- grace join
class GraceJoinAggregate:
    """Grace hash join fused with SELECT p, COUNT(*), SUM(probe.pay *
    build.pay % 2^31) ... GROUP BY p, p = partition_of(H_kg(key), partitions)."""

    def __init__(self, ctx: Context, partitions=64, table_budget=8 << 30, sum_mod31=False):
        """sum_mod31: SUM read as ((sum of products) mod 2^64) mod 2^31 (any
        payloads); otherwise per-pair (p * b) % 2^31, products in [0, 2^31)."""
        self.ctx = ctx
        self.parts = partitions
        h = C.c_void_p()
        check(lib().kg_grace_create(ctx.h, partitions, table_budget, 1 if sum_mod31 else 0,
                                    C.byref(h)))
        self.h = h

    def build_sink(self, key: Col, pay: Col, nrows=None):
        nrows = key.length if nrows is None else nrows
        k, p = key.c(), pay.c()
        check(lib().kg_grace_build_sink(self.h, C.byref(k), C.byref(p), nrows))

    def probe_sink(self, key: Col, pay: Col, nrows=None):
        nrows = key.length if nrows is None else nrows
        k, p = key.c(), pay.c()
        check(lib().kg_grace_probe_sink(self.h, C.byref(k), C.byref(p), nrows))

    def finish(self):
        """(counts, sums, sum_valid, spilled_bytes)."""
        counts = np.zeros(self.parts, dtype=np.int64)
        sums = np.zeros(self.parts, dtype=np.int64)
        valid = np.zeros(self.parts, dtype=np.uint8)
        spilled = C.c_uint64(0)
        check(lib().kg_grace_finish(self.h, counts.ctypes.data, sums.ctypes.data,
                                    valid.ctypes.data, C.byref(spilled)))
        return counts, sums, valid.astype(bool), spilled.value

    def close(self):
        if self.h:
            lib().kg_grace_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- serving
class JoinAggServer:
    """Batched, continuously served join-aggregate requests (kg_server_*):
    each request is SELECT key, SUM(pay), COUNT(*) FROM build JOIN probe ON
    key = fk GROUP BY key over device-resident int64 columns.  Requests are
    packed into batches admitted by arena bytes; resident build tables are
    shared by the requests that name them."""

    def __init__(self, ctx: Context, arena_bytes=8 << 30, max_batch=256, use_graph=True):
        from ._lib import ServeRequest
        self._Req = ServeRequest
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().kg_server_create(ctx.h, arena_bytes, max_batch, 1 if use_graph else 0,
                                     C.byref(h)))
        self.h = h

    def add_table(self, keys: DeviceBuffer, n: int) -> int:
        tid = C.c_int32(-1)
        check(lib().kg_server_add_table(self.h, keys.ptr, n, C.byref(tid)))
        return tid.value

    def submit(self, rid: int, probe_keys: DeviceBuffer, probe_pay: DeviceBuffer,
               probe_rows: int, build_keys: DeviceBuffer | None = None, build_rows: int = 0,
               table: int = -1, probe_valid: DeviceBuffer | None = None):
        r = self._Req(rid, build_keys.ptr if build_keys is not None else None, build_rows,
                      table, 0, probe_keys.ptr, probe_valid.ptr if probe_valid is not None else None,
                      probe_pay.ptr, probe_rows)
        check(lib().kg_server_submit(self.h, C.byref(r)))

    def progress(self, wait=False, max_ids=4096):
        """Ids of the requests finished since the last call (launches the
        next batch when a stream slot is free)."""
        ids = np.zeros(max_ids, dtype=np.int64)
        nd = C.c_int32(0)
        check(lib().kg_server_progress(self.h, 1 if wait else 0, ids.ctypes.data, max_ids,
                                       C.byref(nd)))
        return ids[:nd.value].tolist()

    def result(self, rid: int) -> Batch:
        """The finished request's (key, sum, count) batch, in HBM."""
        h = C.c_void_p()
        check(lib().kg_server_take_result(self.h, rid, C.byref(h)))
        return Batch(h, self.ctx)

    def pending(self) -> int:
        return int(lib().kg_server_pending(self.h))

    def close(self):
        if self.h:
            lib().kg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
