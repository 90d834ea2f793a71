"""kgpu: the MI355X (gfx950) execution path for Krishiv's hash join, hash
aggregation, runtime Bloom filter and hash repartition, behind a C ABI
(include/kgpu.h) implemented by libkgpu.so.  This package is the Python host
mirror of the engine's operator interface over that ABI."""
from ._lib import KgError, lib, exported_symbols  # noqa: F401
from . import _lib as abi  # noqa: F401
from .ops import (Batch, Col, Context, DeviceBuffer, GraceJoinAggregate, HashAggregate,  # noqa: F401
                  HashJoin, GraceJoin, JoinAggServer, JoinLocal, KeyFilter, RowStage, autotune_joinagg, calib_atomic_add, calib_random_read,
                  tune_get, tune_load, tune_save, tune_set,
                  RuntimeBloom, gather, hash_columns, join_partition, prof_enable, prof_get, prof_reset,
                  repartition, route, select_range, exchange_frames, frame_end, crc32c, arrow_col)
