// Generated by gen_lib.py from include/kgpu.h: the FFI of the MI355X path
// (libkgpu.so).  Do not edit by hand; regenerate after changing the header.
#![allow(non_camel_case_types, non_upper_case_globals, dead_code)]

use std::os::raw::{c_char, c_int, c_void};

pub const KG_ABI_VERSION: i32 = 1;
pub const KG_OK: i32 = 0;
pub const KG_ERR_UNSUPPORTED: i32 = 1;
pub const KG_ERR_OVERFLOW: i32 = 2;
pub const KG_ERR_OUT_OF_RANGE: i32 = 3;
pub const KG_ERR_MEMORY: i32 = 4;
pub const KG_ERR_INTERNAL: i32 = 5;
pub const KG_ERR_DEVICE: i32 = 6;
pub const KG_T_BOOL: i32 = 1;
pub const KG_T_INT8: i32 = 2;
pub const KG_T_INT16: i32 = 3;
pub const KG_T_INT32: i32 = 4;
pub const KG_T_INT64: i32 = 5;
pub const KG_T_UINT8: i32 = 6;
pub const KG_T_UINT16: i32 = 7;
pub const KG_T_UINT32: i32 = 8;
pub const KG_T_UINT64: i32 = 9;
pub const KG_T_FLOAT32: i32 = 10;
pub const KG_T_FLOAT64: i32 = 11;
pub const KG_T_DATE32: i32 = 12;
pub const KG_T_TIMESTAMP: i32 = 13;
pub const KG_T_DECIMAL64: i32 = 14;
pub const KG_T_DECIMAL128: i32 = 15;
pub const KG_COL_DEVICE: u32 = 1;
pub const ARROW_FLAG_DICTIONARY_ORDERED: i32 = 1;
pub const ARROW_FLAG_NULLABLE: i32 = 2;
pub const ARROW_FLAG_MAP_KEYS_SORTED: i32 = 4;
pub const ARROW_DEVICE_CPU: i32 = 1;
pub const ARROW_DEVICE_ROCM: i32 = 10;
pub const ARROW_DEVICE_ROCM_HOST: i32 = 11;
pub const KG_AGG_COUNT_STAR: i32 = 1;
pub const KG_AGG_COUNT: i32 = 2;
pub const KG_AGG_SUM: i32 = 3;
pub const KG_AGG_MIN: i32 = 4;
pub const KG_AGG_MAX: i32 = 5;
pub const KG_AGG_AVG: i32 = 6;
pub const KG_AGG_ANSI: u32 = 1;
pub const KG_AGG_PARTIAL: u32 = 2;
pub const KG_AGG_MERGE: u32 = 4;
pub const KG_AGG_MERGE_AUTO: i32 = 0;
pub const KG_AGG_MERGE_WINDOWS: i32 = 1;
pub const KG_AGG_MERGE_BY_KEY: i32 = 2;
pub const KG_FORM_RANGE: i32 = 0;
pub const KG_FORM_SMALL: i32 = 1;
pub const KG_FORM_EXACT: i32 = 2;
pub const KG_FORM_BLOOM: i32 = 3;
pub const KG_FORM_DENSE: i32 = 4;
pub const KG_SIDE_BOTH: i32 = 0;
pub const KG_SIDE_UPPER: i32 = 1;
pub const KG_SIDE_LOWER: i32 = 2;
pub const KG_ROW_STAGE_SLOTS: i32 = 8;
pub const KG_FRAME_HEADER_BYTES: i32 = 44;
pub const KG_FRAME_DATA: i32 = 1;
pub const KG_FRAME_END: i32 = 2;
pub const KG_FRAME_VALUE: i32 = 3;
pub const KG_JOIN_INNER: i32 = 0;
pub const KG_JOIN_LEFT_OUTER: i32 = 1;
pub const KG_JOIN_LEFT_SEMI: i32 = 2;
pub const KG_JOIN_LEFT_ANTI: i32 = 3;
pub const KG_JOIN_NULL_AWARE_ANTI: i32 = 4;
pub const KG_JOIN_EXISTENCE: i32 = 5;
pub const KG_JOIN_RIGHT_SEMI: i32 = 6;
pub const KG_JOIN_RIGHT_ANTI: i32 = 7;
pub const KG_JOIN_RIGHT_OUTER: i32 = 8;
pub const KG_JOIN_FULL_OUTER: i32 = 9;
pub const KG_X_PROBE: i32 = 1;
pub const KG_X_BUILD: i32 = 2;
pub const KG_X_CONST: i32 = 3;
pub const KG_X_ADD: i32 = 4;
pub const KG_X_SUB: i32 = 5;
pub const KG_X_MUL: i32 = 6;
pub const KG_X_MOD: i32 = 7;
pub const KG_X_EQ: i32 = 8;
pub const KG_X_NE: i32 = 9;
pub const KG_X_LT: i32 = 10;
pub const KG_X_LE: i32 = 11;
pub const KG_X_GT: i32 = 12;
pub const KG_X_GE: i32 = 13;
pub const KG_X_AND: i32 = 14;
pub const KG_X_OR: i32 = 15;
pub const KG_X_NOT: i32 = 16;
pub const KG_X_IS_NULL: i32 = 17;
pub const KG_GRACE_SUM_MOD31: u32 = 1;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_column {
    pub r#type: i32,
    pub precision: i32,
    pub scale: i32,
    pub flags: u32,
    pub length: i64,
    pub offset: i64,
    pub values: *const c_void,
    pub validity: *const u8,
}

#[repr(C)]
pub struct kg_ctx {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_mem_stats {
    pub in_use: u64,
    pub cached: u64,
    pub hip_mallocs: u64,
    pub hip_frees: u64,
    pub cache_hits: u64,
    pub releases: u64,
}

#[repr(C)]
pub struct kg_batch {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ArrowSchema {
    pub format: *const c_char,
    pub name: *const c_char,
    pub metadata: *const c_char,
    pub flags: i64,
    pub n_children: i64,
    pub children: *mut *mut ArrowSchema,
    pub dictionary: *mut ArrowSchema,
    pub release: Option<unsafe extern "C" fn(*mut ArrowSchema)>,
    pub private_data: *mut c_void,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ArrowArray {
    pub length: i64,
    pub null_count: i64,
    pub offset: i64,
    pub n_buffers: i64,
    pub n_children: i64,
    pub buffers: *mut *const c_void,
    pub children: *mut *mut ArrowArray,
    pub dictionary: *mut ArrowArray,
    pub release: Option<unsafe extern "C" fn(*mut ArrowArray)>,
    pub private_data: *mut c_void,
}

pub type ArrowDeviceType = i32;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ArrowDeviceArray {
    pub array: ArrowArray,
    pub device_id: i64,
    pub device_type: ArrowDeviceType,
    pub sync_event: *mut c_void,
    pub reserved: [i64; 3],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_agg_spec {
    pub func: i32,
    pub input: i32,
}

#[repr(C)]
pub struct kg_hash_agg {
    _private: [u8; 0],
}

#[repr(C)]
pub struct kg_bloom {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_filter_options {
    pub max_values: u64,
    pub exact_max_values: u64,
    pub bloom: i32,
    pub side: i32,
    pub bloom_max_items: u64,
    pub bloom_max_bytes: u64,
    pub bloom_fpp: f64,
    pub bloom_max_fpp: f64,
    pub bloom_bytes: u64,
}

#[repr(C)]
pub struct kg_key_filter {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_key_filter_info {
    pub form: i32,
    pub has_range: i32,
    pub members: i64,
    pub values: i64,
    pub hi: [uint64_t lo[2],; 2],
    pub fhi: double flo,,
    pub empty: int32_t nulls_pass,,
    pub keys: u64,
    pub bloom_bytes: u64,
    pub row_cost_ns: f64,
}

#[repr(C)]
pub struct kg_row_stage {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_frame_header {
    pub id: u64,
    pub reduce: uint32_t stage_attempt, map, attempt, sequence,,
    pub kind: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_join_spec {
    pub join_type: i32,
    pub num_keys: i32,
    pub probe_keys: [i32; 4],
    pub build_keys: [i32; 4],
    pub null_safe: [u8; 4],
    pub no_direct_index: u8,
}

#[repr(C)]
pub struct kg_hash_join {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_expr {
    pub op: i32,
    pub arg: i32,
    pub imm: i64,
}

#[repr(C)]
pub struct kg_join_local {
    _private: [u8; 0],
}

#[repr(C)]
pub struct kg_grace_join {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_grace_join_stats {
    pub spilled: i32,
    pub max_depth: u32,
    pub partitions_joined: u64,
    pub repartitions: u64,
    pub reversed: u64,
    pub spill_bytes: u64,
}

#[repr(C)]
pub struct kg_grace {
    _private: [u8; 0],
}

#[repr(C)]
pub struct kg_server {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_serve_request {
    pub id: i64,
    pub build_keys: *const i64,
    pub build_rows: i64,
    pub table: i32,
    pub _pad: i32,
    pub probe_keys: *const i64,
    pub probe_valid: *const u8,
    pub probe_pay: *const i64,
    pub probe_rows: i64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct kg_tune_config {
    pub load_factor_pct: i32,
    pub scatter_rows: i32,
    pub merge_lanes: i32,
    pub _pad: i32,
}

#[link(name = "kgpu")]
extern "C" {
    pub fn kg_last_error() -> *const c_char;
    pub fn kg_abi_version() -> c_int;
    pub fn kg_ctx_create(device: c_int, hbm_budget_bytes: u64, out: *mut *mut kg_ctx) -> c_int;
    pub fn kg_ctx_destroy(ctx: *mut kg_ctx);
    pub fn kg_ctx_hbm_in_use(ctx: *const kg_ctx) -> u64;
    pub fn kg_ctx_mem_stats(ctx: *mut kg_ctx, out: *mut kg_mem_stats) -> c_int;
    pub fn kg_dev_alloc(ctx: *mut kg_ctx, bytes: u64, out: *mut *mut c_void) -> c_int;
    pub fn kg_dev_free(ctx: *mut kg_ctx, p: *mut c_void) -> c_int;
    pub fn kg_memcpy_h2d(ctx: *mut kg_ctx, dst: *mut c_void, src: *const c_void, bytes: u64) -> c_int;
    pub fn kg_memcpy_d2h(ctx: *mut kg_ctx, dst: *mut c_void, src: *const c_void, bytes: u64) -> c_int;
    pub fn kg_ctx_synchronize(ctx: *mut kg_ctx) -> c_int;
    pub fn kg_batch_num_rows(b: *const kg_batch) -> i64;
    pub fn kg_batch_num_columns(b: *const kg_batch) -> i32;
    pub fn kg_batch_column(b: *const kg_batch, i: i32, out: *mut kg_column) -> c_int;
    pub fn kg_batch_column_to_host(b: *const kg_batch, i: i32, values: *mut c_void, valid_bytes: *mut u8) -> c_int;
    pub fn kg_batch_release(b: *mut kg_batch);
    pub fn kg_column_from_arrow(array: *const ArrowArray, schema: *const ArrowSchema, on_device: i32, out: *mut kg_column) -> c_int;
    pub fn kg_column_from_arrow_device(array: *const ArrowDeviceArray, schema: *const ArrowSchema, out: *mut kg_column) -> c_int;
    pub fn kg_batch_export_device(b: *mut kg_batch, names: *const *const c_char, out: *mut ArrowDeviceArray, schema_out: *mut ArrowSchema) -> c_int;
    pub fn kg_batch_export_host(b: *const kg_batch, names: *const *const c_char, out: *mut ArrowArray, schema_out: *mut ArrowSchema) -> c_int;
    pub fn kg_hash_columns(ctx: *mut kg_ctx, ncols: i32, cols: *const kg_column, nrows: i64, out_hashes: *mut u64, out_flags: u32) -> c_int;
    pub fn kg_partition(ctx: *mut kg_ctx, nkeys: i32, keys: *const kg_column, nrows: i64, partitions: u32, out_pids: *mut u32, out_flags: u32) -> c_int;
    pub fn kg_join_partition(ctx: *mut kg_ctx, nkeys: i32, keys: *const kg_column, nrows: i64, radix_bits: u32, depth: u32, seed: u64, out_hashes: *mut u64, out_parts: *mut u8, out_flags: u32) -> c_int;
    pub fn kg_split_partitions(ctx: *mut kg_ctx, ncols: i32, cols: *const kg_column, nrows: i64, partition_of_row: *const u8, partitions: u32, out_offsets: *mut i64, out: *mut *mut kg_batch) -> c_int;
    pub fn kg_repartition(ctx: *mut kg_ctx, ncols: i32, cols: *const kg_column, nrows: i64, nkeys: i32, key_cols: *const i32, partitions: u32, out_offsets: *mut i64, out: *mut *mut kg_batch) -> c_int;
    pub fn kg_agg_create(ctx: *mut kg_ctx, key_type: *const kg_column, ninputs: i32, input_types: *const kg_column, naggs: i32, aggs: *const kg_agg_spec, flags: u32, expected_groups: u64, out: *mut *mut kg_hash_agg) -> c_int;
    pub fn kg_agg_sink(a: *mut kg_hash_agg, key: *const kg_column, inputs: *const kg_column, nrows: i64) -> c_int;
    pub fn kg_agg_finalize(a: *mut kg_hash_agg) -> c_int;
    pub fn kg_agg_result_type(a: *const kg_hash_agg, i: i32, out_type: *mut kg_column) -> c_int;
    pub fn kg_agg_result(a: *mut kg_hash_agg, out: *mut *mut kg_batch) -> c_int;
    pub fn kg_agg_set_bypass(a: *mut kg_hash_agg, min_rows: u64, max_group_percent: u32, max_table_bytes: u64) -> c_int;
    pub fn kg_agg_partial_next(a: *mut kg_hash_agg, out: *mut *mut kg_batch, done: *mut i32) -> c_int;
    pub fn kg_agg_partial_stats(a: *const kg_hash_agg, flushes: *mut u64, bypassed_rows: *mut u64) -> c_int;
    pub fn kg_agg_set_merge(a: *mut kg_hash_agg, form: i32, size_from_sketch: i32) -> c_int;
    pub fn kg_agg_destroy(a: *mut kg_hash_agg);
    pub fn kg_bloom_create(ctx: *mut kg_ctx, nbytes: u64, out: *mut *mut kg_bloom) -> c_int;
    pub fn kg_bloom_bytes_for(items: u64, fpp: f64) -> u64;
    pub fn kg_bloom_insert(b: *mut kg_bloom, key: *const kg_column, select: *const kg_column, nrows: i64) -> c_int;
    pub fn kg_bloom_expected_fpp(b: *const kg_bloom, items: u64) -> f64;
    pub fn kg_bloom_export(b: *mut kg_bloom, bytes: *mut u8, len: *mut u64) -> c_int;
    pub fn kg_bloom_probe(b: *mut kg_bloom, key: *const kg_column, nrows: i64, out_rows: *mut i32, out_count: *mut i64) -> c_int;
    pub fn kg_bloom_probe_noalloc(b: *mut kg_bloom, key: *const kg_column, nrows: i64, out_rows: *mut i32, out_count: *mut i64) -> c_int;
    pub fn kg_bloom_destroy(b: *mut kg_bloom);
    pub fn kg_spark_bloom_bits_for(items: u64, fpp: f64) -> u64;
    pub fn kg_spark_bloom_hashes_for(items: u64, num_bits: u64) -> i32;
    pub fn kg_spark_bloom_create(ctx: *mut kg_ctx, num_bits: u64, num_hashes: i32, out: *mut *mut kg_bloom) -> c_int;
    pub fn kg_select_range(ctx: *mut kg_ctx, col: *const kg_column, nrows: i64, lo: i64, hi: i64, out_rows: *mut i32, out_count: *mut i64) -> c_int;
    pub fn kg_gather(ctx: *mut kg_ctx, ncols: i32, cols: *const kg_column, rows: *const i32, nrows: i64, out: *mut *mut kg_batch) -> c_int;
    pub fn kg_filter_default_options(o: *mut kg_filter_options);
    pub fn kg_key_filter_create(ctx: *mut kg_ctx, key_type: *const kg_column, opts: *const kg_filter_options, out: *mut *mut kg_key_filter) -> c_int;
    pub fn kg_key_filter_add(f: *mut kg_key_filter, key: *const kg_column, include: *const kg_column, nrows: i64) -> c_int;
    pub fn kg_key_filter_merge_builder(into: *mut kg_key_filter, other: *const kg_key_filter) -> c_int;
    pub fn kg_key_filter_finish(f: *mut kg_key_filter) -> c_int;
    pub fn kg_key_filter_info_get(f: *const kg_key_filter, out: *mut kg_key_filter_info) -> c_int;
    pub fn kg_key_filter_members(f: *const kg_key_filter, out_words: *mut u64, n: *mut i64) -> c_int;
    pub fn kg_key_filter_evaluate(f: *mut kg_key_filter, col: *const kg_column, nrows: i64, out_rows: *mut i32, out_count: *mut i64) -> c_int;
    pub fn kg_key_filter_encode(f: *const kg_key_filter, bytes: *mut u8, len: *mut u64) -> c_int;
    pub fn kg_key_filter_decode(ctx: *mut kg_ctx, bytes: *const u8, len: u64, out: *mut *mut kg_key_filter) -> c_int;
    pub fn kg_key_filter_merge(a: *const kg_key_filter, b: *const kg_key_filter, opts: *const kg_filter_options, out: *mut *mut kg_key_filter) -> c_int;
    pub fn kg_key_filter_destroy(f: *mut kg_key_filter);
    pub fn kg_row_stage_create(ctx: *mut kg_ctx, min_rows: u64, max_pass_percent: u32, out: *mut *mut kg_row_stage) -> c_int;
    pub fn kg_row_stage_set(s: *mut kg_row_stage, slot: i32, f: *const kg_key_filter, column: i32) -> c_int;
    pub fn kg_row_stage_start_partition(s: *mut kg_row_stage) -> c_int;
    pub fn kg_row_stage_apply(s: *mut kg_row_stage, ncols: i32, cols: *const kg_column, nrows: i64, out_rows: *mut i32, out_count: *mut i64) -> c_int;
    pub fn kg_row_stage_stats(s: *const kg_row_stage, slot: i32, tested: *mut u64, dropped: *mut u64, off: *mut i32) -> c_int;
    pub fn kg_row_stage_destroy(s: *mut kg_row_stage);
    pub fn kg_frame_end(h: *const kg_frame_header, out: *mut u8) -> c_int;
    pub fn kg_crc32c(ctx: *mut kg_ctx, data: *const u8, len: u64, out: *mut u32) -> c_int;
    pub fn kg_exchange_frames(ctx: *mut kg_ctx, ncols: i32, cols: *const kg_column, names: *const *const c_char, part_offsets: *const i64, partitions: u32, base: *const kg_frame_header, out: *mut u8, capacity: u64, frame_offsets: *mut u64, total: *mut u64) -> c_int;
    pub fn kg_join_create(ctx: *mut kg_ctx, spec: *const kg_join_spec, expected_build_rows: u64, out: *mut *mut kg_hash_join) -> c_int;
    pub fn kg_join_build_sink(j: *mut kg_hash_join, ncols: i32, cols: *const kg_column, nrows: i64) -> c_int;
    pub fn kg_join_build_finalize(j: *mut kg_hash_join) -> c_int;
    pub fn kg_join_build_rows(j: *const kg_hash_join) -> i64;
    pub fn kg_join_set_filter(j: *mut kg_hash_join, nops: i32, ops: *const kg_expr) -> c_int;
    pub fn kg_join_local_create(j: *mut kg_hash_join, out: *mut *mut kg_join_local) -> c_int;
    pub fn kg_join_local_destroy(l: *mut kg_join_local);
    pub fn kg_join_probe_push(l: *mut kg_join_local, ncols: i32, cols: *const kg_column, nrows: i64, sel_rows: *const i32, nsel: i64) -> c_int;
    pub fn kg_join_probe_next(l: *mut kg_join_local, max_rows: i64, out: *mut *mut kg_batch, done: *mut i32) -> c_int;
    pub fn kg_join_tail_next(j: *mut kg_hash_join, nprobe_cols: i32, probe_types: *const kg_column, max_rows: i64, out: *mut *mut kg_batch, done: *mut i32) -> c_int;
    pub fn kg_join_probe(j: *mut kg_hash_join, ncols: i32, cols: *const kg_column, nrows: i64, out: *mut *mut kg_batch) -> c_int;
    pub fn kg_join_probe_selected(j: *mut kg_hash_join, ncols: i32, cols: *const kg_column, nrows: i64, sel_rows: *const i32, nsel: i64, out: *mut *mut kg_batch) -> c_int;
    pub fn kg_join_destroy(j: *mut kg_hash_join);
    pub fn kg_grace_join_create(ctx: *mut kg_ctx, spec: *const kg_join_spec, table_budget: u64, out: *mut *mut kg_grace_join) -> c_int;
    pub fn kg_grace_join_set_filter(g: *mut kg_grace_join, nops: i32, ops: *const kg_expr) -> c_int;
    pub fn kg_grace_join_build_sink(g: *mut kg_grace_join, ncols: i32, cols: *const kg_column, nrows: i64) -> c_int;
    pub fn kg_grace_join_build_finalize(g: *mut kg_grace_join) -> c_int;
    pub fn kg_grace_join_probe_push(g: *mut kg_grace_join, ncols: i32, cols: *const kg_column, nrows: i64) -> c_int;
    pub fn kg_grace_join_finish(g: *mut kg_grace_join) -> c_int;
    pub fn kg_grace_join_next(g: *mut kg_grace_join, max_rows: i64, out: *mut *mut kg_batch, done: *mut i32) -> c_int;
    pub fn kg_grace_join_stats_get(g: *const kg_grace_join, out: *mut kg_grace_join_stats) -> c_int;
    pub fn kg_grace_join_destroy(g: *mut kg_grace_join);
    pub fn kg_joinagg_create(j: *mut kg_hash_join, ninputs: i32, input_types: *const kg_column, naggs: i32, aggs: *const kg_agg_spec, flags: u32) -> c_int;
    pub fn kg_joinagg_probe(j: *mut kg_hash_join, ncols: i32, cols: *const kg_column, nrows: i64, input_cols: *const i32) -> c_int;
    pub fn kg_joinagg_result(j: *mut kg_hash_join, out: *mut *mut kg_batch) -> c_int;
    pub fn kg_grace_create(ctx: *mut kg_ctx, partitions: u32, table_budget_bytes: u64, flags: u32, out: *mut *mut kg_grace) -> c_int;
    pub fn kg_grace_build_sink(g: *mut kg_grace, key: *const kg_column, pay: *const kg_column, nrows: i64) -> c_int;
    pub fn kg_grace_probe_sink(g: *mut kg_grace, key: *const kg_column, pay: *const kg_column, nrows: i64) -> c_int;
    pub fn kg_grace_finish(g: *mut kg_grace, counts: *mut i64, sums: *mut i64, sum_valid: *mut u8, spilled_bytes: *mut u64) -> c_int;
    pub fn kg_grace_destroy(g: *mut kg_grace);
    pub fn kg_server_create(ctx: *mut kg_ctx, arena_bytes: u64, max_batch: i32, use_graph: i32, out: *mut *mut kg_server) -> c_int;
    pub fn kg_server_add_table(s: *mut kg_server, keys: *const i64, n: i64, table_id: *mut i32) -> c_int;
    pub fn kg_server_submit(s: *mut kg_server, r: *const kg_serve_request) -> c_int;
    pub fn kg_server_progress(s: *mut kg_server, wait: i32, done_ids: *mut i64, max: i32, ndone: *mut i32) -> c_int;
    pub fn kg_server_take_result(s: *mut kg_server, id: i64, out: *mut *mut kg_batch) -> c_int;
    pub fn kg_server_pending(s: *const kg_server) -> i64;
    pub fn kg_server_destroy(s: *mut kg_server);
    pub fn kg_gen_join_build(ctx: *mut kg_ctx, seed: u64, lo: i64, hi: i64, n_total: i64, key_dev: *mut i64, val_dev: *mut i64) -> c_int;
    pub fn kg_gen_join_probe(ctx: *mut kg_ctx, seed: u64, lo: i64, hi: i64, key_range: i64, null_permille: u32, fk_dev: *mut i64, fk_validity_dev: *mut u8, amt_dev: *mut i64) -> c_int;
    pub fn kg_zipf_cdf(n: u64, s: f64, cdf_host: *mut f64) -> c_int;
    pub fn kg_zipf_cdf_device(ctx: *mut kg_ctx, n: u64, s: f64, cdf_dev: *mut f64) -> c_int;
    pub fn kg_gen_agg(ctx: *mut kg_ctx, seed: u64, lo: i64, hi: i64, cdf_dev: *const f64, ndistinct: u64, key_dev: *mut i64, dec_dev: *mut i64, f_dev: *mut f64, i32_dev: *mut i32, k8_dev: *mut int8_t) -> c_int;
    pub fn kg_gen_zipf_keys(ctx: *mut kg_ctx, seed: u64, key_seed: u64, lo: i64, hi: i64, cdf_dev: *const f64, ndistinct: u64, key_dev: *mut i64) -> c_int;
    pub fn kg_prof_enable(on: c_int);
    pub fn kg_prof_reset();
    pub fn kg_prof_get(family: *const c_char, total_ms: *mut f64, launches: *mut i64) -> c_int;
    pub fn kg_tune_get(ctx: *mut kg_ctx, build_rows: i64, out: *mut kg_tune_config) -> c_int;
    pub fn kg_tune_set(ctx: *mut kg_ctx, build_rows: i64, cfg: *const kg_tune_config) -> c_int;
    pub fn kg_autotune_joinagg(ctx: *mut kg_ctx, build_rows: i64, probe_rows: i64, best: *mut kg_tune_config, best_ms: *mut f64, candidates: *mut i32) -> c_int;
    pub fn kg_tune_save(ctx: *mut kg_ctx, path: *const c_char) -> c_int;
    pub fn kg_tune_load(ctx: *mut kg_ctx, path: *const c_char, classes: *mut i32) -> c_int;
    pub fn kg_calib_random_read(ctx: *mut kg_ctx, table_bytes: u64, read_bytes: i32, ops: i64, reps: i32, ms_per_launch: *mut f64) -> c_int;
    pub fn kg_calib_atomic_add(ctx: *mut kg_ctx, addresses: u64, ops: i64, reps: i32, ms_per_launch: *mut f64) -> c_int;
}
