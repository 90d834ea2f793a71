/*
 * kgpu.h -- C ABI of the MI355X (gfx950) execution path for the Krishiv
 * engine's vectorized hash-join / hash-aggregate / runtime-Bloom /
 * hash-repartition operators.
 *
 * The reference engine is Rust and has no FFI for this path; these entry
 * points are what a thin `extern "C"` crate binds, one `extern` block per
 * declaration below, with `#[repr(C)]` mirrors of the structs.  Each
 * declaration names the reference interface it replaces (paths relative to
 * the reference repository root).  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every function returning `int` returns KG_OK (0) or a KG_ERR_* code;
 *    kg_last_error() then holds "<ERROR_CLASS>: <message>" for the calling
 *    thread (the reference's Spark error classes, aggregate/mod.rs:1150-1192).
 *    No C++ exception or abort crosses this boundary.
 *  - Columns are Arrow layout (kg_column): values buffer, optional LSB
 *    validity bitmap (NULL = no NULLs), an element offset applied to both.
 *    KG_COL_DEVICE marks device (HBM) pointers; otherwise host memory, which
 *    is copied to HBM before the call returns (borrowed input, as
 *    Sink::sink borrows its Batch, crates/krishiv-exec/src/operator.rs:346-380).
 *  - An operator handle is used by one thread at a time; distinct handles may
 *    be used concurrently.  Each handle owns one HIP stream (default
 *    priority).  Calls enqueue device work and return; calls that hand data
 *    back to the host synchronize.
 */
#ifndef KGPU_H
#define KGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KG_ABI_VERSION 1

/* ---- status codes (Spark error classes in kg_last_error) -------------- */
#define KG_OK 0
#define KG_ERR_UNSUPPORTED 1  /* KRISHIV_UNSUPPORTED_FEATURE */
#define KG_ERR_OVERFLOW 2     /* ARITHMETIC_OVERFLOW */
#define KG_ERR_OUT_OF_RANGE 3 /* NUMERIC_VALUE_OUT_OF_RANGE.WITH_SUGGESTION */
#define KG_ERR_MEMORY 4       /* UNABLE_TO_ACQUIRE_MEMORY (HBM budget exhausted) */
#define KG_ERR_INTERNAL 5     /* INTERNAL_ERROR (bad argument, broken invariant) */
#define KG_ERR_DEVICE 6       /* INTERNAL_ERROR raised by the HIP runtime */

/* ---- column types (same numbering as the oracle's KGO_T_*) ------------- */
#define KG_T_BOOL 1
#define KG_T_INT8 2
#define KG_T_INT16 3
#define KG_T_INT32 4
#define KG_T_INT64 5
#define KG_T_UINT8 6
#define KG_T_UINT16 7
#define KG_T_UINT32 8
#define KG_T_UINT64 9
#define KG_T_FLOAT32 10
#define KG_T_FLOAT64 11
#define KG_T_DATE32 12
#define KG_T_TIMESTAMP 13 /* any 64-bit temporal (date64, timestamp, duration) */
#define KG_T_DECIMAL64 14
#define KG_T_DECIMAL128 15

#define KG_COL_DEVICE 1u /* values/validity are device pointers */

typedef struct kg_column {
  int32_t type;
  int32_t precision; /* decimals */
  int32_t scale;     /* decimals */
  uint32_t flags;    /* KG_COL_* */
  int64_t length;
  int64_t offset; /* element offset into values and validity */
  const void *values;
  const uint8_t *validity; /* LSB bitmap, NULL = all valid */
} kg_column;

/* ---- errors and context --------------------------------------------- */
const char *kg_last_error(void);
int kg_abi_version(void);

typedef struct kg_ctx kg_ctx;
/* One GPU and an HBM budget for every operator of the context (0 = 90 % of
 * free HBM): the analogue of the memory pool's reservations
 * (crates/krishiv-memory/src/reservation.rs:12,135); past it operators fail
 * with KG_ERR_MEMORY. */
int kg_ctx_create(int device, uint64_t hbm_budget_bytes, kg_ctx **out);
void kg_ctx_destroy(kg_ctx *ctx);
/* Bytes of HBM currently held by the context's operators. */
uint64_t kg_ctx_hbm_in_use(const kg_ctx *ctx);
/* The context's device allocator: freed blocks are kept by size class (at
 * most 12.5 % over a request above 2 MB) and reused once the work enqueued
 * before the free has completed; hipFree, which waits for the whole device,
 * runs only past a quarter of the device's memory cached or when hipMalloc fails
 * (then the whole cache is released).  The analogue of the reference's
 * memory-pool statistics (PoolSnapshot, crates/krishiv-memory/src/pool.rs:826). */
typedef struct kg_mem_stats {
  uint64_t in_use;      /* bytes of live allocations (size classes) */
  uint64_t cached;      /* freed bytes kept for reuse */
  uint64_t hip_mallocs; /* hipMalloc calls */
  uint64_t hip_frees;   /* hipFree calls */
  uint64_t cache_hits;  /* allocations served from the cache */
  uint64_t releases;    /* whole-cache releases after a failed hipMalloc */
} kg_mem_stats;
int kg_ctx_mem_stats(kg_ctx *ctx, kg_mem_stats *out);

/* Plain device memory, for callers that keep inputs resident in HBM. */
int kg_dev_alloc(kg_ctx *ctx, uint64_t bytes, void **out);
int kg_dev_free(kg_ctx *ctx, void *p);
int kg_memcpy_h2d(kg_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int kg_memcpy_d2h(kg_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int kg_ctx_synchronize(kg_ctx *ctx);

/* ---- device batches (operator output) ----------------------------------
 * A device-resident record batch owned by the library: a join's output,
 * an aggregate's result, a repartition's partitions.  Its columns can be
 * passed on to another operator as KG_COL_DEVICE columns without a copy. */
typedef struct kg_batch kg_batch;
int64_t kg_batch_num_rows(const kg_batch *b);
int32_t kg_batch_num_columns(const kg_batch *b);
/* A device view of column i (valid while the batch lives). */
int kg_batch_column(const kg_batch *b, int32_t i, kg_column *out);
/* Copies column i to host buffers: values (width x rows bytes; BOOL as one
 * byte per row), validity as one byte per row (may be NULL). */
int kg_batch_column_to_host(const kg_batch *b, int32_t i, void *values, uint8_t *valid_bytes);
/* Releases a batch.  Its device memory is reused only after the library's
 * own work that reads or writes it has completed (stream order plus the
 * reads of every operator call made before the release); device work the
 * caller enqueued itself on the columns must be complete first. */
void kg_batch_release(kg_batch *b);

/* ---- Arrow C Data Interface at the boundary ------------------------------
 * The engine's batches are Arrow record batches (krishiv-types Batch over
 * arrow_array); the structures below are the Arrow C Data Interface and C
 * Device Data Interface as the Arrow specification defines them (the same
 * layout everywhere, so a caller's own definitions are interchangeable). */
#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE
#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4
struct ArrowSchema {
  const char *format;
  const char *name;
  const char *metadata;
  int64_t flags;
  int64_t n_children;
  struct ArrowSchema **children;
  struct ArrowSchema *dictionary;
  void (*release)(struct ArrowSchema *);
  void *private_data;
};
struct ArrowArray {
  int64_t length;
  int64_t null_count;
  int64_t offset;
  int64_t n_buffers;
  int64_t n_children;
  const void **buffers;
  struct ArrowArray **children;
  struct ArrowArray *dictionary;
  void (*release)(struct ArrowArray *);
  void *private_data;
};
#endif
#ifndef ARROW_C_DEVICE_DATA_INTERFACE
#define ARROW_C_DEVICE_DATA_INTERFACE
typedef int32_t ArrowDeviceType;
#define ARROW_DEVICE_CPU 1
#define ARROW_DEVICE_ROCM 10
#define ARROW_DEVICE_ROCM_HOST 11
struct ArrowDeviceArray {
  struct ArrowArray array;
  int64_t device_id;
  ArrowDeviceType device_type;
  void *sync_event;
  int64_t reserved[3];
};
#endif

/* A kg_column viewing one primitive Arrow array (no copy): formats b c s i l
 * C S I L f g tdD tsu:<tz> d:p,s[,128] d:p,s,64 (bool, ints, uints, floats,
 * date32, microsecond timestamps, decimal128/64); other formats are
 * KG_ERR_UNSUPPORTED.  on_device: the buffers are device pointers.  The view
 * is valid while the array is. */
int kg_column_from_arrow(const struct ArrowArray *array, const struct ArrowSchema *schema,
                         int32_t on_device, kg_column *out);
/* The same of an ArrowDeviceArray: ARROW_DEVICE_ROCM buffers are device
 * columns (its sync_event, a hipEvent_t *, is waited for first), CPU and
 * ROCM_HOST ones host columns. */
int kg_column_from_arrow_device(const struct ArrowDeviceArray *array,
                                const struct ArrowSchema *schema, kg_column *out);
/* A batch as an Arrow struct array ("+s") of its columns, named names[c]
 * (NULL: "c0", "c1", ...), without a copy: the batch is moved into the
 * export (the caller no longer releases it; the array's release does).
 * device_type ARROW_DEVICE_ROCM, device_id the context's device, sync_event
 * NULL (the batch is complete when exported). */
int kg_batch_export_device(kg_batch *b, const char *const *names, struct ArrowDeviceArray *out,
                           struct ArrowSchema *schema_out);
/* A batch copied to host memory as an Arrow struct array (the batch is left
 * as it was). */
int kg_batch_export_host(const kg_batch *b, const char *const *names, struct ArrowArray *out,
                         struct ArrowSchema *schema_out);

/* ---- routing hash and hash repartition (exchange) ---------------------
 * kg_hash_columns replaces routing::hash_columns
 * (crates/krishiv-kernels/src/hash/routing.rs:258-282): H_kg v1 of each row
 * (XXH64 seed 0 of the normalized key bytes, folded over several columns).
 * out_hashes: device or host (out_flags & KG_COL_DEVICE). */
int kg_hash_columns(kg_ctx *ctx, int32_t ncols, const kg_column *cols, int64_t nrows,
                    uint64_t *out_hashes, uint32_t out_flags);
/* exchange::route for Partitioning::Hash (crates/krishiv-exec/src/exchange/mod.rs:159-212,
 * physical/mod.rs:248-262): partition_of(H_kg(keys), partitions) per row. */
int kg_partition(kg_ctx *ctx, int32_t nkeys, const kg_column *keys, int64_t nrows,
                 uint32_t partitions, uint32_t *out_pids, uint32_t out_flags);
/* The join's row-layout hash h of each row's key columns (RowLayout::encode's
 * stored hash, crates/krishiv-exec/src/rows/layout.rs:487-553,787-797,871-904)
 * and its grace partition (Hashing::partition, crates/krishiv-exec/src/join/
 * keys.rs:118-128): depth 0 takes the top radix_bits (clamped to 3..6) of h,
 * depth d >= 1 the 3 bits at 48 - 3 min(d, 6) of g = finalize(h, seed).
 * out_hashes (may be NULL) and out_parts (one byte a row): device or host
 * (out_flags & KG_COL_DEVICE). */
int kg_join_partition(kg_ctx *ctx, int32_t nkeys, const kg_column *keys, int64_t nrows,
                      uint32_t radix_bits, uint32_t depth, uint64_t seed, uint64_t *out_hashes,
                      uint8_t *out_parts, uint32_t out_flags);
/* The split alone (exchange/mod.rs:480-501) with the partition of each row
 * given, one byte a row in device memory (< partitions <= 256): the rows of
 * `cols` regrouped by partition, stable within a partition; out_offsets
 * (host, partitions+1 entries). */
int kg_split_partitions(kg_ctx *ctx, int32_t ncols, const kg_column *cols, int64_t nrows,
                        const uint8_t *partition_of_row, uint32_t partitions,
                        int64_t *out_offsets, kg_batch **out);
/* ExchangeSink::sink's route + split (exchange/mod.rs:447-501): the rows of
 * `cols` regrouped by partition, stable within a partition.  The output batch
 * holds the permuted columns; out_offsets (host, partitions+1 entries) gives
 * each partition's row range. */
int kg_repartition(kg_ctx *ctx, int32_t ncols, const kg_column *cols, int64_t nrows,
                   int32_t nkeys, const int32_t *key_cols, uint32_t partitions,
                   int64_t *out_offsets, kg_batch **out);

/* ---- hash aggregation ---------------------------------------------------
 * HashAggregate (crates/krishiv-exec/src/aggregate/sink.rs:222,1459-1585) and
 * PartialAggregate (aggregate/partial.rs): a Sink (kg_agg_sink*, called per
 * batch) then finalize, then a Source (kg_agg_result).  Spark semantics and
 * result types as aggregate/mod.rs:588-775. */
#define KG_AGG_COUNT_STAR 1
#define KG_AGG_COUNT 2
#define KG_AGG_SUM 3
#define KG_AGG_MIN 4
#define KG_AGG_MAX 5
#define KG_AGG_AVG 6

#define KG_AGG_ANSI 1u /* spark.sql.ansi.enabled: overflow raises */
/* PartialAggregate (aggregate/partial.rs): the result holds each aggregate's
 * partial-state columns (state_fields, aggregate/mod.rs:747-775) instead of
 * its final value: COUNT -> count; SUM int/float -> sum; SUM decimal ->
 * (sum, nonempty); AVG -> (sum, count); MIN/MAX -> value. */
#define KG_AGG_PARTIAL 2u
/* The final phase of a two-phase aggregate: each sink call's inputs are the
 * partial-state columns above, in aggregate order (input_types and the
 * aggregate list stay the original argument types and functions). */
#define KG_AGG_MERGE 4u

typedef struct kg_agg_spec {
  int32_t func;  /* KG_AGG_* */
  int32_t input; /* input column index (ignored by COUNT_STAR) */
} kg_agg_spec;

typedef struct kg_hash_agg kg_hash_agg;
/* key_type: the grouping column's type (type/precision/scale used; NULL for a
 * global aggregate).  input_types: the value columns' types.
 * expected_groups sizes the table (it grows if exceeded). */
int kg_agg_create(kg_ctx *ctx, const kg_column *key_type, int32_t ninputs,
                  const kg_column *input_types, int32_t naggs, const kg_agg_spec *aggs,
                  uint32_t flags, uint64_t expected_groups, kg_hash_agg **out);
/* Sink::sink of raw rows.  Host batches of fewer than 512K rows (at up to
 * 64 B a row) are copied to pinned staging and gathered in HBM (64 times the
 * first batch's rows, growing to 4M rows or 256 MB), and aggregated when the
 * next batch does not fit, a device batch arrives, or at finalize: an error
 * of that work
 * (NeedMemory, overflow) is returned by that later call.  NeedMemory leaves
 * the gathered rows in place and the call's own batch not taken, so the
 * same batch retried is counted once.  Without a bypass policy only. */
int kg_agg_sink(kg_hash_agg *a, const kg_column *key, const kg_column *inputs, int64_t nrows);
/* Sink::finalize: checks overflow, fixes the group count. */
int kg_agg_finalize(kg_hash_agg *a);
/* Result type of aggregate i (type, precision, scale). */
int kg_agg_result_type(const kg_hash_agg *a, int32_t i, kg_column *out_type);
/* Source: the final values as a device batch [key, agg 0, agg 1, ...]. */
int kg_agg_result(kg_hash_agg *a, kg_batch **out);
/* BypassPolicy of a PartialAggregate (aggregate/partial.rs:20-50, defaults
 * 100000 rows, 80 percent, 16 MiB): after each batch is folded, once
 * min_rows rows are seen and the groups (flushed and held) per 100 rows reach
 * max_group_percent, the table is flushed and every later batch passes
 * through as one partial state per row (AggTable::states_of_rows); otherwise
 * a table whose records pass max_table_bytes is flushed (partial.rs:217-262).
 * Flushed and passed-through state batches are taken with
 * kg_agg_partial_next; finalize + result give what the table holds at the
 * end.  KG_AGG_PARTIAL aggregates with a key only, set before the first
 * sink. */
int kg_agg_set_bypass(kg_hash_agg *a, uint64_t min_rows, uint32_t max_group_percent,
                      uint64_t max_table_bytes);
/* The next emitted partial-state batch (*out NULL if none); *done = 1 when no
 * more are queued.  The caller owns the batch. */
int kg_agg_partial_next(kg_hash_agg *a, kg_batch **out, int32_t *done);
/* PartialAggregate's metrics (partial.rs:289-296): table flushes and rows
 * passed through. */
int kg_agg_partial_stats(const kg_hash_agg *a, uint64_t *flushes, uint64_t *bypassed_rows);
/* How a drain merges the spilled rows and records into the table (set
 * before the first sink).  form: KG_AGG_MERGE_AUTO (each window of the table
 * merged in LDS by its own workgroup when the spill's partitions hold many
 * inputs a window, else by key), KG_AGG_MERGE_WINDOWS or KG_AGG_MERGE_BY_KEY.
 * size_from_sketch 0: the table grows from expected_groups and when full
 * only, not from the HyperLogLog estimate of the spilled keys.  No
 * counterpart in the reference (its row tables are CPU hash indexes): for
 * tuning and for tests of each form. */
#define KG_AGG_MERGE_AUTO 0
#define KG_AGG_MERGE_WINDOWS 1
#define KG_AGG_MERGE_BY_KEY 2
int kg_agg_set_merge(kg_hash_agg *a, int32_t form, int32_t size_from_sketch);
void kg_agg_destroy(kg_hash_agg *a);

/* ---- runtime Bloom filter (split-block, Parquet layout) ----------------
 * RuntimeBloom (crates/krishiv-io/src/runtime_filter/bloom.rs:83-320) built
 * as KeyFilterBuilder::finish_with/bloom() (runtime_filter/filter.rs:533-624)
 * does: XXH64 seed 0 of the normalized key bytes, `bytes` = bytes_for(items,
 * fpp) (or the planned size), kept only if expected_fpp(items) <= max_fpp. */
typedef struct kg_bloom kg_bloom;
int kg_bloom_create(kg_ctx *ctx, uint64_t nbytes, kg_bloom **out);
/* bytes_for (bloom.rs:96-110). */
uint64_t kg_bloom_bytes_for(uint64_t items, double fpp);
/* Inserts the non-NULL keys of `key` (rows where `select` is set, if given:
 * the build side's kept rows). */
int kg_bloom_insert(kg_bloom *b, const kg_column *key, const kg_column *select, int64_t nrows);
/* expected_fpp (bloom.rs:270-308) after `items` insertions. */
double kg_bloom_expected_fpp(const kg_bloom *b, uint64_t items);
/* The Parquet-layout bit array (bloom.rs:311-320) to host memory. */
int kg_bloom_export(kg_bloom *b, uint8_t *bytes, uint64_t *len);
/* check_batch (bloom.rs:211-247) fused with the scan row stage's selection
 * (operators/runtime_filter.rs): writes the ids of rows that may match
 * (NULL keys never pass) to out_rows (device) and their count to
 * *out_count (host). */
int kg_bloom_probe(kg_bloom *b, const kg_column *key, int64_t nrows, int32_t *out_rows,
                   int64_t *out_count);
/* The same probe with no device allocation: no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The key column must be
 * device-resident (KG_COL_DEVICE) and out_rows must hold nrows entries; at
 * most 2^31 - 1 rows a call.  Its scratch (the survivor bitmap and tile
 * counts of 16M-row chunks, 2.03 MB) is one file-scope device array per GPU,
 * held until the process ends: calls that overlap on one GPU race, so the
 * caller serializes them.  Results equal kg_bloom_probe's. */
int kg_bloom_probe_noalloc(kg_bloom *b, const kg_column *key, int64_t nrows, int32_t *out_rows,
                           int64_t *out_count);
void kg_bloom_destroy(kg_bloom *b);

/* Spark-layout Bloom filter (optional).  The reference holds only the hash
 * Spark's filter uses, Murmur3 x86_32 (crates/krishiv-kernels/src/hash/
 * murmur3.rs:9-52); the layout is Spark's sketch BloomFilterImpl (version 1):
 * a long[] of ceil(num_bits / 64) words, putLong(item) sets bits
 * (h1 + i*h2, ~ when negative) mod bitSize for i = 1..num_hashes with
 * h1 = murmur3(item, 0), h2 = murmur3(item, h1); the item of a key is
 * Spark's xxhash64(key) (seed 42), as InjectRuntimeFilter builds it.  Keys:
 * int8/16/32, date32 (hashed as int), int64, timestamp, decimal64 (as long).
 * kg_bloom_insert, kg_bloom_probe, kg_bloom_export (the words, little-endian)
 * and kg_bloom_destroy apply to it. */
uint64_t kg_spark_bloom_bits_for(uint64_t items, double fpp);   /* optimalNumOfBits */
int32_t kg_spark_bloom_hashes_for(uint64_t items, uint64_t num_bits); /* optimalNumOfHashFunctions */
int kg_spark_bloom_create(kg_ctx *ctx, uint64_t num_bits, int32_t num_hashes, kg_bloom **out);

/* The range form of a runtime filter and the scan's row stage
 * (runtime_filter/member.rs Range, operators/runtime_filter.rs,
 * operators/select.rs): ids of the rows whose value lies in [lo, hi]
 * (integer-like columns, compared as int64; NULLs never pass), in input
 * order, to out_rows (device); *out_count (host). */
int kg_select_range(kg_ctx *ctx, const kg_column *col, int64_t nrows, int64_t lo, int64_t hi,
                    int32_t *out_rows, int64_t *out_count);
/* The selected rows (device row ids) of `cols` as a device batch: the take of
 * a selection vector (operators/select.rs). */
int kg_gather(kg_ctx *ctx, int32_t ncols, const kg_column *cols, const int32_t *rows,
              int64_t nrows, kg_batch **out);

/* ---- runtime KeyFilter: forms, codec, merge, row stage -----------------
 * KeyFilterBuilder / KeyFilter (crates/krishiv-io/src/runtime_filter/
 * filter.rs:180-615): a producer's keys (fixed-width key types: bool, ints,
 * date32, timestamp, float32/64, decimal64/128) collected on the device into
 * the min/max range, the distinct keys while few, and the row hashes for a
 * Bloom filter; finish picks the row form (member.rs:1-20): the range for
 * dense integers, a small set (<= 8), an exact set (<= exact_max_values; a
 * bit per value for short integer ranges), else a split-block Bloom filter
 * kept only if its expected fpp <= bloom_max_fpp. */
#define KG_FORM_RANGE 0
#define KG_FORM_SMALL 1
#define KG_FORM_EXACT 2
#define KG_FORM_BLOOM 3
#define KG_FORM_DENSE 4 /* an exact set held as a bit per value */
#define KG_SIDE_BOTH 0  /* RangeSide (filter.rs:76-90) */
#define KG_SIDE_UPPER 1
#define KG_SIDE_LOWER 2

/* KeyFilterOptions (filter.rs:28-107). */
typedef struct kg_filter_options {
  uint64_t max_values;       /* statistics IN list cap (default 50) */
  uint64_t exact_max_values; /* exact set cap (4096; at most 65536 here) */
  int32_t bloom;             /* collect hashes for a Bloom filter (1) */
  int32_t side;              /* KG_SIDE_* */
  uint64_t bloom_max_items;  /* 4,000,000 */
  uint64_t bloom_max_bytes;  /* 8 MiB */
  double bloom_fpp;          /* 0.01 */
  double bloom_max_fpp;      /* 0.05 */
  uint64_t bloom_bytes;      /* the planned size; 0: sized for the keys */
} kg_filter_options;
void kg_filter_default_options(kg_filter_options *o);

typedef struct kg_key_filter kg_key_filter;
/* KeyFilterBuilder::new for keys of key_type (type, precision, scale). */
int kg_key_filter_create(kg_ctx *ctx, const kg_column *key_type, const kg_filter_options *opts,
                         kg_key_filter **out);
/* KeyFilterBuilder::add_hashed (filter.rs:225-284): the keys of rows where
 * `include` (optional; rows whose value is non-zero and valid) is set; a NULL
 * key lets NULLs pass.  The column must be of the key type. */
int kg_key_filter_add(kg_key_filter *f, const kg_column *key, const kg_column *include,
                      int64_t nrows);
/* KeyFilterBuilder::merge (filter.rs:478-510): `other`'s keys into `into`
 * (both unfinished builders of one key type; `other` is left as it was). */
int kg_key_filter_merge_builder(kg_key_filter *into, const kg_key_filter *other);
/* KeyFilterBuilder::finish (filter.rs:526-602): the published filter. */
int kg_key_filter_finish(kg_key_filter *f);

typedef struct kg_key_filter_info {
  int32_t form; /* KG_FORM_* */
  int32_t has_range;
  int64_t members; /* set forms: the members; otherwise 0 */
  int64_t values;  /* the statistics IN list's length, -1 without one */
  uint64_t lo[2], hi[2]; /* integer / decimal range (low, high words) */
  double flo, fhi;       /* float range */
  int32_t nulls_pass, empty;
  uint64_t keys;        /* non-null keys stored, duplicates included */
  uint64_t bloom_bytes; /* the Bloom form's size */
  double row_cost_ns;   /* Member::row_cost_ns (member.rs:410-418) */
} kg_key_filter_info;
int kg_key_filter_info_get(const kg_key_filter *f, kg_key_filter_info *out);
/* A set form's members as (low, high) words, ascending; *n in: capacity in
 * members, out: the count (out_words NULL: the count only). */
int kg_key_filter_members(const kg_key_filter *f, uint64_t *out_words, int64_t *n);
/* KeyFilter::evaluate (filter.rs:764-797, member.rs:439-560): ids of the
 * rows of `col` that may match, in order (device out_rows, host count).
 * The column may be of a type that widens to the key type (key.rs:75-103:
 * narrower ints, ints or float32 to float64, decimals to a wider scale);
 * another type is KG_ERR_UNSUPPORTED (the filter does not apply). */
int kg_key_filter_evaluate(kg_key_filter *f, const kg_column *col, int64_t nrows,
                           int32_t *out_rows, int64_t *out_count);
/* KeyFilter::encode (runtime_filter/codec.rs:104-180): the versioned bytes
 * (format 1, row hash XXH64 seed 0, normalization 1).  bytes NULL: *len gets
 * the size; else *len in: capacity, out: the size written.  Exact sets are
 * written in ascending order. */
int kg_key_filter_encode(const kg_key_filter *f, uint8_t *bytes, uint64_t *len);
/* KeyFilter::decode (codec.rs:186-246): bytes of another version, type or
 * layout are refused (KG_ERR_INTERNAL); a top-N bound is KG_ERR_UNSUPPORTED. */
int kg_key_filter_decode(kg_ctx *ctx, const uint8_t *bytes, uint64_t len, kg_key_filter **out);
/* KeyFilter::merge (merge.rs:35-169): the filter of the union of the keys of
 * two published filters; equal-size planned Bloom filters OR on the device. */
int kg_key_filter_merge(const kg_key_filter *a, const kg_key_filter *b,
                        const kg_filter_options *opts, kg_key_filter **out);
void kg_key_filter_destroy(kg_key_filter *f);

/* RowStage (crates/krishiv-exec/src/operators/runtime_filter.rs:240-395): a
 * scan driver's runtime filters over its batches in one fused pass, cheapest
 * form first; a filter passing more than min(max_pass_percent / 100, 1 -
 * row_cost / 20 ns) of a window of min_rows rows it tested is switched off
 * until start_partition or a new filter in its slot. */
#define KG_ROW_STAGE_SLOTS 8
typedef struct kg_row_stage kg_row_stage;
int kg_row_stage_create(kg_ctx *ctx, uint64_t min_rows, uint32_t max_pass_percent,
                        kg_row_stage **out);
/* Slot `slot` filters column `column` of the batches with `f` (a published
 * filter the stage does not own; NULL: pending, applies nothing). */
int kg_row_stage_set(kg_row_stage *s, int32_t slot, const kg_key_filter *f, int32_t column);
int kg_row_stage_start_partition(kg_row_stage *s);
/* The rows some producer may keep (device out_rows, in order); *out_count =
 * -1 when nothing is filtered (every row passes; out_rows untouched). */
int kg_row_stage_apply(kg_row_stage *s, int32_t ncols, const kg_column *cols, int64_t nrows,
                       int32_t *out_rows, int64_t *out_count);
/* Rows tested and dropped by a slot's filter, and whether it is off. */
int kg_row_stage_stats(const kg_row_stage *s, int32_t slot, uint64_t *tested, uint64_t *dropped,
                       int32_t *off);
void kg_row_stage_destroy(kg_row_stage *s);

/* ---- exchange frames ------------------------------------------------------
 * The unit of local shuffle files and the wire (crates/krishiv-shuffle/src/
 * frame.rs:1-26): a 44-byte little-endian header -- magic "KXFR", format
 * version 1, kind, compression 0, exchange id, stage attempt, map partition,
 * task attempt, sequence, reduce partition, body length, CRC32C of the body
 * -- then the body: for a Data frame an Arrow IPC stream of one batch
 * (frame.rs:205-221), for an End frame nothing. */
#define KG_FRAME_HEADER_BYTES 44
#define KG_FRAME_DATA 1
#define KG_FRAME_END 2
#define KG_FRAME_VALUE 3
typedef struct kg_frame_header {
  uint64_t id; /* the exchange (or value) id */
  uint32_t stage_attempt, map, attempt, sequence, reduce;
  uint32_t kind; /* KG_FRAME_* (kg_exchange_frames writes Data frames) */
} kg_frame_header;
/* frame::header_bytes of an End frame (empty body, CRC 0): 44 bytes to out
 * (host). */
int kg_frame_end(const kg_frame_header *h, uint8_t *out);
/* frame::crc32c (frame.rs:251-274) of len bytes of device memory. */
int kg_crc32c(kg_ctx *ctx, const uint8_t *data, uint64_t len, uint32_t *out);
/* ExchangeSink's Data frames of one split batch (exchange/mod.rs:447-501 then
 * frame::encode, frame.rs:118-150): `cols` holds the rows grouped by reduce
 * partition (kg_repartition's output), partition p's rows at
 * [part_offsets[p], part_offsets[p+1]).  Each non-empty partition's rows
 * leave as one frame: the header of `base` with reduce = p and kind Data,
 * and an Arrow IPC stream (schema, one record batch, end marker) whose
 * fields are named names[c] (NULL: "c0", "c1", ...), all nullable.  The
 * frames are written back to back into `out` (device), partition p's at
 * [frame_offsets[p], frame_offsets[p+1]) (host array of partitions + 1; an
 * empty partition has no frame); *total gets the bytes.  out NULL: the sizes
 * only. */
int kg_exchange_frames(kg_ctx *ctx, int32_t ncols, const kg_column *cols,
                       const char *const *names, const int64_t *part_offsets, uint32_t partitions,
                       const kg_frame_header *base, uint8_t *out, uint64_t capacity,
                       uint64_t *frame_offsets, uint64_t *total);

/* ---- hash join ----------------------------------------------------------
 * HashJoin (crates/krishiv-exec/src/join/mod.rs:1014-1211): the build side is
 * a Sink (kg_join_build_sink per batch, then kg_join_build_finalize,
 * join/build.rs:1225-1695); the probe side an Operator (join/probe.rs:784-948)
 * driven per driver through a local handle: kg_join_probe_push takes a batch,
 * kg_join_probe_next returns its output in batches of at most max_rows rows
 * (Cursor::next / pairs, probe.rs:266-501).  Once every probe batch is done,
 * kg_join_tail_next returns the build rows the join type preserves
 * (emit_build, probe.rs:611-632).  Spark's join types and NULL semantics
 * (join/mod.rs:1-60,217-293):
 *   INNER        pairs                       [probe cols, build cols]
 *   LEFT_OUTER   pairs + unmatched probe rows (NULL build cols)
 *   LEFT_SEMI    probe rows with a match     [probe cols]
 *   LEFT_ANTI    probe rows without a match  [probe cols]
 *   NULL_AWARE_ANTI  NOT IN (one `=` key, no residual filter): an empty build
 *                keeps every probe row; a NULL build key keeps none; else a
 *                probe row with a non-NULL key and no match [probe cols]
 *   EXISTENCE    every probe row once        [probe cols, exists BOOL non-null]
 *   RIGHT_SEMI   build rows with a match (tail)           [build cols]
 *   RIGHT_ANTI   build rows without one, NULL keys too (tail) [build cols]
 *   RIGHT_OUTER  pairs + unmatched build rows (tail; NULL probe cols)
 *   FULL_OUTER   pairs + unmatched probe rows + unmatched build rows (tail)
 * A residual filter (kg_join_set_filter) decides which key-equal pairs
 * match, before any of the above. */
#define KG_JOIN_INNER 0
#define KG_JOIN_LEFT_OUTER 1
#define KG_JOIN_LEFT_SEMI 2
#define KG_JOIN_LEFT_ANTI 3
#define KG_JOIN_NULL_AWARE_ANTI 4
#define KG_JOIN_EXISTENCE 5
#define KG_JOIN_RIGHT_SEMI 6
#define KG_JOIN_RIGHT_ANTI 7
#define KG_JOIN_RIGHT_OUTER 8
#define KG_JOIN_FULL_OUTER 9

typedef struct kg_join_spec {
  int32_t join_type;
  int32_t num_keys;       /* 1 or 2 */
  int32_t probe_keys[4];  /* key column indices of the probe batch */
  int32_t build_keys[4];  /* key column indices of the build batch */
  uint8_t null_safe[4];   /* `<=>` keys */
  /* 1: never probe through a direct index (krishiv.exec.join.directIndex =
   * false, crates/krishiv-common/src/config/registry_join.rs:26).  With 0,
   * a build of at most 4,096 rows on one integer `=` key whose values span
   * at most 16,384 slots and 4 a row is probed by key - min
   * (crates/krishiv-exec/src/join/direct.rs:1-160). */
  uint8_t no_direct_index;
} kg_join_spec;

typedef struct kg_hash_join kg_hash_join;
int kg_join_create(kg_ctx *ctx, const kg_join_spec *spec, uint64_t expected_build_rows,
                   kg_hash_join **out);
/* BuildSink::sink.  Host batches of fewer than 512K rows are gathered as
 * kg_agg_sink gathers them (taken at finalize or when the next batch does
 * not fit; NeedMemory then leaves them gathered and the new batch untaken);
 * kg_join_build_rows counts them. */
int kg_join_build_sink(kg_hash_join *j, int32_t ncols, const kg_column *cols, int64_t nrows);
int kg_join_build_finalize(kg_hash_join *j);
/* Rows sunk into the build side (all of them; those with a NULL `=` key
 * are stored but never indexed). */
int64_t kg_join_build_rows(const kg_hash_join *j);

/* The residual filter (JoinSpec::filter, join/mod.rs:327-350; evaluated per
 * key-equal pair, probe.rs:686+), as a postfix program over the pair's
 * probe and build columns with Spark's three-valued logic: a pair matches
 * when the program yields TRUE (NULL and FALSE do not).  Integer-like
 * columns (integers, dates, timestamps, decimal64 as unscaled values) read
 * as int64; + - * wrap at 64 bits; % takes the dividend's sign and is NULL
 * for a zero divisor; comparisons with NULL are NULL; AND/OR are Kleene's.
 * At most 32 ops, 8 distinct columns, a stack depth of 16.  Set before the
 * first probe; not with NULL_AWARE_ANTI or the fused aggregate. */
#define KG_X_PROBE 1 /* push probe column `arg` */
#define KG_X_BUILD 2 /* push build column `arg` */
#define KG_X_CONST 3 /* push `imm` */
#define KG_X_ADD 4
#define KG_X_SUB 5
#define KG_X_MUL 6
#define KG_X_MOD 7
#define KG_X_EQ 8
#define KG_X_NE 9
#define KG_X_LT 10
#define KG_X_LE 11
#define KG_X_GT 12
#define KG_X_GE 13
#define KG_X_AND 14
#define KG_X_OR 15
#define KG_X_NOT 16
#define KG_X_IS_NULL 17
typedef struct kg_expr {
  int32_t op;
  int32_t arg;
  int64_t imm;
} kg_expr;
int kg_join_set_filter(kg_hash_join *j, int32_t nops, const kg_expr *ops);

/* A driver's probe state (the Operator's per-driver LocalState,
 * crates/krishiv-exec/src/operator.rs:41-44): its own HIP stream and
 * buffers over the shared, finalized build.  Distinct locals of one join may
 * be used from different threads at once. */
typedef struct kg_join_local kg_join_local;
int kg_join_local_create(kg_hash_join *j, kg_join_local **out);
void kg_join_local_destroy(kg_join_local *l);
/* Operator::execute's input: a probe batch (all rows, or the rows of a
 * selection vector: device row ids, the runtime filter's survivors, the
 * fused RowSelect of join/mod.rs:641-657).  Host columns are copied to HBM
 * before the call returns; device columns and sel_rows must stay valid
 * until kg_join_probe_next reports done.  The previous batch must be done. */
int kg_join_probe_push(kg_join_local *l, int32_t ncols, const kg_column *cols, int64_t nrows,
                       const int32_t *sel_rows, int64_t nsel);
/* The pushed batch's next output rows, at most max_rows (the bounded
 * execute of operator.rs:7-13); *done = 1 once the batch's output is all
 * returned (the last call may return a zero-row batch). */
int kg_join_probe_next(kg_join_local *l, int64_t max_rows, kg_batch **out, int32_t *done);
/* After every probe batch of every local: the build rows the join type
 * returns by their matches (right semi: matched; right anti, right outer,
 * full outer: unmatched, NULL `=` keys included), at most max_rows a call;
 * the outer joins pad them with NULL probe columns of the given types.
 * Other join types return done at once. */
int kg_join_tail_next(kg_hash_join *j, int32_t nprobe_cols, const kg_column *probe_types,
                      int64_t max_rows, kg_batch **out, int32_t *done);
/* One batch's whole output at once on the join's own local (push + next
 * without a bound). */
int kg_join_probe(kg_hash_join *j, int32_t ncols, const kg_column *cols, int64_t nrows,
                  kg_batch **out);
int kg_join_probe_selected(kg_hash_join *j, int32_t ncols, const kg_column *cols, int64_t nrows,
                           const int32_t *sel_rows, int64_t nsel, kg_batch **out);
void kg_join_destroy(kg_hash_join *j);

/* ---- grace hash join ------------------------------------------------------
 * GraceHashJoin (crates/krishiv-exec/src/join/grace.rs:1-20,828-1031): a
 * HashJoin (same spec, join types, residual filter and output as
 * kg_hash_join) whose build rows may pass `table_budget` bytes of HBM.
 * Under the budget it is the in-memory join.  Past it every build row, and
 * then every probe row, is partitioned on the top 6 bits of the row hash h
 * (Hashing::partition level 0, keys.rs:118-128) into pinned host memory;
 * after the last probe batch (kg_grace_join_finish) each partition is
 * joined in HBM -- a partition whose build rows do not fit is split 8 ways
 * on the next 3 bits of g (levels 1..6, MAX_DEPTH, grace.rs:1481-1522),
 * and a semi or anti partition with at least 8 build rows a probe row is
 * joined with roles reversed (grace.rs:497-548).  Output rows are those of
 * the in-memory join (order within the output aside). */
typedef struct kg_grace_join kg_grace_join;
int kg_grace_join_create(kg_ctx *ctx, const kg_join_spec *spec, uint64_t table_budget,
                         kg_grace_join **out);
int kg_grace_join_set_filter(kg_grace_join *g, int32_t nops, const kg_expr *ops);
int kg_grace_join_build_sink(kg_grace_join *g, int32_t ncols, const kg_column *cols,
                             int64_t nrows);
int kg_grace_join_build_finalize(kg_grace_join *g);
/* A probe batch (the previous one's output taken first): in memory its
 * output follows with kg_grace_join_next; spilled, its rows go to their
 * partitions and come out after finish. */
int kg_grace_join_probe_push(kg_grace_join *g, int32_t ncols, const kg_column *cols,
                             int64_t nrows);
/* Every probe batch pushed: the preserved build rows and the spilled
 * partitions follow with kg_grace_join_next. */
int kg_grace_join_finish(kg_grace_join *g);
/* The next output batch of at most max_rows rows; *done = 1 when what was
 * pushed (or, after finish, everything) is out (*out may then be NULL). */
int kg_grace_join_next(kg_grace_join *g, int64_t max_rows, kg_batch **out, int32_t *done);
typedef struct kg_grace_join_stats {
  int32_t spilled;            /* the build passed the budget */
  uint32_t max_depth;         /* deepest repartitioning level joined */
  uint64_t partitions_joined; /* partitions joined in HBM (Counter::GracePartitions) */
  uint64_t repartitions;      /* partitions split again (Counter::Repartitions) */
  uint64_t reversed;          /* joined with roles reversed (Counter::ReversedPartitions) */
  uint64_t spill_bytes;       /* bytes written to host memory */
} kg_grace_join_stats;
int kg_grace_join_stats_get(const kg_grace_join *g, kg_grace_join_stats *out);
void kg_grace_join_destroy(kg_grace_join *g);

/* ---- join fused with GROUP BY the build key -----------------------------
 * The plan Aggregate(GROUP BY build key; SUM/COUNT/MIN/MAX/AVG of probe
 * columns) over an inner HashJoin (the shape of SELECT key, SUM(amt),
 * COUNT(*) FROM build JOIN probe ON key = fk GROUP BY key): the aggregate
 * state lives in the build table's slots, so no join pair is materialized.
 * Equal to kg_join_probe followed by kg_agg_sink, without the pairs.
 * Decimal SUM/AVG and ANSI integer SUM (checked sums) are taken while each
 * probe batch is proven unable to overflow them (rows x the most build rows
 * of a key x the largest magnitude, or the batch's sum of magnitudes, within
 * the bound); a batch that cannot be proven is refused with
 * KG_ERR_UNSUPPORTED before anything is applied.  Decimal128 inputs are
 * refused. */
int kg_joinagg_create(kg_hash_join *j, int32_t ninputs, const kg_column *input_types,
                      int32_t naggs, const kg_agg_spec *aggs, uint32_t flags);
/* Host probe batches of fewer than 512K rows are gathered and probed when
 * the next batch does not fit, the input columns change, a device batch
 * arrives, or at kg_joinagg_result, which then returns an error of that
 * work (a checked sum's refusal leaves the gathered rows in place). */
int kg_joinagg_probe(kg_hash_join *j, int32_t ncols, const kg_column *cols, int64_t nrows,
                     const int32_t *input_cols);
int kg_joinagg_result(kg_hash_join *j, kg_batch **out);

/* ---- grace hash join fused with the per-partition aggregate ------------
 * The grace path of the hash join (join/grace.rs:99-1031: spill whole
 * partitions, then per partition build a table and probe it) for the plan
 *   SELECT p, COUNT(*), SUM((probe.pay * build.pay) % 2^31)
 *   FROM build JOIN probe ON build.key = probe.key GROUP BY p
 * with p = partition_of(H_kg(key), partitions) (exchange routing).  Rows are
 * radix-partitioned on p into HBM partition buffers as they are sunk; past
 * the context's HBM budget partitions go to pinned host memory.  finish()
 * takes the partitions one at a time: the build partition is aggregated per
 * key (row count, payload sum) in a table within `table_budget_bytes`, then
 * the probe partition is streamed against it.  Per key, the pairs' count is
 * nb*np and their payload-product sum Sb*Sp, so no pair is materialized;
 * this equals the per-pair `% 2^31` only when every product is below 2^31
 * and non-negative, which finish() checks from the payloads' ranges
 * (KRISHIV_UNSUPPORTED_FEATURE otherwise).  Sums wrap at 64 bits (Spark
 * non-ANSI bigint SUM).  Integer-like int64 keys and payloads. */
typedef struct kg_grace kg_grace;
/* flags: KG_GRACE_SUM_MOD31 reads the SUM as ((sum of every pair's
 * probe.pay * build.pay, wrapping at 2^64) mod 2^31), which factors per key
 * for any payloads; without it, SUM((probe.pay * build.pay) % 2^31) per pair,
 * accepted only when every product lies in [0, 2^31). */
#define KG_GRACE_SUM_MOD31 1u
int kg_grace_create(kg_ctx *ctx, uint32_t partitions, uint64_t table_budget_bytes,
                    uint32_t flags, kg_grace **out);
int kg_grace_build_sink(kg_grace *g, const kg_column *key, const kg_column *pay, int64_t nrows);
int kg_grace_probe_sink(kg_grace *g, const kg_column *key, const kg_column *pay, int64_t nrows);
/* Per output partition (host arrays of `partitions`): the pair count, the
 * sum, and whether the sum is non-NULL (some pair had both payloads).
 * *spilled_bytes: partition bytes that were kept in pinned host memory. */
int kg_grace_finish(kg_grace *g, int64_t *counts, int64_t *sums, uint8_t *sum_valid,
                    uint64_t *spilled_bytes);
void kg_grace_destroy(kg_grace *g);

/* ---- batched, continuously served join-aggregate requests ---------------
 * A request is SELECT key, SUM(pay), COUNT(*) FROM build JOIN probe ON
 * key = fk GROUP BY key (int64 keys/payloads, device-resident inputs).  The
 * server packs every request waiting when the device frees into one batch:
 * persistent kernels pull (request, row range) tiles from a device work list,
 * so one launch per phase (table init, build, probe, output) serves requests
 * of any size mix; the phase sequence is one captured hipGraph replayed per
 * batch.  Requests are admitted while their tables, records and outputs fit
 * the arena (the HBM budget, the analogue of the session's admission
 * controller, crates/krishiv-session/src/admission.rs:168-191); build tables
 * registered with kg_server_add_table stay resident and are shared by the
 * requests that name them (the broadcast build tables of join/broadcast.rs).
 * Two batches can be in flight (two streams), so requests that arrive while
 * one batch runs are admitted into the next as arena space frees.  A request
 * whose probe has 4M rows or more is not batched: a worker thread runs it
 * through the radix-partitioned fused aggregate (kg_joinagg_*), on its own
 * stream, while the batches run on the device.  One worker by default;
 * the environment variable KG_SERVE_LARGE_WORKERS (1 to 4, read at create)
 * runs that many large requests at once (two on the same resident table run
 * one after the other). */
typedef struct kg_server kg_server;
typedef struct kg_serve_request {
  int64_t id;
  const int64_t *build_keys; /* device; NULL: probe the resident table `table` */
  int64_t build_rows;
  int32_t table;
  int32_t _pad;
  const int64_t *probe_keys;   /* device */
  const uint8_t *probe_valid;  /* device LSB bitmap or NULL */
  const int64_t *probe_pay;    /* device */
  int64_t probe_rows;
} kg_serve_request;
int kg_server_create(kg_ctx *ctx, uint64_t arena_bytes, int32_t max_batch, int32_t use_graph,
                     kg_server **out);
/* A resident build table over `n` device keys; *table_id names it. */
int kg_server_add_table(kg_server *s, const int64_t *keys, int64_t n, int32_t *table_id);
/* Queues a request (its inputs must stay valid until it completes). */
int kg_server_submit(kg_server *s, const kg_serve_request *r);
/* Non-blocking: retires finished batches (their ids to done_ids, at most
 * max; ids beyond max are reported by the next calls), then launches the
 * next batch from the queue if a stream is free.  wait != 0 blocks until at
 * least one batch retires (or nothing is left).  A failed large request
 * fails the call that meets it with no id reported yet. */
int kg_server_progress(kg_server *s, int32_t wait, int64_t *done_ids, int32_t max,
                       int32_t *ndone);
/* The finished request's result [key, sum, count] (once; then forgotten). */
int kg_server_take_result(kg_server *s, int64_t id, kg_batch **out);
/* Requests queued, in flight, or finished with their ids not yet reported. */
int64_t kg_server_pending(const kg_server *s);
void kg_server_destroy(kg_server *s);

/* ---- synthetic data on the device (bench and tests) --------------------
 * Counter-based generators; the integer ones are bit-identical to the
 * CPU generators used by the tests. */
int kg_gen_join_build(kg_ctx *ctx, uint64_t seed, int64_t lo, int64_t hi, int64_t n_total,
                      int64_t *key_dev, int64_t *val_dev);
/* fk_dev, fk_validity_dev and amt_dev may each be NULL (not generated). */
int kg_gen_join_probe(kg_ctx *ctx, uint64_t seed, int64_t lo, int64_t hi, int64_t key_range,
                      uint32_t null_permille, int64_t *fk_dev, uint8_t *fk_validity_dev,
                      int64_t *amt_dev);
/* The Zipf(s) cumulative distribution over n ranks (host, double). */
int kg_zipf_cdf(uint64_t n, double s, double *cdf_host);
/* The same table built on the device (for distributions of billions of
 * ranks): the terms summed by a parallel scan, so its entries differ from
 * kg_zipf_cdf's by rounding. */
int kg_zipf_cdf_device(kg_ctx *ctx, uint64_t n, double s, double *cdf_dev);
int kg_gen_agg(kg_ctx *ctx, uint64_t seed, int64_t lo, int64_t hi, const double *cdf_dev,
               uint64_t ndistinct, int64_t *key_dev, int64_t *dec_dev, double *f_dev,
               int32_t *i32_dev, int8_t *k8_dev);
/* Zipf keys (ranks drawn from the row seed) whose rank-to-key permutation
 * comes from key_seed: columns with different seeds and one key_seed share
 * one key distribution (the same keys are hot on both sides of a join);
 * key_seed == seed gives kg_gen_agg's key column. */
int kg_gen_zipf_keys(kg_ctx *ctx, uint64_t seed, uint64_t key_seed, int64_t lo, int64_t hi,
                     const double *cdf_dev, uint64_t ndistinct, int64_t *key_dev);

/* ---- profiling ----------------------------------------------------------
 * Device time of the hot kernels, from HIP events recorded around each launch
 * on the stream it runs on (the GPU analogue of the operators' EXPLAIN
 * ANALYZE timers, crates/krishiv-exec/src/join/mod.rs:148-167).  Off by
 * default. */
void kg_prof_enable(int on);
void kg_prof_reset(void);
/* Total device milliseconds and launch count of a kernel family
 * ("agg_update", "agg_flush", "join_build", "join_probe", "joinagg_probe",
 * "bloom_insert", "bloom_probe", "hash_partition", "scatter"). */
int kg_prof_get(const char *family, double *total_ms, int64_t *launches);

/* ---- shape autotuner of the hash join and its fused aggregate -----------
 * No counterpart in the reference (its operators have no device shapes to
 * pick).  Per size class of the build side (ceil log2 of its rows) a
 * context keeps one shape: the table's load factor (capacity >= rows * 100 /
 * load_factor_pct, a power of two; for a class nobody tuned 25 up to 2^20
 * build rows, 75 up to 2^23, else 50: what the tuner picked for the join
 * aggregate on an MI355X), the staged partition
 * scatter's rows per lane (4, 8 or 12; 0 = by the row width) and the
 * partition merge's workgroup (1024 lanes with the whole LDS, or 512 with
 * half).  kg_autotune_joinagg times all 18 candidates (load 50/25/75 x rows
 * 12/8/4 x lanes 1024/512) on synthetic inputs of the shape (unique build
 * keys, probe keys uniform over 1.2x their range, 1 % NULL, SUM and COUNT;
 * at most 2^26 probe rows, scaled to probe_rows) and keeps the fastest
 * (build + probe time).  Results never depend on the shape. */
typedef struct kg_tune_config {
  int32_t load_factor_pct;
  int32_t scatter_rows;
  int32_t merge_lanes;
  int32_t _pad;
} kg_tune_config;
int kg_tune_get(kg_ctx *ctx, int64_t build_rows, kg_tune_config *out);
/* load_factor_pct 0 returns the class to its untuned default. */
int kg_tune_set(kg_ctx *ctx, int64_t build_rows, const kg_tune_config *cfg);
int kg_autotune_joinagg(kg_ctx *ctx, int64_t build_rows, int64_t probe_rows, kg_tune_config *best,
                        double *best_ms, int32_t *candidates);
/* The context's tuned shapes to and from a text file (one line a size class:
 * "class load_pct scatter_rows merge_lanes"), so a process tunes once per
 * device and later ones load the result. */
int kg_tune_save(kg_ctx *ctx, const char *path);
int kg_tune_load(kg_ctx *ctx, const char *path, int32_t *classes);

/* ---- measured ceilings of the random-access phases ----------------------
 * No counterpart in the reference: calibration kernels that the probe
 * phases' rates are reported against (as the streaming phases' are against
 * HBM bandwidth).  kg_calib_random_read: random reads of read_bytes (8, 16
 * or 32) at hashed entries of a device table of table_bytes (a power of
 * two), 8 in flight per lane on 8 workgroups a CU.  kg_calib_atomic_add:
 * no-return 64-bit global atomic adds at hashed words among `addresses` (a
 * power of two; 1 puts every add on one word).  Each times `reps` launches
 * of `ops` operations (after one untimed launch) with HIP events on the
 * context's stream and returns the mean device ms per launch. */
int kg_calib_random_read(kg_ctx *ctx, uint64_t table_bytes, int32_t read_bytes, int64_t ops,
                         int32_t reps, double *ms_per_launch);
int kg_calib_atomic_add(kg_ctx *ctx, uint64_t addresses, int64_t ops, int32_t reps,
                        double *ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif
