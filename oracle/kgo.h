/*
 * kgo.h -- CPU oracle: a plain-C restatement of the reference engine's
 * hash-join / hash-aggregate / runtime-Bloom / hash-repartition algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path (libkgpu.so) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Parity pins (what checks each function
 * against something other than its own formula) are listed in the header
 * comment of each section.
 */
#ifndef KGO_H
#define KGO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Column types: the same numbering as include/kgpu.h (KG_T_*). */
enum {
  KGO_T_BOOL = 1,
  KGO_T_INT8 = 2,
  KGO_T_INT16 = 3,
  KGO_T_INT32 = 4,
  KGO_T_INT64 = 5,
  KGO_T_UINT8 = 6,
  KGO_T_UINT16 = 7,
  KGO_T_UINT32 = 8,
  KGO_T_UINT64 = 9,
  KGO_T_FLOAT32 = 10,
  KGO_T_FLOAT64 = 11,
  KGO_T_DATE32 = 12,
  KGO_T_TIMESTAMP = 13, /* any 64-bit temporal: date64, timestamps, durations */
  KGO_T_DECIMAL64 = 14,
  KGO_T_DECIMAL128 = 15
};

/* One Arrow-layout column: `values` (bit-packed for BOOL), optional LSB
 * validity bitmap (NULL = all valid), element offset into both. */
typedef struct kgo_col {
  int32_t type;
  int32_t precision;
  int32_t scale;
  int32_t _pad;
  int64_t length;
  int64_t offset;
  const void *values;
  const uint8_t *validity;
} kgo_col;

/* ---------------------------------------------------------------------
 * Routing hash H_kg v1  (crates/krishiv-kernels/src/hash/routing.rs)
 * Pinned by: routing tests' XXH64 vectors and H_kg goldens
 * (crates/krishiv-kernels/tests/routing.rs:30-33,41-145), Python xxhash, and
 * the partition/key-group tiling invariants (tests/routing.rs:385-420).
 * --------------------------------------------------------------------- */
uint64_t kgo_xxh64(const void *data, size_t len, uint64_t seed);
uint64_t kgo_xxh64_fixed(uint64_t lo, uint64_t hi, int width);
uint64_t kgo_fold_column(uint64_t h, uint64_t h_col);
uint32_t kgo_key_group(uint64_t h);
uint32_t kgo_radix(uint64_t h, uint32_t bits);
uint32_t kgo_partition_of(uint64_t h, uint32_t n);
void kgo_key_group_range(uint32_t p, uint32_t n, uint32_t *start, uint32_t *end);
/* H_kg of each row of `ncols` key columns (NULL_HASH for a NULL value). */
int kgo_hash_columns(int ncols, const kgo_col *cols, int64_t n, uint64_t *out);
/* exchange::route for a Hash partitioning: partition of each row. */
int kgo_route(int ncols, const kgo_col *cols, int64_t n, uint32_t partitions, uint32_t *out);
/* Stable split of row ids by partition (ExchangeSink::sink's `split`):
 * out_rows holds the row ids grouped by partition in input order,
 * out_offsets[p]..out_offsets[p+1] its range. */
void kgo_split(const uint32_t *pids, int64_t n, uint32_t partitions, int64_t *out_offsets,
               int64_t *out_rows);

/* ---------------------------------------------------------------------
 * Row-layout table hash h and finalizer g
 * (crates/krishiv-exec/src/rows/layout.rs:43-50,487-553,871-904,1080-1144)
 * Pinned by: the known answer 4733552696251797926 for keys
 * (float64 0.0, int32 NULL) (crates/krishiv-exec/src/join/keys.rs:348).
 * --------------------------------------------------------------------- */
uint64_t kgo_fold(uint64_t a, uint64_t b);
int kgo_row_hash(int ncols, const kgo_col *cols, int64_t n, uint64_t *out);
uint64_t kgo_finalize(uint64_t h, uint64_t seed);
/* join::keys::Hashing::partition (join/keys.rs:118-128) for the row layout */
uint32_t kgo_join_partition(uint64_t h, uint64_t seed, uint32_t radix_bits, uint32_t depth);

/* ---------------------------------------------------------------------
 * Split-block Bloom filter (crates/krishiv-io/src/runtime_filter/bloom.rs)
 * Pinned by: the Parquet SBBF bit-for-bit vector (bloom.rs:329-347,
 * regenerated from the Parquet specification with Python xxhash), the sizing
 * closed form, and the reference's sizing/false-positive bounds
 * (bloom.rs:349-378).
 * --------------------------------------------------------------------- */
size_t kgo_sbbf_bytes_for(uint64_t items, double fpp);
size_t kgo_sbbf_blocks_for_bytes(size_t bytes);
void kgo_sbbf_insert(uint32_t *words, size_t nblocks, uint64_t h);
int kgo_sbbf_check(const uint32_t *words, size_t nblocks, uint64_t h);
void kgo_sbbf_check_batch(const uint32_t *words, size_t nblocks, const uint64_t *h, int64_t n,
                          uint8_t *out);
double kgo_sbbf_expected_fpp(size_t nblocks, uint64_t items);
/* filter.rs:606-624 `bloom()`: build from hashes; returns 1 if kept (expected
 * fpp <= max_fpp), 0 if dropped.  bytes==0 sizes it from the hash count. */
int kgo_sbbf_build(const uint64_t *hashes, int64_t n, size_t bytes, double fpp, size_t max_bytes,
                   double max_fpp, uint32_t *words_out, size_t *nblocks_out);

/* ---------------------------------------------------------------------
 * Runtime KeyFilter (kgo_filter.c): KeyFilterBuilder's forms
 * (crates/krishiv-io/src/runtime_filter/filter.rs:180-615, member.rs),
 * evaluation (member.rs:439-560), the versioned encoding (codec.rs), the
 * merge of published filters (merge.rs:35-169) and the scan row stage with
 * its adaptive switch-off (crates/krishiv-exec/src/operators/runtime_filter.rs:240-395).
 * Pinned by: brute-force membership (exact forms keep exactly the stored
 * keys, every form keeps every stored key), the form cutoffs the reference
 * states (8 / 4096 members, the dense bit-set size rule with the numbers of
 * member.rs:138-146), hand-assembled encodings of the codec layout
 * (codec.rs:6-20,104-180), decode refusals (form_tests.rs:474-537), and
 * merge(a, b) against the filter built from the union of the keys.
 * --------------------------------------------------------------------- */
enum { KGO_SIDE_BOTH = 0, KGO_SIDE_UPPER = 1, KGO_SIDE_LOWER = 2 };
enum { KGO_FORM_RANGE = 0, KGO_FORM_SMALL = 1, KGO_FORM_EXACT = 2, KGO_FORM_BLOOM = 3,
       KGO_FORM_DENSE = 4 };
#define KGO_SMALL_SET_MAX 8                /* member.rs:34 */
#define KGO_DENSE_MAX_BITS (1u << 20)      /* member.rs:47 */
#define KGO_DENSE_SIZE_FACTOR 16           /* member.rs:53 */
#define KGO_DENSE_FREE_BYTES 4096          /* member.rs:56 */
#define KGO_HASH_SET_BYTES_PER_KEY 9       /* member.rs:60 */
#define KGO_FILTER_FORMAT_VERSION 1        /* codec.rs:34 */
#define KGO_FILTER_ROW_HASH_ID 1           /* hash.rs:40-44 */
#define KGO_FILTER_NORMALIZATION_VERSION 1 /* codec.rs:36 */

typedef struct kgo_filter_opts {
  uint64_t max_values;       /* the statistics IN list cap (50) */
  uint64_t exact_max_values; /* exact set cap (4096) */
  int32_t bloom;             /* collect hashes for a Bloom filter */
  int32_t side;              /* KGO_SIDE_* */
  uint64_t bloom_max_items;  /* 4,000,000 */
  uint64_t bloom_max_bytes;  /* 8 MiB */
  double bloom_fpp;          /* 0.01 */
  double bloom_max_fpp;      /* 0.05 */
  uint64_t bloom_bytes;      /* the planned size (0: sized for the keys) */
} kgo_filter_opts;

typedef struct kgo_kf_info {
  int32_t form;     /* KGO_FORM_* */
  int32_t has_range;
  int64_t nmembers; /* small / exact members (a dense set: 0, see kgo_kf_members) */
  int64_t values;   /* the IN list's length, -1 without one */
  uint64_t lo_lo, lo_hi, hi_lo, hi_hi; /* integer / decimal range */
  double flo, fhi;                     /* float range */
  int32_t nulls_pass, empty;
  uint64_t keys;
  uint64_t bloom_bytes;
} kgo_kf_info;

typedef struct kgo_kf kgo_kf;
typedef struct kgo_rs kgo_rs;
void kgo_filter_default_opts(kgo_filter_opts *o);
int kgo_kf_widens(int from, int fp, int fs, int to, int tp, int ts);
kgo_kf *kgo_kf_new(int type, int precision, int scale, const kgo_filter_opts *o);
void kgo_kf_free(kgo_kf *f);
int kgo_kf_add(kgo_kf *f, const kgo_col *c, const uint8_t *include, int64_t n);
int kgo_kf_merge_builder(kgo_kf *a, const kgo_kf *b);
int kgo_kf_finish(kgo_kf *f);
int kgo_kf_get_info(const kgo_kf *f, kgo_kf_info *o);
int64_t kgo_kf_members(const kgo_kf *f, uint64_t *out);
int64_t kgo_kf_bloom_words(const kgo_kf *f, uint32_t *out);
double kgo_kf_row_cost(const kgo_kf *f);
int64_t kgo_kf_evaluate(const kgo_kf *f, const kgo_col *c, int64_t n, uint8_t *out);
int64_t kgo_kf_encode(const kgo_kf *f, uint8_t *out, int64_t cap);
kgo_kf *kgo_kf_decode(const uint8_t *bytes, int64_t len, char *err);
kgo_kf *kgo_kf_merge(const kgo_kf *a, const kgo_kf *b, const kgo_filter_opts *o, char *err);
double kgo_sbbf_pass_rate(const uint32_t *words, size_t nblocks);
kgo_rs *kgo_rs_new(uint64_t min_rows, double max_pass_rate);
void kgo_rs_free(kgo_rs *s);
int kgo_rs_set(kgo_rs *s, int slot, const kgo_kf *f, int column);
void kgo_rs_start_partition(kgo_rs *s);
void kgo_rs_stats(const kgo_rs *s, int slot, uint64_t *tested, uint64_t *dropped, int *off);
int64_t kgo_rs_apply(kgo_rs *s, int ncols, const kgo_col *cols, int64_t n, uint8_t *out);

/* ---------------------------------------------------------------------
 * Exchange frames (kgo_frame.c; crates/krishiv-shuffle/src/frame.rs).
 * Pinned by: the CRC32C check value of "123456789" (0xE3069283) and of the
 * empty input (frame.rs tests.rs:122-126), a hand-assembled header, and the
 * refusals of damaged frames (tests.rs:84-116).
 * --------------------------------------------------------------------- */
typedef struct kgo_frame_hdr {
  uint64_t id;
  uint32_t stage_attempt, map, attempt, sequence, reduce;
  uint32_t kind; /* 1 Data, 2 End, 3 Value */
} kgo_frame_hdr;
enum { KGO_FRAME_ETRUNCATED = 1, KGO_FRAME_EMAGIC = 2, KGO_FRAME_EVERSION = 3,
       KGO_FRAME_EKIND = 4, KGO_FRAME_ECOMPRESSION = 5, KGO_FRAME_ECHECKSUM = 6,
       KGO_FRAME_EBODY = 7 };
uint32_t kgo_crc32c(const uint8_t *p, uint64_t n);
void kgo_frame_header(const kgo_frame_hdr *h, const uint8_t *body, uint64_t len, uint8_t out[44]);
int kgo_frame_decode(const uint8_t *f, uint64_t n, kgo_frame_hdr *h, uint64_t *body_off,
                     uint64_t *body_len);

/* ---------------------------------------------------------------------
 * Synthetic data generators (counter-based; identical integer arithmetic on
 * the device in libkgpu).
 * --------------------------------------------------------------------- */
uint64_t kgo_mix64(uint64_t z);
uint64_t kgo_rand(uint64_t seed, uint64_t stream, uint64_t i);
uint64_t kgo_below(uint64_t r, uint64_t n);
double kgo_unit(uint64_t r);
uint64_t kgo_perm(uint64_t i, uint64_t n, uint64_t seed);
void kgo_zipf_cdf(uint64_t n, double s, double *cdf);
uint64_t kgo_zipf_rank(const double *cdf, uint64_t n, double u);
/* config generators (see bench.py and the tests for the shapes) */
void kgo_gen_join_build(uint64_t seed, int64_t lo, int64_t hi, int64_t n_total, int64_t *key,
                        int64_t *val);
void kgo_gen_join_probe(uint64_t seed, int64_t lo, int64_t hi, int64_t key_range,
                        uint32_t null_permille, int64_t *fk, uint8_t *fk_valid, int64_t *amt);
void kgo_gen_agg(uint64_t seed, int64_t lo, int64_t hi, const double *cdf, uint64_t ndistinct,
                 int64_t *key, int64_t *dec, double *f, int32_t *i32, int8_t *k8);
/* Zipf keys whose rank-to-key permutation comes from key_seed (shared by
 * both sides of a join); key_seed == seed gives kgo_gen_agg's keys. */
void kgo_gen_zipf_keys(uint64_t seed, uint64_t key_seed, int64_t lo, int64_t hi,
                       const double *cdf, uint64_t ndistinct, int64_t *key);

/* ---------------------------------------------------------------------
 * Hash aggregation with Spark semantics
 * (crates/krishiv-exec/src/aggregate/mod.rs:588-775,929-1025,1083-1113;
 *  accumulator/dense.rs:55-345; accumulator/helpers.rs:120-330,595-700;
 *  accumulator/output.rs:230-285)
 * Pinned by: Spark 4.2.0 goldens tests/fixtures/spark-4.2.0/decimal-sum-avg.json
 * (172 cases), exact Python integer/decimal arithmetic, and math.fsum for
 * float sums.
 * --------------------------------------------------------------------- */
enum {
  KGO_AGG_COUNT_STAR = 1,
  KGO_AGG_COUNT = 2,
  KGO_AGG_SUM = 3,
  KGO_AGG_MIN = 4,
  KGO_AGG_MAX = 5,
  KGO_AGG_AVG = 6
};
typedef struct kgo_agg_spec {
  int32_t func;
  int32_t input; /* input column index (ignored for COUNT_STAR) */
} kgo_agg_spec;

/* One output column of the aggregate: which buffer is used depends on the
 * result type (i64 for integers/counts, f64 for doubles, lo/hi for
 * decimals, lo for Decimal64 results). */
typedef struct kgo_out {
  int32_t type; /* KGO_T_* of the result */
  int32_t precision;
  int32_t scale;
  int32_t _pad;
  int64_t *i64;  /* also decimal low word */
  int64_t *hi;   /* decimal high word */
  double *f64;
  uint8_t *valid; /* one byte per row */
} kgo_out;

typedef struct kgo_agg kgo_agg;
/* key: the grouping column (NULL for a global aggregate). */
kgo_agg *kgo_agg_new(const kgo_col *key_type_like, int ninputs, const kgo_col *input_types,
                     int naggs, const kgo_agg_spec *aggs, int ansi);
/* Raw rows, in order (HashAggregate::sink -> RowTable::apply). */
int kgo_agg_update(kgo_agg *a, const kgo_col *key, const kgo_col *inputs, int64_t n);
/* Partial states of `other`, merged into `a` in group-arrival order
 * (combine/finalize of a two-phase aggregate). */
int kgo_agg_merge(kgo_agg *a, const kgo_agg *other);
int64_t kgo_agg_groups(const kgo_agg *a);
/* Final values: out_key (i64 + valid) and one kgo_out per aggregate. */
int kgo_agg_finish(const kgo_agg *a, int64_t *out_key, uint8_t *out_key_valid, kgo_out *outs);
/* Result type of aggregate i: type, precision, scale. */
void kgo_agg_result_type(const kgo_agg *a, int i, int32_t *type, int32_t *precision,
                         int32_t *scale);
void kgo_agg_free(kgo_agg *a);
const char *kgo_last_error(void);
/* Float SUM/AVG as a compensated (TwoSum) sum: a fsum-grade reference value
 * for checking float sums whose addition order differs (0: the reference's
 * row-order double adds, the default). */
void kgo_agg_set_exact_f64(kgo_agg *a, int on);

/* The aggregate on `threads` threads, shaped as the reference's thread pool
 * runs it: thread-local tables in 16 radix partitions (aggregate/mod.rs:11-21,
 * rowtable.rs:137-139,278-284), then a parallel merge per partition
 * (final_tasks, rowtable.rs:1805).  Integer and decimal results equal
 * kgo_agg's; float sums differ by their order (or not at all with
 * exact_f64).  finish writes the groups partition by partition. */
typedef struct kgo_pagg kgo_pagg;
kgo_pagg *kgo_pagg_new(const kgo_col *key_type_like, int ninputs, const kgo_col *input_types,
                       int naggs, const kgo_agg_spec *aggs, int ansi, int exact_f64,
                       int threads);
int kgo_pagg_update(kgo_pagg *p, const kgo_col *key, const kgo_col *inputs, int64_t n);
int kgo_pagg_finalize(kgo_pagg *p);
int64_t kgo_pagg_groups(const kgo_pagg *p);
int kgo_pagg_finish(kgo_pagg *p, int64_t *out_key, uint8_t *out_key_valid, kgo_out *outs);
void kgo_pagg_result_type(const kgo_pagg *p, int i, int32_t *type, int32_t *precision,
                          int32_t *scale);
const char *kgo_pagg_error(const kgo_pagg *p);
void kgo_pagg_free(kgo_pagg *p);

/* decimal avg by one HALF_UP division (accumulator/helpers.rs:595-609).
 * Returns 1 and *out on success, 0 when the shortcut does not apply or the
 * value does not fit decimal(p, s). */
int kgo_avg_fast(int64_t sum_lo, int64_t sum_hi, int sum_scale, int64_t count, int p, int s,
                 int64_t *out_lo, int64_t *out_hi);

/* Spark's hash() (Murmur3 x86_32, crates/krishiv-kernels/src/hash/
 * murmur3.rs:9-52) and xxhash64() (seed 42) of int and bigint values. */
int32_t kgo_murmur3_int(int32_t v, int32_t seed);
int32_t kgo_murmur3_long(int64_t v, int32_t seed);
int64_t kgo_spark_xxhash64(int64_t v, int width);
void kgo_spark_xxhash64_batch(const int64_t *v, int64_t n, int width, int64_t *out);

/* ---------------------------------------------------------------------
 * Hash join (crates/krishiv-exec/src/join/mod.rs:1-60,217-293;
 * join/probe.rs:266-501,532-632).  Spark's ten join types on `=` (NULL never
 * matches) or `<=>` keys over integer-like or float columns, with an
 * optional residual filter.
 * Pinned by: a nested-loop join on small inputs, and a pure-Python
 * restatement of the reference's own oracle and its edge cases
 * (crates/krishiv-exec/tests/hash_join/main.rs:130-290, VLess filter).
 * --------------------------------------------------------------------- */
enum {
  KGO_JOIN_INNER = 0,
  KGO_JOIN_LEFT_OUTER = 1,
  KGO_JOIN_LEFT_SEMI = 2,
  KGO_JOIN_LEFT_ANTI = 3,
  KGO_JOIN_NULL_AWARE_ANTI = 4,
  KGO_JOIN_EXISTENCE = 5,
  KGO_JOIN_RIGHT_SEMI = 6,
  KGO_JOIN_RIGHT_ANTI = 7,
  KGO_JOIN_RIGHT_OUTER = 8,
  KGO_JOIN_FULL_OUTER = 9
};
/* Residual filter program: postfix ops over (probe, build) columns. */
enum {
  KGO_X_PROBE = 1, /* push probe column `arg` */
  KGO_X_BUILD = 2, /* push build column `arg` */
  KGO_X_CONST = 3, /* push `imm` */
  KGO_X_ADD = 4,
  KGO_X_SUB = 5,
  KGO_X_MUL = 6,
  KGO_X_MOD = 7,
  KGO_X_EQ = 8,
  KGO_X_NE = 9,
  KGO_X_LT = 10,
  KGO_X_LE = 11,
  KGO_X_GT = 12,
  KGO_X_GE = 13,
  KGO_X_AND = 14,
  KGO_X_OR = 15,
  KGO_X_NOT = 16,
  KGO_X_IS_NULL = 17
};
typedef struct kgo_expr {
  int32_t op;
  int32_t arg;
  int64_t imm;
} kgo_expr;
/* Emits (probe row, build row) pairs: -1 pads the unmatched side; semi,
 * anti and null-aware anti emit (probe row, -1); right semi/anti (-1, build
 * row); existence (probe row, exists as 1/0).  Probe rows first, then the
 * build-side tail.  Returns the count, or -1 if `cap` is too small.
 * `nested` selects the nested-loop form (the reference test oracle). */
int64_t kgo_join_ex(int join_type, int nkeys, const kgo_col *probe_keys, const kgo_col *build_keys,
                    const uint8_t *null_safe, int64_t n_probe, int64_t n_build, int nested,
                    int nfilter, const kgo_expr *filter, const kgo_col *probe_cols,
                    const kgo_col *build_cols, int64_t cap, int64_t *out_probe,
                    int64_t *out_build);
/* kgo_join_ex without a residual filter. */
int64_t kgo_join(int join_type, int nkeys, const kgo_col *probe_keys, const kgo_col *build_keys,
                 const uint8_t *null_safe, int64_t n_probe, int64_t n_build, int nested,
                 int64_t cap, int64_t *out_probe, int64_t *out_build);

/* ---------------------------------------------------------------------
 * Configuration 4's plan through the grace join: per output partition
 * p = partition_of(H_kg(key), parts), the join pairs' COUNT(*) and the sum
 * of probe.pay * build.pay, factorized per key (nb * np, Sb * Sp).  mode 0:
 * the wrapping 64-bit sum; mode 1: that sum mod 2^31.  sum_valid: some pair
 * had both payloads.  Pinned by the pair walk of kgo_join on small inputs.
 * --------------------------------------------------------------------- */
int kgo_grace_factorized(const kgo_col *bkey, const kgo_col *bpay, int64_t nb, const kgo_col *pkey,
                         const kgo_col *ppay, int64_t np, uint32_t parts, int mode,
                         int64_t *counts, int64_t *sums, uint8_t *sum_valid);

#ifdef __cplusplus
}
#endif
#endif
