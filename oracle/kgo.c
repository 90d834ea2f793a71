/*
 * kgo.c -- CPU oracle for the hash-join / hash-aggregate / runtime-Bloom /
 * hash-repartition path.  TEST INFRASTRUCTURE ONLY (see kgo.h): a plain-C,
 * single-threaded restatement of the reference's Rust operators, used as the
 * checker of the HIP path and as the timed CPU baseline.  Never linked into
 * the product library.
 *
 * Build: make -C oracle   (gcc, -ffp-contract=off so float sums are exactly
 * the source-order IEEE adds).
 */
#include "kgo.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

static __thread char g_err[256];
static int set_err(const char *cls, const char *msg) {
  snprintf(g_err, sizeof g_err, "%s: %s", cls, msg);
  return -1;
}
const char *kgo_last_error(void) { return g_err; }

/* ------------------------------------------------------------------ */
/* column access                                                        */
/* ------------------------------------------------------------------ */
static inline int col_valid(const kgo_col *c, int64_t i) {
  if (!c->validity) return 1;
  int64_t j = i + c->offset;
  return (c->validity[j >> 3] >> (j & 7)) & 1;
}

static int type_width(int t) {
  switch (t) {
    case KGO_T_BOOL: return 0;
    case KGO_T_INT8: case KGO_T_UINT8: return 1;
    case KGO_T_INT16: case KGO_T_UINT16: return 2;
    case KGO_T_INT32: case KGO_T_UINT32: case KGO_T_FLOAT32: case KGO_T_DATE32: return 4;
    case KGO_T_INT64: case KGO_T_UINT64: case KGO_T_FLOAT64: case KGO_T_TIMESTAMP:
    case KGO_T_DECIMAL64: return 8;
    case KGO_T_DECIMAL128: return 16;
    default: return -1;
  }
}

/* Integer-like value widened to i128 (signed types sign-extend). */
static i128 col_int(const kgo_col *c, int64_t i) {
  int64_t j = i + c->offset;
  switch (c->type) {
    case KGO_T_BOOL: return (((const uint8_t *)c->values)[j >> 3] >> (j & 7)) & 1;
    case KGO_T_INT8: return ((const int8_t *)c->values)[j];
    case KGO_T_UINT8: return ((const uint8_t *)c->values)[j];
    case KGO_T_INT16: return ((const int16_t *)c->values)[j];
    case KGO_T_UINT16: return ((const uint16_t *)c->values)[j];
    case KGO_T_INT32: case KGO_T_DATE32: return ((const int32_t *)c->values)[j];
    case KGO_T_UINT32: return ((const uint32_t *)c->values)[j];
    case KGO_T_INT64: case KGO_T_TIMESTAMP: case KGO_T_DECIMAL64:
      return ((const int64_t *)c->values)[j];
    case KGO_T_UINT64: return (i128)((const uint64_t *)c->values)[j];
    case KGO_T_DECIMAL128: {
      const uint64_t *w = (const uint64_t *)c->values + 2 * j;
      return (i128)(((u128)w[1] << 64) | w[0]);
    }
    default: return 0;
  }
}

static double col_f64(const kgo_col *c, int64_t i) {
  int64_t j = i + c->offset;
  if (c->type == KGO_T_FLOAT64) return ((const double *)c->values)[j];
  if (c->type == KGO_T_FLOAT32) return ((const float *)c->values)[j];
  return (double)col_int(c, i);
}

static int is_float(int t) { return t == KGO_T_FLOAT32 || t == KGO_T_FLOAT64; }
static int is_decimal(int t) { return t == KGO_T_DECIMAL64 || t == KGO_T_DECIMAL128; }
static int is_integral(int t) {
  return t == KGO_T_INT8 || t == KGO_T_INT16 || t == KGO_T_INT32 || t == KGO_T_INT64;
}

/* ------------------------------------------------------------------ */
/* XXH64 and H_kg (crates/krishiv-kernels/src/hash/routing.rs)          */
/* ------------------------------------------------------------------ */
#define P1 0x9E3779B185EBCA87ULL
#define P2 0xC2B2AE3D27D4EB4FULL
#define P3 0x165667B19E3779F9ULL
#define P4 0x85EBCA77C2B2AE63ULL
#define P5 0x27D4EB2F165667C5ULL
#define NULL_HASH 0xA5A55A5AC3C33C3CULL /* routing.rs:78 */
#define FOLD_INIT 0x243F6A8885A308D3ULL /* routing.rs:81 */
#define FOLD_MUL 0xB7E151628AED2A6BULL  /* routing.rs:84 */

static inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

/* routing.rs:93-99 */
static inline uint64_t avalanche(uint64_t h) {
  h ^= h >> 33;
  h *= P2;
  h ^= h >> 29;
  h *= P3;
  return h ^ (h >> 32);
}

/* xxh64.rs:24-33: one accumulator round and the merge of four */
static inline uint64_t xround(uint64_t acc, uint64_t in) { return rotl(acc + in * P2, 31) * P1; }
static inline uint64_t xmerge(uint64_t acc, uint64_t v) { return (acc ^ xround(0, v)) * P1 + P4; }
static inline uint64_t le64(const uint8_t *b) {
  uint64_t v;
  memcpy(&v, b, 8);
  return v;
}
static inline uint64_t le32(const uint8_t *b) {
  uint32_t v;
  memcpy(&v, b, 4);
  return v;
}

/* General XXH64 (crates/krishiv-kernels/src/hash/xxh64.rs:73-118, the
 * published algorithm). */
uint64_t kgo_xxh64(const void *data, size_t len, uint64_t seed) {
  const uint8_t *p = (const uint8_t *)data;
  size_t rest = len;
  uint64_t h;
  if (len >= 32) {
    uint64_t v[4] = {seed + P1 + P2, seed + P2, seed, seed - P1};
    while (rest >= 32) {
      for (int k = 0; k < 4; k++) v[k] = xround(v[k], le64(p + 8 * k));
      p += 32;
      rest -= 32;
    }
    h = rotl(v[0], 1) + rotl(v[1], 7) + rotl(v[2], 12) + rotl(v[3], 18);
    for (int k = 0; k < 4; k++) h = xmerge(h, v[k]);
  } else {
    h = seed + P5;
  }
  h += (uint64_t)len;
  while (rest >= 8) {
    h ^= xround(0, le64(p));
    h = rotl(h, 27) * P1 + P4;
    p += 8;
    rest -= 8;
  }
  if (rest >= 4) {
    h ^= le32(p) * P1;
    h = rotl(h, 23) * P2 + P3;
    p += 4;
    rest -= 4;
  }
  while (rest > 0) {
    h ^= (uint64_t)(*p) * P5;
    h = rotl(h, 11) * P1;
    p++;
    rest--;
  }
  return avalanche(h);
}

/* routing.rs:102-106 */
static inline uint64_t lane8(uint64_t h, uint64_t k) {
  uint64_t k1 = rotl(k * P2, 31) * P1;
  return rotl(h ^ k1, 27) * P1 + P4;
}

/* xxh64_fixed<N> (routing.rs:112-142): the short-input path unrolled for a
 * width of 1, 2, 4, 8 or 16 little-endian bytes held in (lo, hi). */
uint64_t kgo_xxh64_fixed(uint64_t lo, uint64_t hi, int width) {
  uint64_t h = P5 + (uint64_t)width;
  int left = width;
  if (left >= 8) {
    h = lane8(h, lo);
    left -= 8;
    lo = hi;
    hi = 0;
  }
  if (left >= 8) {
    h = lane8(h, lo);
    left -= 8;
    lo = 0;
  }
  if (left >= 4) {
    h = rotl(h ^ ((lo & 0xFFFFFFFFULL) * P1), 23) * P2 + P3;
    lo >>= 32;
    left -= 4;
  }
  while (left > 0) {
    h = rotl(h ^ ((lo & 0xFF) * P5), 11) * P1;
    lo >>= 8;
    left--;
  }
  return avalanche(h);
}

/* fold_column (routing.rs:172-183): mix(h ^ h_col), a folded multiply. */
uint64_t kgo_fold_column(uint64_t h, uint64_t h_col) {
  u128 p = (u128)(h ^ h_col) * (u128)FOLD_MUL;
  return (uint64_t)p ^ (uint64_t)(p >> 64);
}

/* routing.rs:186-190 */
uint32_t kgo_key_group(uint64_t h) { return (uint32_t)(h >> 52); }

/* routing.rs:194-205 */
uint32_t kgo_radix(uint64_t h, uint32_t bits) {
  if (bits == 0) return 0;
  uint32_t b = bits < 32 ? bits : 32;
  return (uint32_t)(h >> (64 - b));
}

/* partition_of (routing.rs:217-231) */
uint32_t kgo_partition_of(uint64_t h, uint32_t n) {
  if (n == 0) n = 1;
  if (n <= 4096) return (uint32_t)(((uint64_t)kgo_key_group(h) * n) >> 12);
  return (uint32_t)(((u128)h * (u128)n) >> 64);
}

/* key_group_range (routing.rs:233-256) */
void kgo_key_group_range(uint32_t p, uint32_t n, uint32_t *start, uint32_t *end) {
  if (n == 0) n = 1;
  if (p >= n) {
    *start = *end = 0;
    return;
  }
  uint64_t g = 4096, P = p, N = n;
  if (N <= g) {
    *start = (uint32_t)((P * g + N - 1) / N);
    *end = (uint32_t)(((P + 1) * g + N - 1) / N);
  } else {
    uint32_t kg = (uint32_t)(P * g / N);
    *start = kg;
    *end = kg + 1;
  }
}

/* The column hash of one value (routing.rs:332-420 and :424-447). */
static uint64_t value_hash(const kgo_col *c, int64_t i) {
  if (!col_valid(c, i)) return NULL_HASH;
  int64_t j = i + c->offset;
  switch (c->type) {
    case KGO_T_BOOL: return kgo_xxh64_fixed((uint64_t)col_int(c, i), 0, 1);
    case KGO_T_INT8: case KGO_T_UINT8:
      return kgo_xxh64_fixed(((const uint8_t *)c->values)[j], 0, 1);
    case KGO_T_INT16: case KGO_T_UINT16:
      return kgo_xxh64_fixed(((const uint16_t *)c->values)[j], 0, 2);
    case KGO_T_INT32: case KGO_T_UINT32: case KGO_T_DATE32:
      return kgo_xxh64_fixed(((const uint32_t *)c->values)[j], 0, 4);
    case KGO_T_INT64: case KGO_T_UINT64: case KGO_T_TIMESTAMP:
      return kgo_xxh64_fixed(((const uint64_t *)c->values)[j], 0, 8);
    case KGO_T_DECIMAL64: case KGO_T_DECIMAL128: {
      /* decimal_hash: the unscaled value as 16 LE bytes (routing.rs:431-436) */
      i128 v = col_int(c, i);
      return kgo_xxh64_fixed((uint64_t)v, (uint64_t)((u128)v >> 64), 16);
    }
    case KGO_T_FLOAT32: {
      /* f32_hash (routing.rs:438-443): NaN -> canonical, -0.0 -> 0.0 */
      float f = ((const float *)c->values)[j];
      uint32_t bits;
      if (f != f) {
        bits = 0x7FC00000u;
      } else {
        f = f + 0.0f;
        memcpy(&bits, &f, 4);
      }
      return kgo_xxh64_fixed(bits, 0, 4);
    }
    case KGO_T_FLOAT64: {
      /* f64_hash (routing.rs:445-450) */
      double f = ((const double *)c->values)[j];
      uint64_t bits;
      if (f != f) {
        bits = 0x7FF8000000000000ULL;
      } else {
        f = f + 0.0;
        memcpy(&bits, &f, 8);
      }
      return kgo_xxh64_fixed(bits, 0, 8);
    }
    default: return 0;
  }
}

/* hash_columns (routing.rs:258-282): one column is its own hash; several
 * fold from FOLD_INIT in key order. */
int kgo_hash_columns(int ncols, const kgo_col *cols, int64_t n, uint64_t *out) {
  for (int c = 0; c < ncols; c++)
    if (type_width(cols[c].type) < 0) return set_err("KRISHIV_UNSUPPORTED_FEATURE", "key type");
  if (ncols == 1) {
    for (int64_t i = 0; i < n; i++) out[i] = value_hash(&cols[0], i);
    return 0;
  }
  for (int64_t i = 0; i < n; i++) {
    uint64_t h = FOLD_INIT;
    for (int c = 0; c < ncols; c++) h = kgo_fold_column(h, value_hash(&cols[c], i));
    out[i] = h;
  }
  return 0;
}

/* exchange::route, Partitioning::Hash (exchange/mod.rs:159-212,
 * physical/mod.rs:248-262): partition_of(H_kg(keys), n). */
int kgo_route(int ncols, const kgo_col *cols, int64_t n, uint32_t partitions, uint32_t *out) {
  if (partitions <= 1 || ncols == 0) {
    memset(out, 0, sizeof(uint32_t) * (size_t)n);
    return 0;
  }
  uint64_t *h = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
  if (kgo_hash_columns(ncols, cols, n, h) != 0) {
    free(h);
    return -1;
  }
  for (int64_t i = 0; i < n; i++) out[i] = kgo_partition_of(h[i], partitions);
  free(h);
  return 0;
}

/* ExchangeSink's split (exchange/mod.rs:477-501): a take per partition, rows
 * in input order. */
void kgo_split(const uint32_t *pids, int64_t n, uint32_t partitions, int64_t *out_offsets,
               int64_t *out_rows) {
  memset(out_offsets, 0, sizeof(int64_t) * (partitions + 1));
  for (int64_t i = 0; i < n; i++) out_offsets[pids[i] + 1]++;
  for (uint32_t p = 0; p < partitions; p++) out_offsets[p + 1] += out_offsets[p];
  int64_t *at = (int64_t *)malloc(sizeof(int64_t) * (partitions ? partitions : 1));
  memcpy(at, out_offsets, sizeof(int64_t) * partitions);
  for (int64_t i = 0; i < n; i++) out_rows[at[pids[i]]++] = i;
  free(at);
}

/* ------------------------------------------------------------------ */
/* Row-layout hash h and g (crates/krishiv-exec/src/rows/layout.rs)     */
/* ------------------------------------------------------------------ */
#define L_MIX 0x243F6A8885A308D3ULL    /* layout.rs:45 */
#define L_FINISH 0xB7E151628AED2A6BULL /* layout.rs:46 */
#define L_FIN 0x6A09E667F3BCC909ULL    /* layout.rs:50 */

/* fold (layout.rs:1131-1139): the folded multiply */
uint64_t kgo_fold(uint64_t a, uint64_t b) {
  u128 p = (u128)a * (u128)b;
  return (uint64_t)p ^ (uint64_t)(p >> 64);
}
static inline uint64_t lmix(uint64_t h, uint64_t v) { return kgo_fold(h ^ v, L_MIX); }

/* RowLayout::encode's stored hash (layout.rs:487-553, put_fixed :871-904,
 * write_nulls :787-797): h = fold(0, MIX); per key column in key order a NULL
 * mixes NULL_HASH and a value mixes its LE 8-byte words (floats normalized);
 * finally h = fold(h, FINISH). */
int kgo_row_hash(int ncols, const kgo_col *cols, int64_t n, uint64_t *out) {
  for (int64_t i = 0; i < n; i++) {
    uint64_t h = kgo_fold(0, L_MIX);
    for (int c = 0; c < ncols; c++) {
      const kgo_col *k = &cols[c];
      if (!col_valid(k, i)) {
        h = lmix(h, NULL_HASH);
        continue;
      }
      int64_t j = i + k->offset;
      switch (k->type) {
        case KGO_T_FLOAT64: {
          double f = ((const double *)k->values)[j];
          uint64_t bits;
          if (f != f) bits = 0x7FF8000000000000ULL;
          else {
            f = f + 0.0;
            memcpy(&bits, &f, 8);
          }
          h = lmix(h, bits);
          break;
        }
        case KGO_T_FLOAT32: {
          float f = ((const float *)k->values)[j];
          uint32_t bits;
          if (f != f) bits = 0x7FC00000u;
          else {
            f = f + 0.0f;
            memcpy(&bits, &f, 4);
          }
          h = lmix(h, bits);
          break;
        }
        case KGO_T_DECIMAL128: {
          const uint64_t *w = (const uint64_t *)k->values + 2 * j;
          h = lmix(lmix(h, w[0]), w[1]);
          break;
        }
        case KGO_T_BOOL: h = lmix(h, (uint64_t)col_int(k, i)); break;
        default: {
          int w = type_width(k->type);
          uint64_t word = 0;
          memcpy(&word, (const uint8_t *)k->values + (size_t)j * (size_t)w, (size_t)w);
          h = lmix(h, word);
        }
      }
    }
    out[i] = kgo_fold(h, L_FINISH);
  }
  return 0;
}

/* finalize (layout.rs:1080-1082): g = fold(h ^ seed, FIN) */
uint64_t kgo_finalize(uint64_t h, uint64_t seed) { return kgo_fold(h ^ seed, L_FIN); }

/* Hashing::partition for the row layout (join/keys.rs:49-59,118-128) */
uint32_t kgo_join_partition(uint64_t h, uint64_t seed, uint32_t radix_bits, uint32_t depth) {
  if (radix_bits < 3) radix_bits = 3;
  if (radix_bits > 6) radix_bits = 6;
  if (depth == 0) return (uint32_t)(h >> (64 - radix_bits));
  uint32_t d = depth < 6 ? depth : 6;
  uint32_t shift = 48 - 3 * d;
  return (uint32_t)((kgo_finalize(h, seed) >> shift) & 7);
}

/* ------------------------------------------------------------------ */
/* SBBF (crates/krishiv-io/src/runtime_filter/bloom.rs)                 */
/* ------------------------------------------------------------------ */
static const uint32_t SALT[8] = {0x47b6137bu, 0x44974d91u, 0x8824ad5bu, 0xa2b7289du,
                                 0x705495c7u, 0x2df1424bu, 0x9efc4947u, 0x5c6bfb31u};

/* bytes_for (bloom.rs:96-110) */
size_t kgo_sbbf_bytes_for(uint64_t items, double fpp) {
  if (fpp < 1e-9) fpp = 1e-9;
  if (fpp > 0.5) fpp = 0.5;
  double bits = -8.0 * (double)items / log(1.0 - pow(fpp, 1.0 / 8.0));
  size_t bytes = (size_t)ceil(bits / 8.0);
  size_t blocks = (bytes + 31) / 32;
  if (blocks < 1) blocks = 1;
  return blocks * 32;
}

/* with_bytes (bloom.rs:114-119) */
size_t kgo_sbbf_blocks_for_bytes(size_t bytes) {
  size_t b = (bytes + 31) / 32;
  return b ? b : 1;
}

/* block (bloom.rs:128-131) */
static inline size_t sbbf_block(uint64_t h, size_t nblocks) {
  return (size_t)(((h >> 32) * (uint64_t)nblocks) >> 32);
}

/* insert_hash (bloom.rs:134-148) with mask (bloom.rs:83-89) */
void kgo_sbbf_insert(uint32_t *words, size_t nblocks, uint64_t h) {
  size_t b = sbbf_block(h, nblocks);
  uint32_t key = (uint32_t)h;
  for (int w = 0; w < 8; w++) words[b * 8 + w] |= 1u << ((key * SALT[w]) >> 27);
}

/* check_hash (bloom.rs:184-199) */
int kgo_sbbf_check(const uint32_t *words, size_t nblocks, uint64_t h) {
  size_t b = sbbf_block(h, nblocks);
  uint32_t key = (uint32_t)h, miss = 0;
  for (int w = 0; w < 8; w++) {
    uint32_t bit = 1u << ((key * SALT[w]) >> 27);
    miss |= bit & ~words[b * 8 + w];
  }
  return miss == 0;
}

/* check_batch (bloom.rs:203-247): one result byte per hash */
void kgo_sbbf_check_batch(const uint32_t *words, size_t nblocks, const uint64_t *h, int64_t n,
                          uint8_t *out) {
  for (int64_t i = 0; i < n; i++) out[i] = (uint8_t)kgo_sbbf_check(words, nblocks, h[i]);
}

/* expected_fpp (bloom.rs:270-308) */
double kgo_sbbf_expected_fpp(size_t nblocks, uint64_t items) {
  double lambda = (double)items / (double)nblocks;
  if (lambda == 0.0) return 0.0;
  if (lambda > 1024.0) return pow(1.0 - pow(31.0 / 32.0, lambda), 8);
  double spread = 10.0 * sqrt(lambda) + 20.0;
  double lo = floor(fmax(lambda - spread, 0.0)), hi = lambda + spread;
  double ln_p = -lambda;
  double k = 1.0;
  while (k <= lo) {
    ln_p += log(lambda) - log(k);
    k += 1.0;
  }
  double total = 0.0;
  k = lo;
  while (k <= hi) {
    double word = 1.0 - pow(31.0 / 32.0, k);
    double w2 = word * word, w4 = w2 * w2;
    total += exp(ln_p) * (w4 * w4);
    k += 1.0;
    ln_p += log(lambda) - log(k);
  }
  return total < 1.0 ? total : 1.0;
}

/* bloom() (runtime_filter/filter.rs:606-624): sized exactly for the hash
 * count (duplicates counted) or planned, capped at max_bytes, kept only if
 * expected_fpp <= max_fpp. */
int kgo_sbbf_build(const uint64_t *hashes, int64_t n, size_t bytes, double fpp, size_t max_bytes,
                   double max_fpp, uint32_t *words_out, size_t *nblocks_out) {
  if (n <= 0) {
    *nblocks_out = 0;
    return 0;
  }
  size_t b = bytes ? bytes : kgo_sbbf_bytes_for((uint64_t)n, fpp);
  if (b > max_bytes) b = max_bytes;
  size_t nb = kgo_sbbf_blocks_for_bytes(b);
  memset(words_out, 0, nb * 32);
  for (int64_t i = 0; i < n; i++) kgo_sbbf_insert(words_out, nb, hashes[i]);
  *nblocks_out = nb;
  return kgo_sbbf_expected_fpp(nb, (uint64_t)n) <= max_fpp;
}

/* ------------------------------------------------------------------ */
/* Spark's hash functions over int and bigint values: Murmur3 x86_32 as    */
/* Spark runs it (crates/krishiv-kernels/src/hash/murmur3.rs:9-52) and     */
/* Spark's xxhash64 (seed 42).  Pinned by Spark 4.2.0's hash()/xxhash64()   */
/* goldens (tests/fixtures/spark-4.2.0/hash.json).                          */
/* ------------------------------------------------------------------ */
static uint32_t m3_rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

/* mix_k1 / mix_h1 / fmix (murmur3.rs:12-30) */
static uint32_t m3_mix_k1(uint32_t k1) { return m3_rotl(k1 * 0xcc9e2d51u, 15) * 0x1b873593u; }
static uint32_t m3_mix_h1(uint32_t h1, uint32_t k1) { return m3_rotl(h1 ^ k1, 13) * 5u + 0xe6546b64u; }
static uint32_t m3_fmix(uint32_t h1, uint32_t len) {
  uint32_t h = h1 ^ len;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  return h ^ (h >> 16);
}

/* hash_int (murmur3.rs:39-42) */
int32_t kgo_murmur3_int(int32_t v, int32_t seed) {
  return (int32_t)m3_fmix(m3_mix_h1((uint32_t)seed, m3_mix_k1((uint32_t)v)), 4);
}

/* hash_long (murmur3.rs:44-52): the low, then the high 32 bits */
int32_t kgo_murmur3_long(int64_t v, int32_t seed) {
  uint64_t b = (uint64_t)v;
  uint32_t h1 = m3_mix_h1((uint32_t)seed, m3_mix_k1((uint32_t)(b & 0xffffffffu)));
  h1 = m3_mix_h1(h1, m3_mix_k1((uint32_t)(b >> 32)));
  return (int32_t)m3_fmix(h1, 8);
}

/* Spark's xxhash64(v) of an int (4 bytes) or a bigint (8 bytes), seed 42 */
int64_t kgo_spark_xxhash64(int64_t v, int width) {
  uint8_t b[8];
  uint64_t u = (uint64_t)v;
  for (int i = 0; i < 8; i++) b[i] = (uint8_t)(u >> (8 * i));
  return (int64_t)kgo_xxh64(b, (size_t)width, 42);
}

void kgo_spark_xxhash64_batch(const int64_t *v, int64_t n, int width, int64_t *out) {
  for (int64_t i = 0; i < n; i++) out[i] = kgo_spark_xxhash64(v[i], width);
}

/* ------------------------------------------------------------------ */
/* generators                                                           */
/* ------------------------------------------------------------------ */
/* SplitMix64's output function (Steele, Lea, Flood 2014). */
uint64_t kgo_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
/* Counter-based stream: value i of stream `stream` under `seed`. */
uint64_t kgo_rand(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t base = kgo_mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL));
  return kgo_mix64(base + (i + 1) * 0x9E3779B97F4A7C15ULL);
}
uint64_t kgo_below(uint64_t r, uint64_t n) { return (uint64_t)(((u128)r * n) >> 64); }
double kgo_unit(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

/* A bijection of [0, n): a 4-round Feistel network on the smallest even
 * bit width covering n, cycle-walked back into range. */
uint64_t kgo_perm(uint64_t i, uint64_t n, uint64_t seed) {
  if (n <= 1) return 0;
  int bits = 2;
  while (bits < 64 && (1ULL << bits) < n) bits++;
  if (bits & 1) bits++;
  int half = bits / 2;
  uint64_t mask = (1ULL << half) - 1;
  uint64_t x = i;
  do {
    uint64_t L = x >> half, R = x & mask;
    for (uint64_t r = 0; r < 4; r++) {
      uint64_t f = kgo_mix64(R ^ (seed * 0x9E3779B97F4A7C15ULL) ^ (r << 56)) & mask;
      uint64_t t = L ^ f;
      L = R;
      R = t;
    }
    x = (L << half) | R;
  } while (x >= n);
  return x;
}

/* Zipf(s) over ranks 0..n-1 (probability of rank k proportional to
 * (k+1)^-s): the cumulative distribution, in one left-to-right pass. */
void kgo_zipf_cdf(uint64_t n, double s, double *cdf) {
  double acc = 0.0;
  for (uint64_t k = 0; k < n; k++) {
    acc += pow((double)(k + 1), -s);
    cdf[k] = acc;
  }
  double inv = 1.0 / acc;
  for (uint64_t k = 0; k < n; k++) cdf[k] *= inv;
  cdf[n - 1] = 1.0;
}
/* Smallest rank k with cdf[k] > u. */
uint64_t kgo_zipf_rank(const double *cdf, uint64_t n, double u) {
  uint64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] > u) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

/* Join build side: key = a seeded permutation of 0..n_total-1, val in
 * [0, 1e6).  Rows lo..hi-1 of the full table. */
void kgo_gen_join_build(uint64_t seed, int64_t lo, int64_t hi, int64_t n_total, int64_t *key,
                        int64_t *val) {
  for (int64_t i = lo; i < hi; i++) {
    key[i - lo] = (int64_t)kgo_perm((uint64_t)i, (uint64_t)n_total, seed);
    if (val) val[i - lo] = (int64_t)kgo_below(kgo_rand(seed, 1, (uint64_t)i), 1000000);
  }
}

/* Join probe side: fk in [0, key_range), NULL with probability
 * null_permille/1000 (validity as one byte per row), amt in [1, 1e4]. */
void kgo_gen_join_probe(uint64_t seed, int64_t lo, int64_t hi, int64_t key_range,
                        uint32_t null_permille, int64_t *fk, uint8_t *fk_valid, int64_t *amt) {
  for (int64_t i = lo; i < hi; i++) {
    uint64_t u = (uint64_t)i;
    fk[i - lo] = (int64_t)kgo_below(kgo_rand(seed, 0, u), (uint64_t)key_range);
    if (fk_valid) fk_valid[i - lo] = kgo_below(kgo_rand(seed, 1, u), 1000) >= null_permille;
    if (amt) amt[i - lo] = 1 + (int64_t)kgo_below(kgo_rand(seed, 2, u), 10000);
  }
}

/* Aggregation input: key = a permutation of a Zipf rank over ndistinct
 * values; dec (decimal(12,2) cents) in [1, 1e7); f in [0, 1); i32 in
 * [0, 50]; k8 (the low-cardinality int8 key) in [0, 4). */
void kgo_gen_zipf_keys(uint64_t seed, uint64_t key_seed, int64_t lo, int64_t hi,
                       const double *cdf, uint64_t ndistinct, int64_t *key) {
  for (int64_t i = lo; i < hi; i++) {
    uint64_t r = kgo_zipf_rank(cdf, ndistinct, kgo_unit(kgo_rand(seed, 0, (uint64_t)i)));
    key[i - lo] = (int64_t)kgo_perm(r, ndistinct, key_seed ^ 0x5BD1E995ULL);
  }
}

void kgo_gen_agg(uint64_t seed, int64_t lo, int64_t hi, const double *cdf, uint64_t ndistinct,
                 int64_t *key, int64_t *dec, double *f, int32_t *i32, int8_t *k8) {
  for (int64_t i = lo; i < hi; i++) {
    uint64_t u = (uint64_t)i;
    if (key) {
      uint64_t r = kgo_zipf_rank(cdf, ndistinct, kgo_unit(kgo_rand(seed, 0, u)));
      key[i - lo] = (int64_t)kgo_perm(r, ndistinct, seed ^ 0x5BD1E995ULL);
    }
    if (dec) dec[i - lo] = 1 + (int64_t)kgo_below(kgo_rand(seed, 1, u), 9999999);
    if (f) f[i - lo] = kgo_unit(kgo_rand(seed, 2, u));
    if (i32) i32[i - lo] = (int32_t)kgo_below(kgo_rand(seed, 3, u), 51);
    if (k8) k8[i - lo] = (int8_t)kgo_below(kgo_rand(seed, 4, u), 4);
  }
}

/* ------------------------------------------------------------------ */
/* Hash aggregation                                                     */
/* ------------------------------------------------------------------ */
static i128 pow10i(int p) {
  i128 v = 1;
  for (int i = 0; i < p; i++) v *= 10;
  return v;
}
static inline u128 uabs(i128 v) { return v < 0 ? (u128)0 - (u128)v : (u128)v; }

/* Spark's float ordering for min/max: NaN is the largest value, -0.0 and
 * 0.0 compare equal (aggregate/mod.rs:985-995). Returns <0, 0, >0. */
static int fcmp(double a, double b) {
  int an = a != a, bn = b != b;
  if (an || bn) return an - bn;
  return (a > b) - (a < b);
}

enum { K_COUNT, K_SUM_INT, K_SUM_F, K_SUM_DEC, K_MINMAX_INT, K_MINMAX_F, K_AVG_F, K_AVG_DEC };

typedef struct {
  int func, input, kind;
  int in_type, in_p, in_s;
  int out_type, out_p, out_s;
  int sum_p;    /* precision of the decimal sum type (avg_sum_type) */
  i128 bound;   /* 10^sum_p: running decimal sums stay strictly inside */
} agg_desc;

typedef struct {
  int64_t cnt;  /* COUNT, AVG count, SUM seen (as 0/1) */
  union {
    i128 v;     /* int sum (low 64 bits wrap), decimal sum, min/max value */
    struct {
      double f; /* float sum / float min/max */
      double c; /* compensation of a float sum (exact_f64 mode only) */
    };
  };
  uint8_t st;   /* 0 none, 1 value, 2 overflowed (decimal) */
} agg_state;

struct kgo_agg {
  int has_key, key_type, naggs, ninputs, ansi;
  /* float SUM/AVG as a compensated (TwoSum) sum: a reference value for
   * checking sums whose order differs, not the reference's own row-order
   * double adds */
  int exact_f64;
  agg_desc *d;
  int64_t ngroups, cap_groups;
  int64_t *keys;
  uint8_t *key_valid;
  agg_state *st; /* ngroups x naggs */
  /* open-addressing map key -> group (the NULL key has its own group) */
  int64_t *slots; /* group index + 1, 0 = empty */
  int64_t nslots;
  int64_t null_group; /* -1 if none */
};

static int describe(agg_desc *d, int func, int input, const kgo_col *types) {
  d->func = func;
  d->input = input;
  if (func == KGO_AGG_COUNT_STAR || func == KGO_AGG_COUNT) {
    d->kind = K_COUNT;
    d->out_type = KGO_T_INT64;
    return 0;
  }
  const kgo_col *t = &types[input];
  d->in_type = t->type;
  d->in_p = t->precision;
  d->in_s = t->scale;
  /* computed_type (aggregate/mod.rs:595-640) */
  if (func == KGO_AGG_SUM) {
    if (is_integral(t->type)) {
      d->kind = K_SUM_INT;
      d->out_type = KGO_T_INT64;
    } else if (is_float(t->type)) {
      d->kind = K_SUM_F;
      d->out_type = KGO_T_FLOAT64;
    } else if (is_decimal(t->type)) {
      /* sum_decimal: decimal(min(p+10, 38), s) on the input's carrier rule */
      d->kind = K_SUM_DEC;
      d->out_p = d->sum_p = t->precision + 10 > 38 ? 38 : t->precision + 10;
      d->out_s = t->scale;
      d->out_type = (t->type == KGO_T_DECIMAL64 && d->out_p <= 18) ? KGO_T_DECIMAL64
                                                                    : KGO_T_DECIMAL128;
      d->bound = pow10i(d->sum_p);
    } else {
      return set_err("KRISHIV_UNSUPPORTED_FEATURE", "sum of this type");
    }
  } else if (func == KGO_AGG_AVG) {
    if (is_integral(t->type) || is_float(t->type)) {
      d->kind = K_AVG_F;
      d->out_type = KGO_T_FLOAT64;
    } else if (is_decimal(t->type)) {
      d->kind = K_AVG_DEC;
      d->sum_p = t->precision + 10 > 38 ? 38 : t->precision + 10;
      d->bound = pow10i(d->sum_p);
      d->out_p = t->precision + 4 > 38 ? 38 : t->precision + 4;
      d->out_s = t->scale + 4 > 38 ? 38 : t->scale + 4;
      d->out_type = (t->type == KGO_T_DECIMAL64 && d->out_p <= 18) ? KGO_T_DECIMAL64
                                                                    : KGO_T_DECIMAL128;
    } else {
      return set_err("KRISHIV_UNSUPPORTED_FEATURE", "avg of this type");
    }
  } else if (func == KGO_AGG_MIN || func == KGO_AGG_MAX) {
    d->kind = is_float(t->type) ? K_MINMAX_F : K_MINMAX_INT;
    d->out_type = t->type;
    d->out_p = t->precision;
    d->out_s = t->scale;
  } else {
    return set_err("KRISHIV_UNSUPPORTED_FEATURE", "aggregate function");
  }
  return 0;
}

kgo_agg *kgo_agg_new(const kgo_col *key_type_like, int ninputs, const kgo_col *input_types,
                     int naggs, const kgo_agg_spec *aggs, int ansi) {
  kgo_agg *a = (kgo_agg *)calloc(1, sizeof(kgo_agg));
  a->has_key = key_type_like != NULL;
  a->key_type = key_type_like ? key_type_like->type : 0;
  a->naggs = naggs;
  a->ninputs = ninputs;
  a->ansi = ansi;
  a->d = (agg_desc *)calloc((size_t)(naggs ? naggs : 1), sizeof(agg_desc));
  for (int i = 0; i < naggs; i++) {
    if (describe(&a->d[i], aggs[i].func, aggs[i].input, input_types) != 0) {
      kgo_agg_free(a);
      return NULL;
    }
  }
  a->nslots = 1024;
  a->slots = (int64_t *)calloc((size_t)a->nslots, sizeof(int64_t));
  a->null_group = -1;
  return a;
}

void kgo_agg_free(kgo_agg *a) {
  if (!a) return;
  free(a->d);
  free(a->keys);
  free(a->key_valid);
  free(a->st);
  free(a->slots);
  free(a);
}

int64_t kgo_agg_groups(const kgo_agg *a) { return a->ngroups; }

static inline uint64_t key_slot_hash(int64_t k) {
  return kgo_mix64((uint64_t)k * 0x9E3779B97F4A7C15ULL);
}

static int64_t new_group(kgo_agg *a, int64_t key, int valid) {
  if (a->ngroups == a->cap_groups) {
    int64_t nc = a->cap_groups ? a->cap_groups * 2 : 64;
    a->keys = (int64_t *)realloc(a->keys, sizeof(int64_t) * (size_t)nc);
    a->key_valid = (uint8_t *)realloc(a->key_valid, (size_t)nc);
    a->st = (agg_state *)realloc(a->st, sizeof(agg_state) * (size_t)nc * (size_t)(a->naggs ? a->naggs : 1));
    a->cap_groups = nc;
  }
  int64_t g = a->ngroups++;
  a->keys[g] = key;
  a->key_valid[g] = (uint8_t)valid;
  memset(&a->st[g * a->naggs], 0, sizeof(agg_state) * (size_t)a->naggs);
  return g;
}

static void rehash(kgo_agg *a) {
  int64_t ns = a->nslots * 2;
  int64_t *s = (int64_t *)calloc((size_t)ns, sizeof(int64_t));
  for (int64_t g = 0; g < a->ngroups; g++) {
    if (!a->key_valid[g]) continue;
    uint64_t p = key_slot_hash(a->keys[g]) & (uint64_t)(ns - 1);
    while (s[p]) p = (p + 1) & (uint64_t)(ns - 1);
    s[p] = g + 1;
  }
  free(a->slots);
  a->slots = s;
  a->nslots = ns;
}

static int64_t find_group(kgo_agg *a, int64_t key, int valid) {
  if (!a->has_key) return a->ngroups ? 0 : new_group(a, 0, 1);
  if (!valid) {
    if (a->null_group < 0) a->null_group = new_group(a, 0, 0);
    return a->null_group;
  }
  uint64_t p = key_slot_hash(key) & (uint64_t)(a->nslots - 1);
  while (a->slots[p]) {
    int64_t g = a->slots[p] - 1;
    if (a->keys[g] == key) return g;
    p = (p + 1) & (uint64_t)(a->nslots - 1);
  }
  int64_t g = new_group(a, key, 1);
  a->slots[p] = g + 1;
  if ((a->ngroups - (a->null_group >= 0)) * 2 > a->nslots) rehash(a);
  return g;
}

/* A decimal running sum after adding x: checked, and strictly inside the
 * bound (DecFold::sum / DecFold::fits, accumulator/helpers.rs:141-226). */
static int dec_add(i128 *sum, i128 x, i128 bound) {
  i128 s;
  if (__builtin_add_overflow(*sum, x, &s)) return 0;
  if (uabs(s) >= (u128)bound) return 0;
  *sum = s;
  return 1;
}

/* A float sum's add: the plain double add (the reference), or with
 * exact_f64 Knuth's TwoSum, whose rounding error collects in `c`. */
static inline void fadd(const kgo_agg *a, agg_state *s, double x) {
  if (!a->exact_f64) {
    s->f += x;
    return;
  }
  double t = s->f + x, bp = t - s->f;
  s->c += (s->f - (t - bp)) + (x - bp);
  s->f = t;
}

static int fold_row(kgo_agg *a, agg_state *s, const agg_desc *d, const kgo_col *in, int64_t i) {
  if (d->kind == K_COUNT) {
    if (d->func == KGO_AGG_COUNT_STAR || col_valid(&in[d->input], i)) s->cnt += 1;
    return 0;
  }
  const kgo_col *c = &in[d->input];
  if (!col_valid(c, i)) return 0;
  switch (d->kind) {
    case K_SUM_INT: {
      /* sum_ints (accumulator/dense.rs:55-125): wrapping; ANSI fails on the
       * first add that overflows i64 in row order. */
      int64_t x = (int64_t)col_int(c, i), cur = (int64_t)s->v, r;
      if (__builtin_add_overflow(cur, x, &r)) {
        if (a->ansi) return set_err("ARITHMETIC_OVERFLOW", "long overflow in sum");
      }
      s->v = (int64_t)((uint64_t)cur + (uint64_t)x);
      s->cnt = 1;
      return 0;
    }
    case K_SUM_F:
      /* sum_f64 (dense.rs:289-345): row order, NULL skipped */
      fadd(a, s, col_f64(c, i));
      s->cnt = 1;
      return 0;
    case K_SUM_DEC:
      if (s->st != 2) {
        if (dec_add(&s->v, col_int(c, i), d->bound)) s->st = 1;
        else {
          s->v = 0;
          s->st = 2;
        }
      }
      return 0;
    case K_MINMAX_INT: {
      /* fold_extremum (dense.rs:214-263): strictly better replaces */
      i128 x = col_int(c, i);
      int better = d->func == KGO_AGG_MAX ? x > s->v : x < s->v;
      if (!s->st || better) {
        s->v = x;
        s->st = 1;
      }
      return 0;
    }
    case K_MINMAX_F: {
      double x = col_f64(c, i);
      int cmp = fcmp(x, s->f);
      int better = d->func == KGO_AGG_MAX ? cmp > 0 : cmp < 0;
      if (!s->st || better) {
        s->f = x;
        s->st = 1;
      }
      return 0;
    }
    case K_AVG_F:
      /* AvgFloat (accumulator/fold.rs:401-424): double sum in row order */
      fadd(a, s, col_f64(c, i));
      s->cnt += 1;
      return 0;
    case K_AVG_DEC:
      /* DecFold::avg (helpers.rs:229-287): an overflowing add leaves the sum
       * and sets overflow; the count always advances. */
      if (!dec_add(&s->v, col_int(c, i), d->bound)) s->st = 2;
      s->cnt += 1;
      return 0;
  }
  return 0;
}

static inline int update_row(kgo_agg *a, const kgo_col *key, const kgo_col *inputs, int64_t i) {
  int64_t k = 0;
  int kv = 1;
  if (a->has_key) {
    kv = col_valid(key, i);
    k = kv ? (int64_t)col_int(key, i) : 0;
  }
  int64_t g = find_group(a, k, kv);
  for (int j = 0; j < a->naggs; j++)
    if (fold_row(a, &a->st[g * a->naggs + j], &a->d[j], inputs, i) != 0) return -1;
  return 0;
}

int kgo_agg_update(kgo_agg *a, const kgo_col *key, const kgo_col *inputs, int64_t n) {
  for (int64_t i = 0; i < n; i++)
    if (update_row(a, key, inputs, i) != 0) return -1;
  return 0;
}

void kgo_agg_set_exact_f64(kgo_agg *a, int on) { a->exact_f64 = on; }

/* Merge partial states (the state-column folds of accumulator/fold.rs and
 * DecFold with `w = None`, helpers.rs:168-190,244-254). */
int kgo_agg_merge(kgo_agg *a, const kgo_agg *o) {
  for (int64_t og = 0; og < o->ngroups; og++) {
    int64_t g = find_group(a, o->keys[og], o->key_valid[og]);
    for (int j = 0; j < a->naggs; j++) {
      agg_state *s = &a->st[g * a->naggs + j];
      const agg_state *t = &o->st[og * o->naggs + j];
      const agg_desc *d = &a->d[j];
      switch (d->kind) {
        case K_COUNT: s->cnt += t->cnt; break;
        case K_SUM_INT:
          if (t->cnt) {
            int64_t cur = (int64_t)s->v, x = (int64_t)t->v, r;
            if (__builtin_add_overflow(cur, x, &r) && a->ansi)
              return set_err("ARITHMETIC_OVERFLOW", "long overflow in sum");
            s->v = (int64_t)((uint64_t)cur + (uint64_t)x);
            s->cnt = 1;
          }
          break;
        case K_SUM_F:
          if (t->cnt) {
            fadd(a, s, t->f);
            s->c += t->c;
            s->cnt = 1;
          }
          break;
        case K_SUM_DEC:
          /* partial (sum, nonempty): NULL sum with nonempty is sticky */
          if (t->st == 2) {
            s->st = 2;
            s->v = 0;
          } else if (t->st == 1 && s->st != 2) {
            if (dec_add(&s->v, t->v, d->bound)) s->st = 1;
            else {
              s->v = 0;
              s->st = 2;
            }
          }
          break;
        case K_MINMAX_INT:
          if (t->st) {
            int better = d->func == KGO_AGG_MAX ? t->v > s->v : t->v < s->v;
            if (!s->st || better) {
              s->v = t->v;
              s->st = 1;
            }
          }
          break;
        case K_MINMAX_F:
          if (t->st) {
            int cmp = fcmp(t->f, s->f);
            int better = d->func == KGO_AGG_MAX ? cmp > 0 : cmp < 0;
            if (!s->st || better) {
              s->f = t->f;
              s->st = 1;
            }
          }
          break;
        case K_AVG_F:
          if (t->cnt) {
            fadd(a, s, t->f);
            s->c += t->c;
          }
          s->cnt += t->cnt;
          break;
        case K_AVG_DEC:
          if (t->st == 2) {
            if (t->cnt != 0) s->st = 2;
          } else if (!dec_add(&s->v, t->v, d->bound)) {
            s->st = 2;
          }
          s->cnt += t->cnt;
          break;
      }
    }
  }
  return 0;
}

/* avg_fast (accumulator/helpers.rs:595-609) */
static int avg_fast128(i128 sum, int ss, int64_t count, int p, int s, i128 *out) {
  if (s > 19 || s < ss || count == 0) return 0;
  u128 m = uabs(sum), f = (u128)pow10i(s - ss), n;
  if (m && f > ((u128)0 - 1) / m) return 0; /* checked_mul */
  n = m * f;
  u128 dd = (u128)(count < 0 ? -(i128)count : (i128)count);
  u128 q = n / dd, r = n % dd;
  q += (r >= dd - r) ? 1 : 0;
  if (q >= (u128)pow10i(p)) return 0;
  i128 v = (i128)q;
  *out = ((sum < 0) != (count < 0)) ? -v : v;
  return 1;
}

int kgo_avg_fast(int64_t sum_lo, int64_t sum_hi, int sum_scale, int64_t count, int p, int s,
                 int64_t *out_lo, int64_t *out_hi) {
  i128 sum = (i128)(((u128)(uint64_t)sum_hi << 64) | (uint64_t)sum_lo), v;
  if (!avg_fast128(sum, sum_scale, count, p, s, &v)) return 0;
  *out_lo = (int64_t)(uint64_t)v;
  *out_hi = (int64_t)(uint64_t)((u128)v >> 64);
  return 1;
}

/* avg_decimal's value (the digit loop of accumulator/helpers.rs:520-579):
 * quotient rounded HALF_UP to scale 39, then HALF_UP to s.  Returns 1 and
 * the value when it fits decimal(p, s). */
static int avg_digits(i128 sum, int ss, int64_t count, int p, int s, i128 *out) {
  int negative = (sum < 0) != (count < 0);
  u128 a = uabs(sum), b = (u128)(count < 0 ? -(i128)count : (i128)count);
  u128 in = a / b, rem = a % b;
  int nfrac = 39 - ss;
  uint8_t frac[64];
  for (int k = 0; k < nfrac; k++) {
    rem *= 10;
    frac[k] = (uint8_t)(rem / b);
    rem %= b;
  }
  rem *= 10;
  if ((uint8_t)(rem / b) >= 5) {
    int carry = 1;
    for (int k = nfrac - 1; k >= 0 && carry; k--) {
      if (frac[k] == 9) frac[k] = 0;
      else {
        frac[k]++;
        carry = 0;
      }
    }
    if (carry) in += 1;
  }
  int keep = s - ss > 0 ? s - ss : 0;
  int up = keep < nfrac && frac[keep] >= 5;
  u128 v = in;
  u128 lim = (u128)pow10i(p);
  for (int k = 0; k < keep; k++) {
    if (v > (((u128)0 - 1) - frac[k]) / 10) return 0;
    v = v * 10 + frac[k];
  }
  v += (u128)up;
  if (v >= lim) return 0;
  *out = negative ? -(i128)v : (i128)v;
  return 1;
}

void kgo_agg_result_type(const kgo_agg *a, int i, int32_t *type, int32_t *precision,
                         int32_t *scale) {
  *type = a->d[i].out_type;
  *precision = a->d[i].out_p;
  *scale = a->d[i].out_s;
}

/* Final values (accumulator/output.rs:220-285). */
int kgo_agg_finish(const kgo_agg *a, int64_t *out_key, uint8_t *out_key_valid, kgo_out *outs) {
  for (int64_t g = 0; g < a->ngroups; g++) {
    if (out_key) out_key[g] = a->keys[g];
    if (out_key_valid) out_key_valid[g] = a->key_valid[g];
  }
  for (int j = 0; j < a->naggs; j++) {
    const agg_desc *d = &a->d[j];
    kgo_out *o = &outs[j];
    for (int64_t g = 0; g < a->ngroups; g++) {
      const agg_state *s = &a->st[g * a->naggs + j];
      uint8_t valid = 1;
      switch (d->kind) {
        case K_COUNT: o->i64[g] = s->cnt; break;
        case K_SUM_INT:
          valid = s->cnt != 0;
          o->i64[g] = valid ? (int64_t)s->v : 0;
          break;
        case K_SUM_F:
          valid = s->cnt != 0;
          o->f64[g] = valid ? s->f + s->c : 0.0;
          break;
        case K_SUM_DEC:
          if (s->st == 0) valid = 0;
          else if (s->st == 1 && uabs(s->v) < (u128)pow10i(d->out_p)) {
            o->i64[g] = (int64_t)(uint64_t)s->v;
            if (o->hi) o->hi[g] = (int64_t)(uint64_t)((u128)s->v >> 64);
          } else if (!a->ansi) {
            valid = 0;
          } else if (s->st == 1) {
            return set_err("NUMERIC_VALUE_OUT_OF_RANGE.WITH_SUGGESTION", "sum");
          } else {
            return set_err("ARITHMETIC_OVERFLOW", "Overflow in sum of decimals");
          }
          break;
        case K_MINMAX_INT:
          valid = s->st != 0;
          if (o->i64) o->i64[g] = (int64_t)(uint64_t)s->v;
          if (o->hi) o->hi[g] = (int64_t)(uint64_t)((u128)s->v >> 64);
          break;
        case K_MINMAX_F:
          valid = s->st != 0;
          o->f64[g] = s->f;
          break;
        case K_AVG_F:
          valid = s->cnt != 0;
          o->f64[g] = valid ? (s->f + s->c) / (double)s->cnt : 0.0;
          break;
        case K_AVG_DEC: {
          if (s->cnt == 0) {
            valid = 0;
            break;
          }
          if (s->st == 2) {
            if (a->ansi) return set_err("ARITHMETIC_OVERFLOW", "Overflow in sum of decimals");
            valid = 0;
            break;
          }
          i128 v;
          int ss = d->in_s;
          if (avg_fast128(s->v, ss, s->cnt, d->out_p, d->out_s, &v) ||
              avg_digits(s->v, ss, s->cnt, d->out_p, d->out_s, &v)) {
            o->i64[g] = (int64_t)(uint64_t)v;
            if (o->hi) o->hi[g] = (int64_t)(uint64_t)((u128)v >> 64);
          } else if (a->ansi) {
            return set_err("NUMERIC_VALUE_OUT_OF_RANGE.WITH_SUGGESTION", "avg");
          } else {
            valid = 0;
          }
          break;
        }
      }
      if (o->valid) o->valid[g] = valid;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Parallel hash aggregation, shaped as the reference runs it on a thread */
/* pool: each driver thread folds its share of the rows into its own      */
/* table, split into 16 radix partitions by the top 4 bits of the key's   */
/* hash (aggregate/mod.rs:11-21, rowtable.rs:137-139,278-284); finalize    */
/* merges partition p of every thread's table, in thread order, one task   */
/* a partition on the pool (final_tasks, rowtable.rs:1805).  Integer and   */
/* decimal results equal the single-threaded aggregate's; float sums       */
/* depend on the merge order, as the reference's do (dense.rs:289-293).    */
/* ------------------------------------------------------------------ */
#include <pthread.h>

#define KGO_PAGG_PARTS 16

struct kgo_pagg {
  int threads, has_key, naggs;
  kgo_agg **tl; /* threads x KGO_PAGG_PARTS */
  int64_t *off; /* KGO_PAGG_PARTS + 1 group offsets after finalize */
  int finalized;
  char err[256];
  pthread_mutex_t mu;
};

static inline int pagg_part(const kgo_pagg *p, const kgo_col *key, int64_t i) {
  if (!p->has_key || !col_valid(key, i)) return 0;
  return (int)(key_slot_hash((int64_t)col_int(key, i)) >> 60);
}

typedef struct {
  kgo_pagg *p;
  int t;
  const kgo_col *key, *inputs;
  int64_t lo, hi;
  int rc;
} pagg_job;

static void pagg_fail(kgo_pagg *p) {
  pthread_mutex_lock(&p->mu);
  if (!p->err[0]) snprintf(p->err, sizeof p->err, "%s", g_err);
  pthread_mutex_unlock(&p->mu);
}

static void *pagg_update_job(void *arg) {
  pagg_job *j = (pagg_job *)arg;
  kgo_agg **tab = &j->p->tl[(size_t)j->t * KGO_PAGG_PARTS];
  for (int64_t i = j->lo; i < j->hi; i++) {
    if (update_row(tab[pagg_part(j->p, j->key, i)], j->key, j->inputs, i) != 0) {
      j->rc = -1;
      pagg_fail(j->p);
      return NULL;
    }
  }
  return NULL;
}

kgo_pagg *kgo_pagg_new(const kgo_col *key_type_like, int ninputs, const kgo_col *input_types,
                       int naggs, const kgo_agg_spec *aggs, int ansi, int exact_f64,
                       int threads) {
  if (threads < 1) threads = 1;
  kgo_pagg *p = (kgo_pagg *)calloc(1, sizeof(kgo_pagg));
  p->threads = threads;
  p->has_key = key_type_like != NULL;
  p->naggs = naggs;
  pthread_mutex_init(&p->mu, NULL);
  p->tl = (kgo_agg **)calloc((size_t)threads * KGO_PAGG_PARTS, sizeof(kgo_agg *));
  p->off = (int64_t *)calloc(KGO_PAGG_PARTS + 1, sizeof(int64_t));
  for (int k = 0; k < threads * KGO_PAGG_PARTS; k++) {
    p->tl[k] = kgo_agg_new(key_type_like, ninputs, input_types, naggs, aggs, ansi);
    if (!p->tl[k]) {
      kgo_pagg_free(p);
      return NULL;
    }
    p->tl[k]->exact_f64 = exact_f64;
  }
  return p;
}

void kgo_pagg_free(kgo_pagg *p) {
  if (!p) return;
  for (int k = 0; k < p->threads * KGO_PAGG_PARTS; k++) kgo_agg_free(p->tl[k]);
  free(p->tl);
  free(p->off);
  pthread_mutex_destroy(&p->mu);
  free(p);
}

const char *kgo_pagg_error(const kgo_pagg *p) { return p->err; }

int kgo_pagg_update(kgo_pagg *p, const kgo_col *key, const kgo_col *inputs, int64_t n) {
  if (p->finalized) return set_err("INTERNAL_ERROR", "update after finalize");
  int T = p->threads;
  pthread_t *th = (pthread_t *)calloc((size_t)T, sizeof(pthread_t));
  pagg_job *jobs = (pagg_job *)calloc((size_t)T, sizeof(pagg_job));
  for (int t = 0; t < T; t++) {
    jobs[t] = (pagg_job){p, t, key, inputs, n * t / T, n * (t + 1) / T, 0};
    pthread_create(&th[t], NULL, pagg_update_job, &jobs[t]);
  }
  int rc = 0;
  for (int t = 0; t < T; t++) {
    pthread_join(th[t], NULL);
    rc |= jobs[t].rc;
  }
  free(th);
  free(jobs);
  if (rc) snprintf(g_err, sizeof g_err, "%s", p->err);
  return rc ? -1 : 0;
}

typedef struct {
  kgo_pagg *p;
  int part;
  int64_t *out_key;
  uint8_t *out_key_valid;
  const kgo_out *outs;
  int rc;
} pagg_part_job;

static void *pagg_merge_job(void *arg) {
  pagg_part_job *j = (pagg_part_job *)arg;
  kgo_pagg *p = j->p;
  kgo_agg *dst = p->tl[j->part];
  for (int t = 1; t < p->threads; t++) {
    kgo_agg *src = p->tl[(size_t)t * KGO_PAGG_PARTS + j->part];
    if (kgo_agg_merge(dst, src) != 0) {
      j->rc = -1;
      pagg_fail(p);
      return NULL;
    }
    /* the thread's partition is merged: its memory goes back now */
    kgo_agg *fresh = kgo_agg_new(NULL, 0, NULL, 0, NULL, 0);
    kgo_agg_free(src);
    p->tl[(size_t)t * KGO_PAGG_PARTS + j->part] = fresh;
  }
  return NULL;
}

static int run_parts(kgo_pagg *p, void *(*fn)(void *), pagg_part_job *jobs) {
  pthread_t th[KGO_PAGG_PARTS];
  int rc = 0;
  /* at most `threads` partitions at once */
  for (int a = 0; a < KGO_PAGG_PARTS; a += p->threads) {
    int b = a + p->threads < KGO_PAGG_PARTS ? a + p->threads : KGO_PAGG_PARTS;
    for (int q = a; q < b; q++) pthread_create(&th[q], NULL, fn, &jobs[q]);
    for (int q = a; q < b; q++) {
      pthread_join(th[q], NULL);
      rc |= jobs[q].rc;
    }
  }
  if (rc) snprintf(g_err, sizeof g_err, "%s", p->err);
  return rc ? -1 : 0;
}

int kgo_pagg_finalize(kgo_pagg *p) {
  if (p->finalized) return 0;
  pagg_part_job jobs[KGO_PAGG_PARTS];
  for (int q = 0; q < KGO_PAGG_PARTS; q++) jobs[q] = (pagg_part_job){p, q, NULL, NULL, NULL, 0};
  if (run_parts(p, pagg_merge_job, jobs) != 0) return -1;
  p->off[0] = 0;
  for (int q = 0; q < KGO_PAGG_PARTS; q++) p->off[q + 1] = p->off[q] + p->tl[q]->ngroups;
  p->finalized = 1;
  return 0;
}

int64_t kgo_pagg_groups(const kgo_pagg *p) { return p->finalized ? p->off[KGO_PAGG_PARTS] : -1; }

static void *pagg_finish_job(void *arg) {
  pagg_part_job *j = (pagg_part_job *)arg;
  kgo_pagg *p = j->p;
  const int64_t o = p->off[j->part];
  kgo_out outs[64];
  int na = p->naggs;
  for (int k = 0; k < na; k++) {
    outs[k] = j->outs[k];
    if (outs[k].i64) outs[k].i64 += o;
    if (outs[k].hi) outs[k].hi += o;
    if (outs[k].f64) outs[k].f64 += o;
    if (outs[k].valid) outs[k].valid += o;
  }
  if (kgo_agg_finish(p->tl[j->part], j->out_key ? j->out_key + o : NULL,
                     j->out_key_valid ? j->out_key_valid + o : NULL, outs) != 0) {
    j->rc = -1;
    pagg_fail(p);
  }
  return NULL;
}

int kgo_pagg_finish(kgo_pagg *p, int64_t *out_key, uint8_t *out_key_valid, kgo_out *outs) {
  if (!p->finalized) return set_err("INTERNAL_ERROR", "finish before finalize");
  if (p->naggs > 64) return set_err("KRISHIV_UNSUPPORTED_FEATURE", "more than 64 aggregates");
  pagg_part_job jobs[KGO_PAGG_PARTS];
  for (int q = 0; q < KGO_PAGG_PARTS; q++)
    jobs[q] = (pagg_part_job){p, q, out_key, out_key_valid, outs, 0};
  return run_parts(p, pagg_finish_job, jobs);
}

void kgo_pagg_result_type(const kgo_pagg *p, int i, int32_t *type, int32_t *precision,
                          int32_t *scale) {
  kgo_agg_result_type(p->tl[0], i, type, precision, scale);
}

/* ------------------------------------------------------------------ */
/* Hash join                                                            */
/* ------------------------------------------------------------------ */
/* A join key value normalized for equality (Spark key normalization:
 * -0.0 is 0.0, every NaN is one NaN; join/mod.rs:24-27). */
static i128 key_norm(const kgo_col *c, int64_t i) {
  if (c->type == KGO_T_FLOAT64) {
    double f = col_f64(c, i);
    uint64_t bits;
    if (f != f) bits = 0x7FF8000000000000ULL;
    else {
      f = f + 0.0;
      memcpy(&bits, &f, 8);
    }
    return (i128)bits;
  }
  if (c->type == KGO_T_FLOAT32) {
    float f = ((const float *)c->values)[i + c->offset];
    uint32_t bits;
    if (f != f) bits = 0x7FC00000u;
    else {
      f = f + 0.0f;
      memcpy(&bits, &f, 4);
    }
    return (i128)bits;
  }
  return col_int(c, i);
}

/* Whether probe row i and build row j match on every key: `=` keys never
 * match a NULL, `<=>` keys match NULL to NULL (join/mod.rs:21-27). */
static int keys_match(int nkeys, const kgo_col *pk, const kgo_col *bk, const uint8_t *ns,
                      int64_t i, int64_t j) {
  for (int k = 0; k < nkeys; k++) {
    int pv = col_valid(&pk[k], i), bv = col_valid(&bk[k], j);
    if (!pv || !bv) {
      if (ns && ns[k] && !pv && !bv) continue;
      return 0;
    }
    if (key_norm(&pk[k], i) != key_norm(&bk[k], j)) return 0;
  }
  return 1;
}

static uint64_t row_key_hash(int nkeys, const kgo_col *keys, int64_t i) {
  uint64_t h = 0x12345;
  for (int k = 0; k < nkeys; k++) {
    i128 v = col_valid(&keys[k], i) ? key_norm(&keys[k], i) : 0;
    h = kgo_mix64(h ^ (uint64_t)v ^ kgo_mix64((uint64_t)((u128)v >> 64)) ^
                  (col_valid(&keys[k], i) ? 0 : 0xABCDEFULL));
  }
  return h;
}

/* The residual filter (join/mod.rs:50-51, probe.rs:686+) as a postfix
 * program over the candidate pair's columns (probe columns, then build
 * columns), Spark's three-valued logic: a pair passes when the program
 * yields TRUE; NULL and FALSE do not pass.  Integer-like columns are read as
 * int64; + - * wrap at 64 bits (non-ANSI); % takes the dividend's sign and
 * is NULL for a zero divisor. */
typedef struct {
  int64_t v;
  int null;
} xval;

static int expr_pass(const kgo_expr *x, int nx, const kgo_col *pc, const kgo_col *bc, int64_t i,
                     int64_t j) {
  xval st[32];
  int sp = 0;
  for (int k = 0; k < nx; k++) {
    const kgo_expr *e = &x[k];
    if (e->op == KGO_X_PROBE || e->op == KGO_X_BUILD) {
      const kgo_col *c = e->op == KGO_X_PROBE ? &pc[e->arg] : &bc[e->arg];
      int64_t r = e->op == KGO_X_PROBE ? i : j;
      int ok = col_valid(c, r);
      st[sp++] = (xval){ok ? (int64_t)col_int(c, r) : 0, !ok};
      continue;
    }
    if (e->op == KGO_X_CONST) {
      st[sp++] = (xval){e->imm, 0};
      continue;
    }
    if (e->op == KGO_X_NOT) {
      st[sp - 1].v = !st[sp - 1].v;
      continue;
    }
    if (e->op == KGO_X_IS_NULL) {
      st[sp - 1] = (xval){st[sp - 1].null, 0};
      continue;
    }
    xval b = st[--sp], a = st[sp - 1], r = {0, 0};
    uint64_t ua = (uint64_t)a.v, ub = (uint64_t)b.v;
    switch (e->op) {
      case KGO_X_AND: /* Kleene: FALSE wins over NULL */
        if ((!a.null && !a.v) || (!b.null && !b.v)) r = (xval){0, 0};
        else r = (xval){1, a.null || b.null};
        break;
      case KGO_X_OR: /* TRUE wins over NULL */
        if ((!a.null && a.v) || (!b.null && b.v)) r = (xval){1, 0};
        else r = (xval){0, a.null || b.null};
        break;
      default:
        if (a.null || b.null) {
          r = (xval){0, 1};
          break;
        }
        switch (e->op) {
          case KGO_X_ADD: r.v = (int64_t)(ua + ub); break;
          case KGO_X_SUB: r.v = (int64_t)(ua - ub); break;
          case KGO_X_MUL: r.v = (int64_t)(ua * ub); break;
          case KGO_X_MOD:
            if (b.v == 0) r.null = 1;
            else if (b.v == -1) r.v = 0;
            else r.v = a.v % b.v;
            break;
          case KGO_X_EQ: r.v = a.v == b.v; break;
          case KGO_X_NE: r.v = a.v != b.v; break;
          case KGO_X_LT: r.v = a.v < b.v; break;
          case KGO_X_LE: r.v = a.v <= b.v; break;
          case KGO_X_GT: r.v = a.v > b.v; break;
          case KGO_X_GE: r.v = a.v >= b.v; break;
        }
    }
    st[sp - 1] = r;
  }
  return sp == 1 && !st[0].null && st[0].v != 0;
}

/* Spark's join types (join/mod.rs:1-60,217-249): pairs and tails as the
 * probe (probe.rs:266-501,532-599) and the build-side tail (emit_build,
 * probe.rs:611-632) return them.  Output (probe row, build row): -1 pads an
 * unmatched side; semi/anti/null-aware anti return (probe row, -1); right
 * semi/anti (-1, build row); existence (probe row, 1 or 0 = exists).  The
 * probe rows come first, in order, then the build-side tail in build order. */
int64_t kgo_join_ex(int join_type, int nkeys, const kgo_col *probe_keys, const kgo_col *build_keys,
                    const uint8_t *null_safe, int64_t n_probe, int64_t n_build, int nested,
                    int nfilter, const kgo_expr *filter, const kgo_col *probe_cols,
                    const kgo_col *build_cols, int64_t cap, int64_t *out_probe,
                    int64_t *out_build) {
  int64_t nout = 0;
  int64_t nslots = 16, *heads = NULL, *next = NULL;
  uint8_t *bmatch = (uint8_t *)calloc((size_t)(n_build ? n_build : 1), 1);
  int build_has_null_key = 0;
  if (!nested) {
    while (nslots < 2 * n_build) nslots *= 2;
    heads = (int64_t *)malloc(sizeof(int64_t) * (size_t)nslots);
    for (int64_t s = 0; s < nslots; s++) heads[s] = -1;
    next = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_build ? n_build : 1));
  }
  for (int64_t j = 0; j < n_build; j++) {
    /* a build row whose `=` key is NULL matches nothing: not indexed */
    int drop = 0;
    for (int k = 0; k < nkeys; k++)
      if (!col_valid(&build_keys[k], j)) {
        if (!(null_safe && null_safe[k])) drop = 1;
        build_has_null_key = 1;
      }
    if (nested) continue;
    next[j] = -1;
    if (drop) continue;
    int64_t s = (int64_t)(row_key_hash(nkeys, build_keys, j) & (uint64_t)(nslots - 1));
    next[j] = heads[s];
    heads[s] = j;
  }
  const int pairs = join_type == KGO_JOIN_INNER || join_type == KGO_JOIN_LEFT_OUTER ||
                    join_type == KGO_JOIN_RIGHT_OUTER || join_type == KGO_JOIN_FULL_OUTER;
  const int marks_build = join_type >= KGO_JOIN_RIGHT_SEMI;
#define EMIT(a, b)              \
  do {                          \
    if (nout >= cap) goto full; \
    out_probe[nout] = (a);      \
    out_build[nout++] = (b);    \
  } while (0)
  for (int64_t i = 0; i < n_probe; i++) {
    int64_t matches = 0;
    int64_t j = nested ? 0 : heads[row_key_hash(nkeys, probe_keys, i) & (uint64_t)(nslots - 1)];
    while (nested ? j < n_build : j >= 0) {
      if (keys_match(nkeys, probe_keys, build_keys, null_safe, i, j) &&
          (nfilter == 0 || expr_pass(filter, nfilter, probe_cols, build_cols, i, j))) {
        matches++;
        if (marks_build) bmatch[j] = 1;
        if (pairs) EMIT(i, j);
        else if (!marks_build) break; /* semi/anti/existence: one match is enough */
      }
      j = nested ? j + 1 : next[j];
    }
    int pkey_null = 0;
    for (int k = 0; k < nkeys; k++) pkey_null |= !col_valid(&probe_keys[k], i);
    switch (join_type) {
      case KGO_JOIN_LEFT_OUTER:
      case KGO_JOIN_FULL_OUTER:
        if (!matches) EMIT(i, -1);
        break;
      case KGO_JOIN_LEFT_SEMI:
        if (matches) EMIT(i, -1);
        break;
      case KGO_JOIN_LEFT_ANTI:
        if (!matches) EMIT(i, -1);
        break;
      case KGO_JOIN_NULL_AWARE_ANTI:
        /* NOT IN: an empty build keeps every row; a NULL build key keeps
         * none; otherwise a non-NULL key that matches nothing */
        if (n_build == 0 || (!build_has_null_key && !pkey_null && !matches)) EMIT(i, -1);
        break;
      case KGO_JOIN_EXISTENCE:
        EMIT(i, matches ? 1 : 0);
        break;
      default:
        break;
    }
  }
  if (marks_build) {
    for (int64_t j = 0; j < n_build; j++) {
      if (join_type == KGO_JOIN_RIGHT_SEMI ? bmatch[j] : !bmatch[j]) EMIT(-1, j);
    }
  }
#undef EMIT
  free(heads);
  free(next);
  free(bmatch);
  return nout;
full:
  free(heads);
  free(next);
  free(bmatch);
  return -1;
}

int64_t kgo_join(int join_type, int nkeys, const kgo_col *probe_keys, const kgo_col *build_keys,
                 const uint8_t *null_safe, int64_t n_probe, int64_t n_build, int nested,
                 int64_t cap, int64_t *out_probe, int64_t *out_build) {
  return kgo_join_ex(join_type, nkeys, probe_keys, build_keys, null_safe, n_probe, n_build, nested,
                     0, NULL, NULL, NULL, cap, out_probe, out_build);
}

/* ------------------------------------------------------------------ */
/* Grace join fused with the per-partition aggregate (configuration 4's  */
/* plan): SELECT p, COUNT(*), SUM(probe.pay * build.pay) ... GROUP BY p,  */
/* p = partition_of(H_kg(key), parts) (routing.rs:222-231), inner join on */
/* `=` int64 keys (NULL keys never match).  Per key the pairs' count is   */
/* nb * np and their payload-product sum Sb * Sp (pairs whose payloads    */
/* are both non-NULL), so no pair is walked.  mode 0: the 64-bit wrapping */
/* sum; mode 1: that sum mod 2^31.  Pinned by the pair walk (kgo_join)    */
/* on small inputs.                                                       */
/* ------------------------------------------------------------------ */
typedef struct {
  int64_t key;
  int64_t nb, nbv; /* build rows, build rows with a payload */
  uint64_t sb;     /* wrapping sum of their payloads */
  int used;
} gkey;

int kgo_grace_factorized(const kgo_col *bkey, const kgo_col *bpay, int64_t nb, const kgo_col *pkey,
                         const kgo_col *ppay, int64_t np, uint32_t parts, int mode,
                         int64_t *counts, int64_t *sums, uint8_t *sum_valid) {
  int64_t cap = 16;
  while (cap < 2 * nb) cap *= 2;
  gkey *t = (gkey *)calloc((size_t)cap, sizeof(gkey));
  if (!t) return set_err("UNABLE_TO_ACQUIRE_MEMORY", "grace oracle table");
  for (int64_t i = 0; i < nb; i++) {
    if (!col_valid(bkey, i)) continue;
    int64_t k = (int64_t)col_int(bkey, i);
    uint64_t h = key_slot_hash(k) & (uint64_t)(cap - 1);
    while (t[h].used && t[h].key != k) h = (h + 1) & (uint64_t)(cap - 1);
    t[h].used = 1;
    t[h].key = k;
    t[h].nb++;
    if (col_valid(bpay, i)) {
      t[h].nbv++;
      t[h].sb += (uint64_t)(int64_t)col_int(bpay, i);
    }
  }
  uint64_t *s = (uint64_t *)calloc(parts, sizeof(uint64_t));
  int64_t *vc = (int64_t *)calloc(parts, sizeof(int64_t));
  for (uint32_t q = 0; q < parts; q++) counts[q] = 0;
  for (int64_t i = 0; i < np; i++) {
    if (!col_valid(pkey, i)) continue;
    int64_t k = (int64_t)col_int(pkey, i);
    uint64_t h = key_slot_hash(k) & (uint64_t)(cap - 1);
    while (t[h].used && t[h].key != k) h = (h + 1) & (uint64_t)(cap - 1);
    if (!t[h].used) continue;
    uint8_t b[8];
    for (int q = 0; q < 8; q++) b[q] = (uint8_t)((uint64_t)k >> (8 * q));
    uint32_t p = kgo_partition_of(kgo_xxh64(b, 8, 0), parts);
    counts[p] += t[h].nb;
    if (col_valid(ppay, i)) {
      s[p] += (uint64_t)(int64_t)col_int(ppay, i) * t[h].sb;
      vc[p] += t[h].nbv;
    }
  }
  for (uint32_t q = 0; q < parts; q++) {
    sums[q] = (int64_t)(mode == 1 ? (s[q] & 0x7FFFFFFFull) : s[q]);
    sum_valid[q] = vc[q] > 0;
  }
  free(t);
  free(s);
  free(vc);
  return 0;
}
