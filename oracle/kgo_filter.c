/*
 * kgo_filter.c -- CPU oracle of the runtime KeyFilter: the builder's forms,
 * evaluation on rows, the versioned encoding, the merge of published
 * filters, and the scan row stage's ordering and adaptive switch-off.
 *
 * TEST INFRASTRUCTURE ONLY (see kgo.h): the checker of libkgpu's device
 * filter, never part of the product path.
 *
 * Restates (paths relative to the reference repository root):
 *   crates/krishiv-io/src/runtime_filter/filter.rs:180-615  KeyFilterBuilder
 *   crates/krishiv-io/src/runtime_filter/member.rs:1-560    forms, evaluate
 *   crates/krishiv-io/src/runtime_filter/key.rs:70-330      widening, hashing
 *   crates/krishiv-io/src/runtime_filter/codec.rs:1-461     encode / decode
 *   crates/krishiv-io/src/runtime_filter/merge.rs:35-169    KeyFilter::merge
 *   crates/krishiv-io/src/runtime_filter/bloom.rs:150-180   union_with, pass_rate
 *   crates/krishiv-exec/src/operators/runtime_filter.rs:240-395  RowStage
 * One deliberate difference: an exact set is encoded with its members in
 * ascending order (the reference writes its hash table's iteration order;
 * its decoder accepts any order and rebuilds the same set).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "kgo.h"

typedef __int128 i128;

enum { CLS_INT = 0, CLS_FLOAT = 1, CLS_DEC = 2 };

struct kgo_kf {
  int type, precision, scale;
  int cls, width, single;
  int tz_len; /* -1: no time zone */
  char tz[64];
  kgo_filter_opts opt;
  /* builder */
  int has_range, nan;
  i128 lo, hi;
  double flo, fhi;
  int distinct_ok;
  i128 *dist;
  int64_t ndist, capdist;
  int hashes_ok;
  uint64_t *hashes;
  int64_t nh, caph;
  int nulls;
  uint64_t rows;
  /* the filter (after finish or decode) */
  int finished;
  int values_ok;
  i128 *values;
  int64_t nvalues;
  int member;
  i128 *members; /* small / exact / dense members */
  int64_t nmembers;
  int64_t dense_lo;
  uint64_t *dense_bits;
  int64_t dense_words;
  uint32_t *bloom;
  size_t nblocks;
  int empty;
};

/* ------------------------------------------------------------ key types */

/* Class::of (member.rs:68-86) and the key type's width. */
static int class_of(int t, int *width, int *single) {
  *single = 0;
  switch (t) {
    case KGO_T_BOOL: case KGO_T_INT8: *width = 1; return CLS_INT;
    case KGO_T_INT16: *width = 2; return CLS_INT;
    case KGO_T_INT32: case KGO_T_DATE32: *width = 4; return CLS_INT;
    case KGO_T_INT64: case KGO_T_TIMESTAMP: *width = 8; return CLS_INT;
    case KGO_T_FLOAT32: *width = 4; *single = 1; return CLS_FLOAT;
    case KGO_T_FLOAT64: *width = 8; return CLS_FLOAT;
    case KGO_T_DECIMAL64: case KGO_T_DECIMAL128: *width = 16; return CLS_DEC;
    default: return -1;
  }
}

static int int_rank(int t) {
  switch (t) {
    case KGO_T_INT8: return 1;
    case KGO_T_INT16: return 2;
    case KGO_T_INT32: return 3;
    case KGO_T_INT64: return 4;
    default: return 0;
  }
}

static int int_digits(int t) {
  switch (t) {
    case KGO_T_INT8: return 3;
    case KGO_T_INT16: return 5;
    case KGO_T_INT32: return 10;
    case KGO_T_INT64: return 19;
    default: return 0;
  }
}

static int is_dec(int t) { return t == KGO_T_DECIMAL64 || t == KGO_T_DECIMAL128; }

/* widens_to (key.rs:75-103) for the fixed-width types. */
int kgo_kf_widens(int from, int fp, int fs, int to, int tp, int ts) {
  int w, s;
  if (class_of(to, &w, &s) < 0) return 0;
  if (from == to && (!is_dec(to) || (fp == tp && fs == ts))) return 1;
  if (to == KGO_T_FLOAT64 &&
      (from == KGO_T_FLOAT32 || from == KGO_T_INT8 || from == KGO_T_INT16 || from == KGO_T_INT32))
    return 1;
  if (to == KGO_T_FLOAT32 && (from == KGO_T_INT8 || from == KGO_T_INT16)) return 1;
  if (is_dec(from) && is_dec(to)) return ts >= fs && tp - ts >= fp - fs;
  if (is_dec(to)) return int_digits(from) > 0 && tp - ts >= int_digits(from);
  return int_rank(from) > 0 && int_rank(to) > 0 && int_rank(from) < int_rank(to);
}

static int col_valid(const kgo_col *c, int64_t i) {
  if (!c->validity) return 1;
  int64_t j = i + c->offset;
  return (c->validity[j >> 3] >> (j & 7)) & 1;
}

static int64_t col_int(const kgo_col *c, int64_t i) {
  int64_t j = i + c->offset;
  switch (c->type) {
    case KGO_T_BOOL: return (((const uint8_t *)c->values)[j >> 3] >> (j & 7)) & 1;
    case KGO_T_INT8: return ((const int8_t *)c->values)[j];
    case KGO_T_INT16: return ((const int16_t *)c->values)[j];
    case KGO_T_INT32: case KGO_T_DATE32: return ((const int32_t *)c->values)[j];
    default: return ((const int64_t *)c->values)[j];
  }
}

static i128 col_dec(const kgo_col *c, int64_t i) {
  int64_t j = i + c->offset;
  if (c->type == KGO_T_DECIMAL128) {
    const uint64_t *w = (const uint64_t *)c->values + 2 * j;
    return (i128)(((unsigned __int128)w[1] << 64) | w[0]);
  }
  if (c->type == KGO_T_DECIMAL64) return ((const int64_t *)c->values)[j];
  return col_int(c, i);
}

/* spark_float (key.rs:146-154): one NaN, -0.0 as 0.0 */
static double spark_float(double x) {
  if (x != x) return NAN;
  if (x == 0.0) return 0.0;
  return x;
}

static double col_float(const kgo_col *c, int64_t i) {
  int64_t j = i + c->offset;
  if (c->type == KGO_T_FLOAT32) return spark_float((double)((const float *)c->values)[j]);
  if (c->type == KGO_T_FLOAT64) return spark_float(((const double *)c->values)[j]);
  return (double)col_int(c, i);
}

static uint64_t f64_bits(double v) {
  uint64_t b;
  memcpy(&b, &v, 8);
  return b;
}

static uint32_t f32_bits(float v) {
  uint32_t b;
  memcpy(&b, &v, 4);
  return b;
}

/* float_bits (key.rs:363-371) */
static i128 float_bits(double v, int single) {
  return single ? (i128)f32_bits((float)v) : (i128)f64_bits(v);
}

/* int_hash / decimal_hash / float_hash (key.rs:290-330) */
static uint64_t int_hash(int64_t v, int width) {
  uint64_t u = (uint64_t)v;
  if (width == 1) return kgo_xxh64_fixed(u & 0xFF, 0, 1);
  if (width == 2) return kgo_xxh64_fixed(u & 0xFFFF, 0, 2);
  if (width == 4) return kgo_xxh64_fixed(u & 0xFFFFFFFFULL, 0, 4);
  return kgo_xxh64_fixed(u, 0, 8);
}

static uint64_t dec_hash(i128 v) {
  unsigned __int128 u = (unsigned __int128)v;
  return kgo_xxh64_fixed((uint64_t)u, (uint64_t)(u >> 64), 16);
}

static uint64_t float_hash(double v, int single) {
  if (single) return kgo_xxh64_fixed(f32_bits((float)v), 0, 4);
  return kgo_xxh64_fixed(f64_bits(v), 0, 8);
}

/* Class::hash64 (member.rs:89-105) / decimal_hash of a member value */
static uint64_t member_hash(const kgo_kf *f, i128 m) {
  if (f->cls == CLS_DEC) return dec_hash(m);
  if (f->cls == CLS_FLOAT) {
    if (f->single) return kgo_xxh64_fixed((uint64_t)m & 0xFFFFFFFFULL, 0, 4);
    return kgo_xxh64_fixed((uint64_t)m, 0, 8);
  }
  return int_hash((int64_t)m, f->width);
}

/* ------------------------------------------------------------ helpers */

static int cmp_i128(const void *a, const void *b) {
  i128 x = *(const i128 *)a, y = *(const i128 *)b;
  return x < y ? -1 : x > y;
}

/* index of v in the sorted set, or -(insertion point) - 1 */
static int64_t find_sorted(const i128 *s, int64_t n, i128 v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t m = (lo + hi) / 2;
    if (s[m] < v) lo = m + 1;
    else hi = m;
  }
  if (lo < n && s[lo] == v) return lo;
  return -lo - 1;
}

static void *grow(void *p, int64_t *cap, int64_t need, size_t elem) {
  if (need <= *cap) return p;
  int64_t c = *cap ? *cap : 16;
  while (c < need) c *= 2;
  *cap = c;
  return realloc(p, (size_t)c * elem);
}

/* Distinct::insert (filter.rs:125-140): false once more than max keys. */
static int dist_insert(kgo_kf *f, i128 v, int64_t max) {
  int64_t at = find_sorted(f->dist, f->ndist, v);
  if (at >= 0) return 1;
  if (f->ndist >= max) return 0;
  at = -at - 1;
  f->dist = grow(f->dist, &f->capdist, f->ndist + 1, sizeof(i128));
  memmove(f->dist + at + 1, f->dist + at, (size_t)(f->ndist - at) * sizeof(i128));
  f->dist[at] = v;
  f->ndist++;
  return 1;
}

static int64_t distinct_max(const kgo_kf *f) {
  return f->opt.max_values > f->opt.exact_max_values ? f->opt.max_values : f->opt.exact_max_values;
}

/* ------------------------------------------------------------ builder */

void kgo_filter_default_opts(kgo_filter_opts *o) {
  memset(o, 0, sizeof *o);
  o->max_values = 50;          /* filter.rs:92-107 */
  o->exact_max_values = 4096;  /* member.rs:39 EXACT_SET_MAX */
  o->bloom = 1;
  o->bloom_max_items = 4000000;
  o->bloom_max_bytes = 8u << 20;
  o->bloom_fpp = 0.01;
  o->bloom_max_fpp = 0.05;
  o->bloom_bytes = 0;
  o->side = KGO_SIDE_BOTH;
}

static kgo_kf *kf_alloc(int type, int precision, int scale, const kgo_filter_opts *o) {
  int w, s;
  int cls = class_of(type, &w, &s);
  if (cls < 0) return NULL;
  kgo_kf *f = calloc(1, sizeof *f);
  f->type = type;
  f->precision = precision;
  f->scale = scale;
  f->cls = cls;
  f->width = w;
  f->single = s;
  f->tz_len = -1;
  if (o) f->opt = *o;
  else kgo_filter_default_opts(&f->opt);
  return f;
}

/* KeyFilterBuilder::new (filter.rs:180-206) */
kgo_kf *kgo_kf_new(int type, int precision, int scale, const kgo_filter_opts *o) {
  kgo_kf *f = kf_alloc(type, precision, scale, o);
  if (!f) return NULL;
  int members = f->opt.side == KGO_SIDE_BOTH;
  f->hashes_ok = f->opt.bloom && members;
  f->distinct_ok = members && distinct_max(f) > 0;
  return f;
}

void kgo_kf_free(kgo_kf *f) {
  if (!f) return;
  free(f->dist);
  free(f->hashes);
  free(f->values);
  free(f->members);
  free(f->dense_bits);
  free(f->bloom);
  free(f);
}

/* normalized value of row i of a column of the key's class */
static i128 norm_member(const kgo_kf *f, const kgo_col *c, int64_t i, double *fv) {
  if (f->cls == CLS_FLOAT) {
    *fv = col_float(c, i);
    return float_bits(*fv, f->single);
  }
  if (f->cls == CLS_DEC) return col_dec(c, i);
  return col_int(c, i);
}

static uint64_t row_hash_of(const kgo_kf *f, i128 v, double fv) {
  if (f->cls == CLS_FLOAT) return float_hash(fv, f->single);
  if (f->cls == CLS_DEC) return dec_hash(v);
  return int_hash((int64_t)v, f->width);
}

/* KeyFilterBuilder::add_hashed (filter.rs:225-284), add_range (286-336),
 * add_values (339-384).  `include`: LSB bitmap of kept rows (NULL: all). */
int kgo_kf_add(kgo_kf *f, const kgo_col *c, const uint8_t *include, int64_t n) {
  if (f->finished) return -1;
  if (c->type != f->type) return -1;
  int64_t kept = 0;
  int64_t *rows = malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) {
    if (include && !((include[i >> 3] >> (i & 7)) & 1)) continue;
    if (!col_valid(c, i)) f->nulls = 1;
    else rows[kept++] = i;
  }
  if (kept == 0) {
    free(rows);
    return 0;
  }
  f->rows += (uint64_t)kept;
  /* add_range */
  if (f->cls == CLS_FLOAT) {
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t k = 0; k < kept; k++) {
      double v = col_float(c, rows[k]);
      if (v != v) f->nan = 1;
      lo = fmin(lo, v);
      hi = fmax(hi, v);
    }
    if (!f->has_range) {
      f->flo = lo;
      f->fhi = hi;
    } else {
      f->flo = fmin(f->flo, lo);
      f->fhi = fmax(f->fhi, hi);
    }
  } else {
    i128 lo = 0, hi = 0;
    for (int64_t k = 0; k < kept; k++) {
      i128 v = f->cls == CLS_DEC ? col_dec(c, rows[k]) : (i128)col_int(c, rows[k]);
      if (k == 0 || v < lo) lo = v;
      if (k == 0 || v > hi) hi = v;
    }
    if (!f->has_range) {
      f->lo = lo;
      f->hi = hi;
    } else {
      if (lo < f->lo) f->lo = lo;
      if (hi > f->hi) f->hi = hi;
    }
  }
  f->has_range = 1;
  if (f->hashes_ok && f->rows > f->opt.bloom_max_items) {
    f->hashes_ok = 0;
    free(f->hashes);
    f->hashes = NULL;
    f->nh = f->caph = 0;
  }
  if (!f->distinct_ok && !f->hashes_ok) {
    free(rows);
    return 0;
  }
  int64_t max = distinct_max(f);
  for (int64_t k = 0; k < kept; k++) {
    double fv = 0;
    i128 v = norm_member(f, c, rows[k], &fv);
    if (f->distinct_ok && !dist_insert(f, v, max)) {
      f->distinct_ok = 0;
      free(f->dist);
      f->dist = NULL;
      f->ndist = f->capdist = 0;
    }
    if (f->hashes_ok) {
      f->hashes = grow(f->hashes, &f->caph, f->nh + 1, 8);
      f->hashes[f->nh++] = row_hash_of(f, v, fv);
    }
  }
  free(rows);
  return 0;
}

/* KeyFilterBuilder::merge (filter.rs:478-510) */
int kgo_kf_merge_builder(kgo_kf *a, const kgo_kf *b) {
  if (a->finished || b->finished || a->type != b->type) return -1;
  if (b->has_range) {
    if (!a->has_range) {
      a->lo = b->lo;
      a->hi = b->hi;
      a->flo = b->flo;
      a->fhi = b->fhi;
    } else {
      if (b->lo < a->lo) a->lo = b->lo;
      if (b->hi > a->hi) a->hi = b->hi;
      a->flo = fmin(a->flo, b->flo);
      a->fhi = fmax(a->fhi, b->fhi);
    }
    a->has_range = 1;
  }
  a->nan |= b->nan;
  a->nulls |= b->nulls;
  a->rows += b->rows;
  if (a->distinct_ok && b->distinct_ok) {
    int64_t max = distinct_max(a);
    for (int64_t k = 0; k < b->ndist; k++)
      if (!dist_insert(a, b->dist[k], max)) {
        a->distinct_ok = 0;
        break;
      }
  } else {
    a->distinct_ok = 0;
  }
  if (!a->distinct_ok) {
    free(a->dist);
    a->dist = NULL;
    a->ndist = a->capdist = 0;
  }
  int keep = a->rows <= a->opt.bloom_max_items;
  if (a->hashes_ok && b->hashes_ok && keep) {
    a->hashes = grow(a->hashes, &a->caph, a->nh + b->nh, 8);
    memcpy(a->hashes + a->nh, b->hashes, (size_t)b->nh * 8);
    a->nh += b->nh;
  } else {
    a->hashes_ok = 0;
    free(a->hashes);
    a->hashes = NULL;
    a->nh = a->caph = 0;
  }
  return 0;
}

/* Dense::of (member.rs:150-176): the bit set of sorted distinct ints. */
static int dense_of(kgo_kf *f, const i128 *v, int64_t n) {
  if (n == 0) return 0;
  i128 span = v[n - 1] - v[0] + 1;
  if (span > (i128)KGO_DENSE_MAX_BITS) return 0;
  int64_t words = (int64_t)((span + 63) / 64);
  int64_t hashed = n * KGO_HASH_SET_BYTES_PER_KEY;
  int64_t allow = hashed * KGO_DENSE_SIZE_FACTOR;
  if (allow < KGO_DENSE_FREE_BYTES) allow = KGO_DENSE_FREE_BYTES;
  if (words * 8 > allow) return 0;
  f->dense_lo = (int64_t)v[0];
  f->dense_words = words;
  f->dense_bits = calloc((size_t)words, 8);
  for (int64_t k = 0; k < n; k++) {
    uint64_t d = (uint64_t)((int64_t)v[k]) - (uint64_t)f->dense_lo;
    f->dense_bits[d >> 6] |= 1ULL << (d & 63);
  }
  return 1;
}

/* Member::exact (member.rs:320-375) + Exact::densify (210-219) over the
 * members already in f->members. */
static void set_exact(kgo_kf *f) {
  if (f->nmembers <= KGO_SMALL_SET_MAX) {
    f->member = KGO_FORM_SMALL;
    return;
  }
  f->member = KGO_FORM_EXACT;
  if (f->cls == CLS_INT && dense_of(f, f->members, f->nmembers)) f->member = KGO_FORM_DENSE;
}

static void set_bloom_from_hashes(kgo_kf *f, const uint64_t *h, int64_t n) {
  /* bloom() (filter.rs:606-624) */
  size_t bytes = f->opt.bloom_bytes ? f->opt.bloom_bytes : kgo_sbbf_bytes_for((uint64_t)n, f->opt.bloom_fpp);
  if (bytes > f->opt.bloom_max_bytes) bytes = f->opt.bloom_max_bytes;
  size_t nb = kgo_sbbf_blocks_for_bytes(bytes);
  f->bloom = calloc(nb, 32);
  f->nblocks = nb;
  for (int64_t k = 0; k < n; k++) kgo_sbbf_insert(f->bloom, nb, h[k]);
  if (kgo_sbbf_expected_fpp(nb, (uint64_t)n) <= f->opt.bloom_max_fpp) {
    f->member = KGO_FORM_BLOOM;
  } else {
    free(f->bloom);
    f->bloom = NULL;
    f->nblocks = 0;
    f->member = KGO_FORM_RANGE;
  }
}

/* KeyFilterBuilder::finish (filter.rs:526-602) for an undrained builder;
 * the range opened per RangeSide (key.rs:403-432). */
int kgo_kf_finish(kgo_kf *f) {
  if (f->finished) return -1;
  f->finished = 1;
  int has_range = f->has_range && !f->nan;
  if (has_range && f->opt.side != KGO_SIDE_BOTH) {
    int upper = f->opt.side == KGO_SIDE_UPPER;
    if (f->cls == CLS_INT) {
      i128 lim_lo, lim_hi;
      switch (f->type) {
        case KGO_T_INT8: lim_lo = INT8_MIN; lim_hi = INT8_MAX; break;
        case KGO_T_INT16: lim_lo = INT16_MIN; lim_hi = INT16_MAX; break;
        case KGO_T_INT32: case KGO_T_DATE32: lim_lo = INT32_MIN; lim_hi = INT32_MAX; break;
        case KGO_T_INT64: case KGO_T_TIMESTAMP: lim_lo = INT64_MIN; lim_hi = INT64_MAX; break;
        default: has_range = 0; lim_lo = lim_hi = 0;
      }
      if (upper) f->lo = lim_lo;
      else f->hi = lim_hi;
    } else if (f->cls == CLS_DEC) {
      i128 lim = 1;
      for (int k = 0; k < f->precision; k++) lim *= 10;
      lim -= 1;
      if (upper) f->lo = -lim;
      else f->hi = lim;
    } else {
      if (upper) f->flo = -INFINITY;
      else has_range = 0;
    }
  }
  f->has_range = has_range;
  if (f->distinct_ok) {
    f->nmembers = f->ndist;
    f->members = malloc((size_t)(f->ndist > 0 ? f->ndist : 1) * sizeof(i128));
    if (f->ndist > 0) memcpy(f->members, f->dist, (size_t)f->ndist * sizeof(i128));
    if ((uint64_t)f->ndist <= f->opt.max_values) {
      f->values_ok = 1;
      f->nvalues = f->ndist;
      f->values = malloc((size_t)(f->ndist > 0 ? f->ndist : 1) * sizeof(i128));
      if (f->ndist > 0) memcpy(f->values, f->dist, (size_t)f->ndist * sizeof(i128));
    }
  }
  int dense = f->distinct_ok && f->has_range && f->cls == CLS_INT &&
              f->hi - f->lo + 1 == (i128)f->ndist;
  if (dense) {
    f->member = KGO_FORM_RANGE;
  } else if (f->distinct_ok && (uint64_t)f->ndist <= f->opt.exact_max_values) {
    set_exact(f);
  } else if (f->hashes_ok && f->nh > 0) {
    set_bloom_from_hashes(f, f->hashes, f->nh);
  } else {
    f->member = KGO_FORM_RANGE;
  }
  if (f->member == KGO_FORM_RANGE || f->member == KGO_FORM_BLOOM) {
    free(f->members);
    f->members = NULL;
    f->nmembers = 0;
  }
  f->empty = f->rows == 0;
  return 0;
}

/* ------------------------------------------------------------ inspection */

int kgo_kf_get_info(const kgo_kf *f, kgo_kf_info *o) {
  memset(o, 0, sizeof *o);
  o->form = f->member;
  o->nmembers = f->nmembers;
  o->values = f->values_ok ? f->nvalues : -1;
  o->has_range = f->has_range;
  o->lo_lo = (uint64_t)f->lo;
  o->lo_hi = (uint64_t)((unsigned __int128)f->lo >> 64);
  o->hi_lo = (uint64_t)f->hi;
  o->hi_hi = (uint64_t)((unsigned __int128)f->hi >> 64);
  o->flo = f->flo;
  o->fhi = f->fhi;
  o->nulls_pass = f->nulls;
  o->empty = f->empty;
  o->keys = f->rows;
  o->bloom_bytes = f->nblocks * 32;
  return 0;
}

/* members (small / exact / dense) as (lo, hi) words, ascending */
int64_t kgo_kf_members(const kgo_kf *f, uint64_t *out) {
  if (f->member == KGO_FORM_DENSE) {
    int64_t k = 0;
    for (int64_t w = 0; w < f->dense_words; w++)
      for (int b = 0; b < 64; b++)
        if ((f->dense_bits[w] >> b) & 1) {
          i128 v = (i128)(int64_t)((uint64_t)f->dense_lo + (uint64_t)(w * 64 + b));
          if (out) {
            out[2 * k] = (uint64_t)v;
            out[2 * k + 1] = (uint64_t)((unsigned __int128)v >> 64);
          }
          k++;
        }
    return k;
  }
  for (int64_t k = 0; k < f->nmembers && out; k++) {
    out[2 * k] = (uint64_t)f->members[k];
    out[2 * k + 1] = (uint64_t)((unsigned __int128)f->members[k] >> 64);
  }
  return f->nmembers;
}

int64_t kgo_kf_bloom_words(const kgo_kf *f, uint32_t *out) {
  if (out && f->bloom) memcpy(out, f->bloom, f->nblocks * 32);
  return (int64_t)f->nblocks * 8;
}

/* Member::row_cost_ns (member.rs:410-418) */
double kgo_kf_row_cost(const kgo_kf *f) {
  switch (f->member) {
    case KGO_FORM_RANGE: return 0.5;
    case KGO_FORM_SMALL: return 3.0;
    case KGO_FORM_BLOOM: return 4.5;
    case KGO_FORM_DENSE: return 1.2;
    default: return 5.0;
  }
}

/* ------------------------------------------------------------ evaluate */

static int has_member(const kgo_kf *f, i128 v) {
  for (int64_t k = 0; k < f->nmembers; k++)
    if (f->members[k] == v) return 1;
  return 0;
}

/* KeyFilter::evaluate (filter.rs:764-797) over a column that widens to the
 * key type: member.rs:439-560 range_pass / member_pass.  out[i] = 1 when row
 * i may match.  Returns the rows passing, or -1 when the column does not
 * widen (or a rescaled decimal overflows: key.rs:270-282). */
int64_t kgo_kf_evaluate(const kgo_kf *f, const kgo_col *c, int64_t n, uint8_t *out) {
  if (!kgo_kf_widens(c->type, c->precision, c->scale, f->type, f->precision, f->scale)) return -1;
  i128 factor = 1;
  if (f->cls == CLS_DEC) {
    int from_scale = is_dec(c->type) ? c->scale : 0;
    for (int k = from_scale; k < f->scale; k++) factor *= 10;
    for (int64_t i = 0; i < n; i++) {
      if (!col_valid(c, i)) continue;
      i128 v = col_dec(c, i), r;
      if (__builtin_mul_overflow(v, factor, &r)) return -1;
    }
  }
  int64_t passed = 0;
  for (int64_t i = 0; i < n; i++) {
    int ok;
    if (!col_valid(c, i)) {
      ok = f->nulls;
    } else if (f->empty) {
      ok = 0;
    } else {
      i128 v = 0;
      double fv = 0;
      uint64_t h;
      if (f->cls == CLS_FLOAT) {
        fv = col_float(c, i);
        ok = !f->has_range || (fv >= f->flo && fv <= f->fhi);
        v = float_bits(fv, f->single);
        h = float_hash(fv, f->single);
      } else if (f->cls == CLS_DEC) {
        v = col_dec(c, i) * factor;
        ok = !f->has_range || (v >= f->lo && v <= f->hi);
        h = dec_hash(v);
      } else {
        v = col_int(c, i);
        ok = !f->has_range || (v >= f->lo && v <= f->hi);
        h = int_hash((int64_t)v, f->width);
      }
      if (ok) {
        switch (f->member) {
          case KGO_FORM_SMALL:
          case KGO_FORM_EXACT: ok = has_member(f, v); break;
          case KGO_FORM_DENSE: {
            uint64_t d = (uint64_t)(int64_t)v - (uint64_t)f->dense_lo;
            ok = (int64_t)(d >> 6) < f->dense_words && ((f->dense_bits[d >> 6] >> (d & 63)) & 1);
            break;
          }
          case KGO_FORM_BLOOM: ok = kgo_sbbf_check(f->bloom, f->nblocks, h); break;
          default: break;
        }
      }
    }
    out[i] = (uint8_t)ok;
    passed += ok;
  }
  return passed;
}

/* ------------------------------------------------------------ codec */

typedef struct {
  uint8_t *p;
  int64_t n, cap;
} wbuf;

static void put(wbuf *w, const void *src, int64_t len) {
  if (w->p && w->n + len <= w->cap) memcpy(w->p + w->n, src, (size_t)len);
  w->n += len;
}
static void put_u8(wbuf *w, uint8_t v) { put(w, &v, 1); }
static void put_u32(wbuf *w, uint32_t v) { put(w, &v, 4); }
static void put_u64(wbuf *w, uint64_t v) { put(w, &v, 8); }
static void put_i128(wbuf *w, i128 v) { put(w, &v, 16); }

static int type_tag(int t) {
  switch (t) {
    case KGO_T_BOOL: return 1;
    case KGO_T_INT8: return 2;
    case KGO_T_INT16: return 3;
    case KGO_T_INT32: return 4;
    case KGO_T_INT64: return 5;
    case KGO_T_FLOAT32: return 6;
    case KGO_T_FLOAT64: return 7;
    case KGO_T_DECIMAL128: return 8;
    case KGO_T_DATE32: return 9;
    case KGO_T_TIMESTAMP: return 10;
    case KGO_T_DECIMAL64: return 17;
    default: return 0;
  }
}

/* KeyFilter::encode (codec.rs:104-180).  Returns the encoded length; writes
 * when it fits `cap`. */
int64_t kgo_kf_encode(const kgo_kf *f, uint8_t *out, int64_t cap) {
  wbuf w = {out, 0, cap};
  put_u8(&w, KGO_FILTER_FORMAT_VERSION);
  put_u8(&w, KGO_FILTER_ROW_HASH_ID);
  put_u8(&w, KGO_FILTER_NORMALIZATION_VERSION);
  put_u8(&w, (uint8_t)type_tag(f->type));
  if (is_dec(f->type)) {
    put_u8(&w, (uint8_t)f->precision);
    put_u8(&w, (uint8_t)(int8_t)f->scale);
  } else if (f->type == KGO_T_TIMESTAMP) {
    if (f->tz_len < 0) {
      put_u8(&w, 0);
    } else {
      put_u8(&w, 1);
      put_u32(&w, (uint32_t)f->tz_len);
      put(&w, f->tz, f->tz_len);
    }
  }
  put_u8(&w, (uint8_t)((f->nulls ? 1 : 0) | (f->empty ? 2 : 0)));
  put_u64(&w, f->rows);
  if (!f->has_range) {
    put_u8(&w, 0);
  } else if (f->cls == CLS_INT) {
    put_u8(&w, 1);
    put_u64(&w, (uint64_t)(int64_t)f->lo);
    put_u64(&w, (uint64_t)(int64_t)f->hi);
  } else if (f->cls == CLS_DEC) {
    put_u8(&w, 2);
    put_i128(&w, f->lo);
    put_i128(&w, f->hi);
  } else {
    put_u8(&w, 3);
    put_u64(&w, f64_bits(f->flo));
    put_u64(&w, f64_bits(f->fhi));
  }
  if (!f->values_ok) {
    put_u8(&w, 0);
  } else {
    put_u8(&w, 1);
    put_u32(&w, (uint32_t)f->nvalues);
    for (int64_t k = 0; k < f->nvalues; k++) {
      put_u8(&w, 0);
      put_i128(&w, f->values[k]);
    }
  }
  switch (f->member) {
    case KGO_FORM_RANGE: put_u8(&w, 0); break;
    case KGO_FORM_SMALL:
    case KGO_FORM_EXACT:
      put_u8(&w, f->member == KGO_FORM_SMALL ? 1 : 2);
      put_u32(&w, (uint32_t)f->nmembers);
      for (int64_t k = 0; k < f->nmembers; k++) {
        if (f->cls == CLS_DEC) put_i128(&w, f->members[k]);
        else put_u64(&w, (uint64_t)(int64_t)f->members[k]);
      }
      break;
    case KGO_FORM_DENSE: {
      put_u8(&w, 2);
      int64_t nm = kgo_kf_members(f, NULL);
      put_u32(&w, (uint32_t)nm);
      for (int64_t wd = 0; wd < f->dense_words; wd++)
        for (int b = 0; b < 64; b++)
          if ((f->dense_bits[wd] >> b) & 1) put_u64(&w, (uint64_t)f->dense_lo + (uint64_t)(wd * 64 + b));
      break;
    }
    case KGO_FORM_BLOOM:
      put_u8(&w, 3);
      put_u32(&w, (uint32_t)(f->nblocks * 32));
      put(&w, f->bloom, (int64_t)f->nblocks * 32);
      break;
  }
  return w.n;
}

typedef struct {
  const uint8_t *p;
  int64_t n;
  int bad;
} rbuf;

static const uint8_t *take(rbuf *r, int64_t len) {
  if (r->bad || r->n < len) {
    r->bad = 1;
    return NULL;
  }
  const uint8_t *at = r->p;
  r->p += len;
  r->n -= len;
  return at;
}
static uint8_t get_u8(rbuf *r) {
  const uint8_t *q = take(r, 1);
  return q ? q[0] : 0;
}
static uint32_t get_u32(rbuf *r) {
  uint32_t v = 0;
  const uint8_t *q = take(r, 4);
  if (q) memcpy(&v, q, 4);
  return v;
}
static uint64_t get_u64(rbuf *r) {
  uint64_t v = 0;
  const uint8_t *q = take(r, 8);
  if (q) memcpy(&v, q, 8);
  return v;
}
static i128 get_i128(rbuf *r) {
  i128 v = 0;
  const uint8_t *q = take(r, 16);
  if (q) memcpy(&v, q, 16);
  return v;
}
/* Reader::count (codec.rs:322-330): at most the bytes left */
static int64_t get_count(rbuf *r) {
  int64_t n = get_u32(r);
  if (n > r->n) r->bad = 1;
  return r->bad ? 0 : n;
}

static void set_err(char *err, const char *msg) {
  if (err) {
    strncpy(err, msg, 127);
    err[127] = 0;
  }
}

/* KeyFilter::decode (codec.rs:186-246) with Reader (codec.rs:286-461). */
kgo_kf *kgo_kf_decode(const uint8_t *bytes, int64_t len, char *err) {
  rbuf r = {bytes, len, 0};
  uint8_t v0 = get_u8(&r), v1 = get_u8(&r), v2 = get_u8(&r);
  if (r.bad || v0 != KGO_FILTER_FORMAT_VERSION || v1 != KGO_FILTER_ROW_HASH_ID ||
      v2 != KGO_FILTER_NORMALIZATION_VERSION) {
    set_err(err, "an unsupported version");
    return NULL;
  }
  int tag = get_u8(&r), type = 0, p = 0, s = 0;
  switch (tag) {
    case 1: type = KGO_T_BOOL; break;
    case 2: type = KGO_T_INT8; break;
    case 3: type = KGO_T_INT16; break;
    case 4: type = KGO_T_INT32; break;
    case 5: type = KGO_T_INT64; break;
    case 6: type = KGO_T_FLOAT32; break;
    case 7: type = KGO_T_FLOAT64; break;
    case 8: type = KGO_T_DECIMAL128; break;
    case 9: type = KGO_T_DATE32; break;
    case 10: type = KGO_T_TIMESTAMP; break;
    case 17: type = KGO_T_DECIMAL64; break;
    default: set_err(err, "an unknown key type"); return NULL;
  }
  kgo_kf *f = kf_alloc(type, 0, 0, NULL);
  f->finished = 1;
  if (is_dec(type)) {
    p = get_u8(&r);
    s = (int8_t)get_u8(&r);
    f->precision = p;
    f->scale = s;
  } else if (type == KGO_T_TIMESTAMP) {
    int z = get_u8(&r);
    if (z == 1) {
      int64_t zl = get_u32(&r);
      const uint8_t *zb = take(&r, zl);
      if (!r.bad && zl > 63) {
        set_err(err, "a time zone longer than 63 bytes");
        kgo_kf_free(f);
        return NULL;
      }
      if (zb) {
        memcpy(f->tz, zb, (size_t)zl);
        f->tz_len = (int)zl;
      }
    } else if (z != 0) {
      r.bad = 1;
    }
  }
  uint8_t flags = get_u8(&r);
  if (!r.bad && flags > 7) {
    set_err(err, "unknown flags");
    kgo_kf_free(f);
    return NULL;
  }
  if (flags & 4) {
    set_err(err, "a top-N bound (not a join filter)");
    kgo_kf_free(f);
    return NULL;
  }
  f->nulls = flags & 1;
  f->empty = (flags >> 1) & 1;
  f->rows = get_u64(&r);
  int rt = get_u8(&r);
  if (rt == 0) {
    f->has_range = 0;
  } else if (rt == 1 && f->cls == CLS_INT) {
    f->lo = (int64_t)get_u64(&r);
    f->hi = (int64_t)get_u64(&r);
    f->has_range = 1;
  } else if (rt == 2 && f->cls == CLS_DEC) {
    f->lo = get_i128(&r);
    f->hi = get_i128(&r);
    f->has_range = 1;
  } else if (rt == 3 && f->cls == CLS_FLOAT) {
    uint64_t a = get_u64(&r), b = get_u64(&r);
    memcpy(&f->flo, &a, 8);
    memcpy(&f->fhi, &b, 8);
    f->has_range = 1;
  } else if (!r.bad) {
    set_err(err, "a range of another type");
    kgo_kf_free(f);
    return NULL;
  }
  int vt = get_u8(&r);
  if (vt == 1) {
    int64_t n = get_count(&r);
    f->values_ok = 1;
    f->nvalues = n;
    f->values = malloc((size_t)(n > 0 ? n : 1) * sizeof(i128));
    for (int64_t k = 0; k < n && !r.bad; k++) {
      if (get_u8(&r) != 0) {
        set_err(err, "a value of another type");
        kgo_kf_free(f);
        return NULL;
      }
      f->values[k] = get_i128(&r);
    }
  } else if (vt != 0 && !r.bad) {
    set_err(err, "an unknown value list");
    kgo_kf_free(f);
    return NULL;
  }
  int mt = get_u8(&r);
  if (mt == 0) {
    f->member = KGO_FORM_RANGE;
  } else if (mt == 1 || mt == 2) {
    int64_t n = get_count(&r);
    if (mt == 1 && n > KGO_SMALL_SET_MAX) {
      set_err(err, "a small set beyond its cap");
      kgo_kf_free(f);
      return NULL;
    }
    f->members = malloc((size_t)(n > 0 ? n : 1) * sizeof(i128));
    int64_t m = 0;
    for (int64_t k = 0; k < n && !r.bad; k++) {
      i128 v = f->cls == CLS_DEC ? get_i128(&r) : (i128)(int64_t)get_u64(&r);
      if (r.bad) break;
      if (mt == 2) {
        /* re-placed by hash: a repeated member is held once */
        int dup = 0;
        for (int64_t q = 0; q < m; q++) dup |= f->members[q] == v;
        if (dup) continue;
      }
      f->members[m++] = v;
    }
    f->nmembers = m;
    if (mt == 1) {
      f->member = KGO_FORM_SMALL;
    } else {
      qsort(f->members, (size_t)m, sizeof(i128), cmp_i128);
      f->member = KGO_FORM_EXACT;
      if (f->cls == CLS_INT && dense_of(f, f->members, m)) f->member = KGO_FORM_DENSE;
    }
  } else if (mt == 3) {
    int64_t bl = get_u32(&r);
    const uint8_t *bb = take(&r, bl);
    if (!r.bad && (bl == 0 || bl % 32 != 0)) {
      set_err(err, "a malformed Bloom");
      kgo_kf_free(f);
      return NULL;
    }
    if (bb) {
      f->nblocks = (size_t)bl / 32;
      f->bloom = malloc((size_t)bl);
      memcpy(f->bloom, bb, (size_t)bl);
      f->member = KGO_FORM_BLOOM;
    }
  } else if (!r.bad) {
    set_err(err, "an unknown row form");
    kgo_kf_free(f);
    return NULL;
  }
  if (r.bad) {
    set_err(err, "truncated");
    kgo_kf_free(f);
    return NULL;
  }
  if (r.n != 0) {
    set_err(err, "trailing bytes");
    kgo_kf_free(f);
    return NULL;
  }
  return f;
}

/* ------------------------------------------------------------ merge */

/* RuntimeBloom::pass_rate (bloom.rs:165-179) */
double kgo_sbbf_pass_rate(const uint32_t *words, size_t nblocks) {
  double sum = 0;
  for (size_t b = 0; b < nblocks; b++) {
    double prod = 1.0;
    for (int k = 0; k < 8; k++) prod *= (double)__builtin_popcount(words[b * 8 + k]) / 32.0;
    sum += prod;
  }
  return sum / (double)nblocks;
}

static kgo_kf *kf_clone(const kgo_kf *a) {
  kgo_kf *f = malloc(sizeof *f);
  *f = *a;
  f->dist = NULL;
  f->ndist = f->capdist = 0;
  f->hashes = NULL;
  f->nh = f->caph = 0;
  f->values = NULL;
  if (a->values) {
    f->values = malloc((size_t)(a->nvalues > 0 ? a->nvalues : 1) * sizeof(i128));
    memcpy(f->values, a->values, (size_t)a->nvalues * sizeof(i128));
  }
  f->members = NULL;
  if (a->members) {
    f->members = malloc((size_t)(a->nmembers > 0 ? a->nmembers : 1) * sizeof(i128));
    memcpy(f->members, a->members, (size_t)a->nmembers * sizeof(i128));
  }
  f->dense_bits = NULL;
  if (a->dense_bits) {
    f->dense_bits = malloc((size_t)a->dense_words * 8);
    memcpy(f->dense_bits, a->dense_bits, (size_t)a->dense_words * 8);
  }
  f->bloom = NULL;
  if (a->bloom) {
    f->bloom = malloc(a->nblocks * 32);
    memcpy(f->bloom, a->bloom, a->nblocks * 32);
  }
  return f;
}

/* members of a set form, sorted (a dense set written out) */
static int64_t set_members(const kgo_kf *f, i128 **out) {
  int64_t n = kgo_kf_members(f, NULL);
  uint64_t *w = malloc((size_t)(n > 0 ? n : 1) * 16);
  kgo_kf_members(f, w);
  i128 *m = malloc((size_t)(n > 0 ? n : 1) * sizeof(i128));
  for (int64_t k = 0; k < n; k++) m[k] = (i128)(((unsigned __int128)w[2 * k + 1] << 64) | w[2 * k]);
  free(w);
  *out = m;
  return n;
}

static int is_set(int form) {
  return form == KGO_FORM_SMALL || form == KGO_FORM_EXACT || form == KGO_FORM_DENSE;
}

/* KeyFilter::merge (merge.rs:35-117) with merge_members (119-169). */
kgo_kf *kgo_kf_merge(const kgo_kf *a, const kgo_kf *b, const kgo_filter_opts *o, char *err) {
  if (a->type != b->type || a->precision != b->precision || a->scale != b->scale ||
      a->tz_len != b->tz_len) {
    set_err(err, "runtime filters of different key types merged");
    return NULL;
  }
  int nulls = a->nulls || b->nulls;
  uint64_t keys = a->rows + b->rows;
  if (keys < a->rows) keys = UINT64_MAX;
  if (b->empty || a->empty) {
    kgo_kf *f = kf_clone(b->empty ? a : b);
    f->nulls = nulls;
    f->rows = keys;
    return f;
  }
  kgo_kf *f = kf_clone(a);
  f->opt = *o;
  f->nulls = nulls;
  f->rows = keys;
  f->empty = 0;
  /* range: union, none when either has none */
  if (a->has_range && b->has_range) {
    f->lo = a->lo < b->lo ? a->lo : b->lo;
    f->hi = a->hi > b->hi ? a->hi : b->hi;
    f->flo = fmin(a->flo, b->flo);
    f->fhi = fmax(a->fhi, b->fhi);
  } else {
    f->has_range = 0;
  }
  /* the statistics IN list */
  free(f->values);
  f->values = NULL;
  f->values_ok = 0;
  if (a->values_ok && b->values_ok) {
    int64_t n = a->nvalues + b->nvalues;
    i128 *v = malloc((size_t)(n > 0 ? n : 1) * sizeof(i128));
    memcpy(v, a->values, (size_t)a->nvalues * sizeof(i128));
    memcpy(v + a->nvalues, b->values, (size_t)b->nvalues * sizeof(i128));
    qsort(v, (size_t)n, sizeof(i128), cmp_i128);
    int64_t m = 0;
    for (int64_t k = 0; k < n; k++)
      if (m == 0 || v[m - 1] != v[k]) v[m++] = v[k];
    if ((uint64_t)m <= o->max_values) {
      f->values = v;
      f->nvalues = m;
      f->values_ok = 1;
    } else {
      free(v);
    }
  }
  /* the row form */
  size_t planned = 0;
  if (o->bloom_bytes) {
    size_t pb = o->bloom_bytes < o->bloom_max_bytes ? o->bloom_bytes : o->bloom_max_bytes;
    planned = kgo_sbbf_blocks_for_bytes(pb) * 32;
  }
  int ua = a->member == KGO_FORM_BLOOM && a->nblocks * 32 != planned;
  int ub = b->member == KGO_FORM_BLOOM && b->nblocks * 32 != planned;
  free(f->members);
  f->members = NULL;
  f->nmembers = 0;
  free(f->dense_bits);
  f->dense_bits = NULL;
  f->dense_words = 0;
  if (a->member != KGO_FORM_BLOOM) {
    free(f->bloom);
    f->bloom = NULL;
    f->nblocks = 0;
  }
  if (ua || ub || a->member == KGO_FORM_RANGE || b->member == KGO_FORM_RANGE) {
    free(f->bloom);
    f->bloom = NULL;
    f->nblocks = 0;
    f->member = KGO_FORM_RANGE;
  } else if (a->member == KGO_FORM_BLOOM && b->member == KGO_FORM_BLOOM) {
    if (a->nblocks == b->nblocks) {
      for (size_t k = 0; k < a->nblocks * 8; k++) f->bloom[k] |= b->bloom[k];
      f->member = KGO_FORM_BLOOM;
    } else {
      free(f->bloom);
      f->bloom = NULL;
      f->nblocks = 0;
      f->member = KGO_FORM_RANGE;
    }
  } else if (a->member == KGO_FORM_BLOOM || b->member == KGO_FORM_BLOOM) {
    const kgo_kf *bl = a->member == KGO_FORM_BLOOM ? a : b;
    const kgo_kf *st = a->member == KGO_FORM_BLOOM ? b : a;
    if (!f->bloom) {
      f->bloom = malloc(bl->nblocks * 32);
      memcpy(f->bloom, bl->bloom, bl->nblocks * 32);
      f->nblocks = bl->nblocks;
    }
    i128 *m;
    int64_t n = set_members(st, &m);
    for (int64_t k = 0; k < n; k++) kgo_sbbf_insert(f->bloom, f->nblocks, member_hash(st, m[k]));
    free(m);
    f->member = KGO_FORM_BLOOM;
  } else if (is_set(a->member) && is_set(b->member)) {
    i128 *ma, *mb;
    int64_t na = set_members(a, &ma), nb = set_members(b, &mb);
    int64_t n = na + nb;
    i128 *all = malloc((size_t)(n > 0 ? n : 1) * sizeof(i128));
    memcpy(all, ma, (size_t)na * sizeof(i128));
    memcpy(all + na, mb, (size_t)nb * sizeof(i128));
    free(ma);
    free(mb);
    qsort(all, (size_t)n, sizeof(i128), cmp_i128);
    int64_t m = 0;
    for (int64_t k = 0; k < n; k++)
      if (m == 0 || all[m - 1] != all[k]) all[m++] = all[k];
    if ((uint64_t)m <= o->exact_max_values) {
      f->members = all;
      f->nmembers = m;
      set_exact(f);
      if (f->member == KGO_FORM_DENSE) {
        free(f->members);
        f->members = NULL;
        f->nmembers = 0;
      }
    } else if (planned) {
      f->nblocks = planned / 32;
      f->bloom = calloc(f->nblocks, 32);
      for (int64_t k = 0; k < m; k++) kgo_sbbf_insert(f->bloom, f->nblocks, member_hash(f, all[k]));
      free(all);
      f->member = KGO_FORM_BLOOM;
    } else {
      free(all);
      f->member = KGO_FORM_RANGE;
    }
  } else {
    f->member = KGO_FORM_RANGE;
  }
  if (f->member == KGO_FORM_BLOOM && kgo_sbbf_pass_rate(f->bloom, f->nblocks) > o->bloom_max_fpp) {
    free(f->bloom);
    f->bloom = NULL;
    f->nblocks = 0;
    f->member = KGO_FORM_RANGE;
  }
  return f;
}

/* ------------------------------------------------------------ row stage */

#define RS_MAX 8
#define DROP_SAVES_NS 20.0 /* operators/runtime_filter.rs:252 */

struct kgo_rs {
  uint64_t min_rows;
  double max_pass_rate;
  int n;
  const kgo_kf *f[RS_MAX];
  int column[RS_MAX];
  uint64_t generation[RS_MAX];
  int64_t off_at[RS_MAX]; /* -1: on */
  uint64_t window_rows[RS_MAX], window_passed[RS_MAX];
  uint64_t tested[RS_MAX], dropped[RS_MAX];
};

kgo_rs *kgo_rs_new(uint64_t min_rows, double max_pass_rate) {
  kgo_rs *s = calloc(1, sizeof *s);
  s->min_rows = min_rows;
  s->max_pass_rate = max_pass_rate;
  return s;
}

void kgo_rs_free(kgo_rs *s) { free(s); }

/* A filter slot: attached pending (f NULL) or published; publishing a new
 * filter is a new generation. */
int kgo_rs_set(kgo_rs *s, int slot, const kgo_kf *f, int column) {
  if (slot < 0 || slot >= RS_MAX) return -1;
  if (slot >= s->n) {
    for (int k = s->n; k <= slot; k++) {
      s->f[k] = NULL;
      s->off_at[k] = -1;
      s->generation[k] = 0;
    }
    s->n = slot + 1;
  }
  s->f[slot] = f;
  s->column[slot] = column;
  if (f) s->generation[slot]++;
  return 0;
}

/* RowStage::start_partition (runtime_filter.rs:307-315) */
void kgo_rs_start_partition(kgo_rs *s) {
  for (int k = 0; k < s->n; k++) {
    s->window_rows[k] = s->window_passed[k] = 0;
    s->off_at[k] = -1;
  }
}

void kgo_rs_stats(const kgo_rs *s, int slot, uint64_t *tested, uint64_t *dropped, int *off) {
  *tested = s->tested[slot];
  *dropped = s->dropped[slot];
  *off = s->off_at[slot] >= 0 && (uint64_t)s->off_at[slot] == s->generation[slot];
}

/* RowStage::mask (runtime_filter.rs:329-395): out[i] = 1 for the rows some
 * producer may keep.  Returns the rows kept, or -1 when every row passes
 * (nothing filtered). */
int64_t kgo_rs_apply(kgo_rs *s, int ncols, const kgo_col *cols, int64_t n, uint8_t *out) {
  if (n == 0) return -1;
  int order[RS_MAX], m = 0;
  for (int k = 0; k < s->n; k++) {
    if (s->off_at[k] >= 0 && (uint64_t)s->off_at[k] == s->generation[k]) continue;
    s->off_at[k] = -1;
    if (s->f[k]) order[m++] = k;
  }
  /* cheapest first, stable */
  for (int a = 1; a < m; a++) {
    int x = order[a], b = a;
    while (b > 0 && kgo_kf_row_cost(s->f[order[b - 1]]) > kgo_kf_row_cost(s->f[x])) {
      order[b] = order[b - 1];
      b--;
    }
    order[b] = x;
  }
  int have = 0;
  int64_t alive = n;
  uint8_t *pass = malloc((size_t)n);
  for (int q = 0; q < m; q++) {
    int k = order[q];
    const kgo_kf *f = s->f[k];
    double cut = s->max_pass_rate;
    double c2 = 1.0 - kgo_kf_row_cost(f) / DROP_SAVES_NS;
    if (c2 < cut) cut = c2;
    if (s->column[k] < 0 || s->column[k] >= ncols) {
      free(pass);
      return -2;
    }
    if (kgo_kf_evaluate(f, &cols[s->column[k]], n, pass) < 0) continue;
    int narrowed = have && alive * 2 < n;
    uint64_t tested = 0, passed = 0;
    for (int64_t i = 0; i < n; i++) {
      if (narrowed && !out[i]) continue;
      tested++;
      passed += pass[i];
    }
    s->tested[k] += tested;
    s->dropped[k] += tested - passed;
    s->window_rows[k] += tested;
    s->window_passed[k] += passed;
    if (s->window_rows[k] >= s->min_rows) {
      double rate = (double)s->window_passed[k] / (double)s->window_rows[k];
      if (rate > cut) s->off_at[k] = (int64_t)s->generation[k];
      s->window_rows[k] = s->window_passed[k] = 0;
    }
    alive = 0;
    for (int64_t i = 0; i < n; i++) {
      out[i] = have ? (uint8_t)(out[i] & pass[i]) : pass[i];
      alive += out[i];
    }
    have = 1;
  }
  free(pass);
  if (!have || alive == n) return -1;
  return alive;
}
