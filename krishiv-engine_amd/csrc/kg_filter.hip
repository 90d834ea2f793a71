// kg_filter.hip -- the runtime KeyFilter on the device: the producer's
// builder (KeyFilterBuilder, crates/krishiv-io/src/runtime_filter/
// filter.rs:180-615), the row forms it publishes (member.rs:1-20), their
// evaluation on rows (member.rs:439-560), the versioned encoding (codec.rs),
// the merge of published filters (merge.rs:35-169), and a scan driver's row
// stage with its adaptive switch-off (crates/krishiv-exec/src/operators/
// runtime_filter.rs:240-395).
//
// Build side: one streaming pass a batch folds the included rows into
// per-workgroup summaries (min/max, NaN, NULL, row count), inserts the keys
// into a device set while they are few (64-bit compare-and-swap claims;
// decimals claim their high word, then their low word), and appends the
// rows' XXH64 hashes for the Bloom form.  Finish reads the few summaries and
// the set back and picks the form; the Bloom bits are set on the device
// from the hashes.  Probe side: one pass a batch evaluates every active
// filter of the stage on each row (range compare, a member test: up to 8
// compares, one bit of a dense set, a probe of an exact set, or one 32-byte
// Bloom block) and writes a 64-row ballot word per wave; the survivors'
// row ids follow from a scan of the per-tile counts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "kg_common.hpp"
#include "kg_device.hpp"
#include "kg_scan.hpp"

namespace kg {

typedef __int128 i128;
typedef unsigned __int128 u128;

enum { KF_INT = 0, KF_FLOAT = 1, KF_DEC = 2 };
constexpr int kKfBlock = 256;
constexpr int kKfTile = kKfBlock * 32;  // rows a workgroup tile
constexpr int kKfParts = 1024;          // workgroups of a build pass (their summaries)
constexpr int kSmallMax = 8;            // member.rs:34 SMALL_SET_MAX
constexpr uint64_t kDenseMaxBits = 1ull << 20;  // member.rs:47
constexpr int64_t kDenseSizeFactor = 16;        // member.rs:53
constexpr int64_t kDenseFreeBytes = 4096;       // member.rs:56
constexpr int64_t kHashSetBytesPerKey = 9;      // member.rs:60
constexpr uint64_t kExactMaxCap = 1 << 16;      // largest exact_max_values accepted
constexpr double kDropSavesNs = 20.0;           // operators/runtime_filter.rs:252
// the builder's distinct set: an empty int/float slot holds INT64_MIN (a
// float key never has that pattern: -0.0 is normalized to 0.0); an empty
// decimal slot holds a high word no decimal of 38 digits has
constexpr uint64_t kSetEmpty = 0x8000000000000000ULL;
constexpr uint64_t kHiEmpty = 0x7FFFFFFFFFFFFFFFULL;
enum { SET_OVER = 1, SET_SENTINEL = 2, SET_WIDE = 4 };

static int key_class(int t, int *width, int *single) {
  *single = 0;
  switch (t) {
    case KG_T_BOOL: case KG_T_INT8: *width = 1; return KF_INT;
    case KG_T_INT16: *width = 2; return KF_INT;
    case KG_T_INT32: case KG_T_DATE32: *width = 4; return KF_INT;
    case KG_T_INT64: case KG_T_TIMESTAMP: *width = 8; return KF_INT;
    case KG_T_FLOAT32: *width = 4; *single = 1; return KF_FLOAT;
    case KG_T_FLOAT64: *width = 8; return KF_FLOAT;
    case KG_T_DECIMAL64: case KG_T_DECIMAL128: *width = 16; return KF_DEC;
    default: return -1;
  }
}

static bool is_dec_type(int t) { return t == KG_T_DECIMAL64 || t == KG_T_DECIMAL128; }
static int int_rank(int t) {
  switch (t) {
    case KG_T_INT8: return 1;
    case KG_T_INT16: return 2;
    case KG_T_INT32: return 3;
    case KG_T_INT64: return 4;
    default: return 0;
  }
}
static int int_digits(int t) {
  static const int d[] = {0, 3, 5, 10, 19};
  return d[int_rank(t)];
}

// widens_to (key.rs:75-103) for the fixed-width types.
static bool widens(int from, int fp, int fs, int to, int tp, int ts) {
  int w, s;
  if (key_class(to, &w, &s) < 0) return false;
  if (from == to && (!is_dec_type(to) || (fp == tp && fs == ts))) return true;
  if (to == KG_T_FLOAT64 &&
      (from == KG_T_FLOAT32 || from == KG_T_INT8 || from == KG_T_INT16 || from == KG_T_INT32))
    return true;
  if (to == KG_T_FLOAT32 && (from == KG_T_INT8 || from == KG_T_INT16)) return true;
  if (is_dec_type(from) && is_dec_type(to)) return ts >= fs && tp - ts >= fp - fs;
  if (is_dec_type(to)) return int_digits(from) > 0 && tp - ts >= int_digits(from);
  return int_rank(from) > 0 && int_rank(to) > 0 && int_rank(from) < int_rank(to);
}

// ---------------------------------------------------------------- device forms
struct KfDev {  // a published filter as the kernels test it
  int32_t cls, key_width, single, form;
  int32_t has_range, empty, nulls_pass, nsmall;
  int64_t ilo, ihi;
  uint64_t dlo[2], dhi[2];
  double flo, fhi;
  uint64_t small[kSmallMax][2];
  int64_t dense_lo;
  uint64_t dense_words;
  const uint64_t *dense_bits;
  uint64_t tab_mask;
  const uint64_t *tab;     // exact set: (low, high) words a slot
  const uint8_t *tab_used;
  const uint32_t *bloom;
  uint64_t nblocks;
};

struct KfProbe {  // a probe column normalized to the key type (key.rs:188-282)
  DCol col;
  int32_t rescale;     // decimals: multiply by fac (|v| <= lim first)
  uint64_t fac[2], lim[2];
};

__device__ __forceinline__ i128 mk128(uint64_t lo, uint64_t hi) {
  return (i128)(((u128)hi << 64) | lo);
}

__device__ __forceinline__ uint64_t int_hash(int64_t v, int width) {
  const uint64_t u = (uint64_t)v;
  if (width == 1) return dev::xxh64_fixed(u & 0xFF, 0, 1);
  if (width == 2) return dev::xxh64_fixed(u & 0xFFFF, 0, 2);
  if (width == 4) return dev::xxh64_fixed(u & 0xFFFFFFFFULL, 0, 4);
  return dev::xxh64_fixed(u, 0, 8);
}

// spark_float then the key's bit pattern (key.rs:146-154, 363-371)
__device__ __forceinline__ double spark_f(double v) {
  if (v != v) return __longlong_as_double(0x7FF8000000000000LL);
  return v + 0.0;
}
__device__ __forceinline__ uint64_t float_key_bits(double v, int single) {
  if (v != v) return single ? 0x7FC00000ULL : 0x7FF8000000000000ULL;
  return single ? (uint64_t)__float_as_uint((float)v) : (uint64_t)__double_as_longlong(v);
}

// The row form's member test of a value (low, high words) with row hash h.
__device__ __forceinline__ bool member_test(const KfDev &F, uint64_t lo, uint64_t hi, uint64_t h) {
  switch (F.form) {
    case KG_FORM_SMALL: {
      bool hit = false;
      for (int k = 0; k < F.nsmall; k++) hit |= F.small[k][0] == lo && F.small[k][1] == hi;
      return hit;
    }
    case KG_FORM_DENSE: {
      const uint64_t d = lo - (uint64_t)F.dense_lo;  // below dense_lo wraps past the words
      return (d >> 6) < F.dense_words && ((F.dense_bits[d >> 6] >> (d & 63)) & 1);
    }
    case KG_FORM_EXACT: {
      uint64_t s = h & F.tab_mask;
      for (uint64_t p = 0; p <= F.tab_mask; p++) {
        if (!F.tab_used[s]) return false;
        if (F.tab[2 * s] == lo && F.tab[2 * s + 1] == hi) return true;
        s = (s + 1) & F.tab_mask;
      }
      return false;
    }
    case KG_FORM_BLOOM: return dev::sbbf_test(F.bloom, F.nblocks, h);
    default: return true;
  }
}

// KeyFilter::evaluate_keys of one row (filter.rs:780-797).
__device__ __forceinline__ bool kf_row(const KfDev &F, const KfProbe &P, int64_t i,
                                       unsigned int *err) {
  const DCol &c = P.col;
  if (!dev::valid(c, i)) return F.nulls_pass != 0;
  if (F.empty) return false;
  if (F.cls == KF_FLOAT) {
    const double v = spark_f(dev::load_f64(c, i));
    if (F.has_range && !(v >= F.flo && v <= F.fhi)) return false;
    if (F.form == KG_FORM_RANGE) return true;
    const uint64_t bits = float_key_bits(v, F.single);
    const uint64_t h = F.single ? dev::xxh64_fixed(bits, 0, 4) : dev::xxh64_fixed(bits, 0, 8);
    return member_test(F, bits, 0, h);
  }
  if (F.cls == KF_DEC) {
    uint64_t lo;
    int64_t hi;
    dev::load_i128(c, i, lo, hi);
    i128 v = mk128(lo, (uint64_t)hi);
    if (P.rescale) {
      const u128 lim = (u128)mk128(P.lim[0], P.lim[1]);
      const u128 mag = v < 0 ? (u128)(-v) : (u128)v;
      if (mag > lim) {
        atomicOr(err, 1u);
        return true;
      }
      v *= mk128(P.fac[0], P.fac[1]);
    }
    if (F.has_range && (v < mk128(F.dlo[0], F.dlo[1]) || v > mk128(F.dhi[0], F.dhi[1])))
      return false;
    if (F.form == KG_FORM_RANGE) return true;
    const uint64_t vl = (uint64_t)v, vh = (uint64_t)((u128)v >> 64);
    return member_test(F, vl, vh, dev::xxh64_fixed(vl, vh, 16));
  }
  const int64_t v = dev::load_i64(c, i);
  if (F.has_range && (v < F.ilo || v > F.ihi)) return false;
  if (F.form == KG_FORM_RANGE) return true;
  const uint64_t h = (F.form == KG_FORM_EXACT || F.form == KG_FORM_BLOOM) ? int_hash(v, F.key_width) : 0;
  return member_test(F, (uint64_t)v, (uint64_t)(v >> 63), h);
}

// One filter over rows [0, n): a ballot word per 64 rows, a count per tile.
__global__ void __launch_bounds__(kKfBlock) k_kf_eval(const KfDev *F, const KfProbe *P, int64_t n,
                                                      uint64_t *bitmap, uint32_t *tile_count,
                                                      unsigned int *err) {
  __shared__ uint32_t s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kKfTile;
  const int lane = threadIdx.x & 63;
  uint32_t cnt = 0;
  for (int k = 0; k < kKfTile / kKfBlock; k++) {
    const int64_t i = base + (int64_t)k * kKfBlock + threadIdx.x;
    const bool pass = i < n && kf_row(*F, *P, i, err);
    const uint64_t m = __ballot(pass);
    const int64_t w0 = base + (int64_t)k * kKfBlock + (threadIdx.x & ~63);
    if (lane == 0 && w0 < n) {
      bitmap[w0 >> 6] = m;
      cnt += (uint32_t)__popcll(m);
    }
  }
  if (lane == 0 && cnt) atomicAdd(&s_cnt, cnt);
  __syncthreads();
  if (threadIdx.x == 0) tile_count[blockIdx.x] = s_cnt;
}

// ---------------------------------------------------------------- row stage
struct RsDev {
  int32_t k;
  KfDev f[KG_ROW_STAGE_SLOTS];
  KfProbe p[KG_ROW_STAGE_SLOTS];
};

// Every active filter on every row, in the stage's order: the row passes
// when all pass.  counts[2k] = rows filter k passes; counts[2k + 1] = rows
// passing filters 0..k (what the stage needs to replay the narrowing of
// runtime_filter.rs:357-371 exactly).
__global__ void __launch_bounds__(kKfBlock) k_rs_eval(const RsDev *R, int64_t n, uint64_t *bitmap,
                                                      uint32_t *tile_count,
                                                      unsigned long long *counts,
                                                      unsigned int *err) {
  __shared__ uint32_t s_cnt[2 * KG_ROW_STAGE_SLOTS + 1];
  if (threadIdx.x < 2 * KG_ROW_STAGE_SLOTS + 1) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int K = R->k;
  const int64_t base = (int64_t)blockIdx.x * kKfTile;
  const int lane = threadIdx.x & 63;
  for (int t = 0; t < kKfTile / kKfBlock; t++) {
    const int64_t i = base + (int64_t)t * kKfBlock + threadIdx.x;
    const bool in = i < n;
    bool surv = in;
    for (int k = 0; k < K; k++) {
      const bool p = in && kf_row(R->f[k], R->p[k], i, err);
      const uint64_t ma = __ballot(p), mb = __ballot(surv && p);
      surv = surv && p;
      if (lane == 0) {
        if (ma) atomicAdd(&s_cnt[2 * k], (uint32_t)__popcll(ma));
        if (mb) atomicAdd(&s_cnt[2 * k + 1], (uint32_t)__popcll(mb));
      }
    }
    const uint64_t m = __ballot(surv);
    const int64_t w0 = base + (int64_t)t * kKfBlock + (threadIdx.x & ~63);
    if (lane == 0 && w0 < n) {
      bitmap[w0 >> 6] = m;
      if (m) atomicAdd(&s_cnt[2 * KG_ROW_STAGE_SLOTS], (uint32_t)__popcll(m));
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * K && s_cnt[threadIdx.x])
    atomicAdd(counts + threadIdx.x, (unsigned long long)s_cnt[threadIdx.x]);
  if (threadIdx.x == 0) tile_count[blockIdx.x] = s_cnt[2 * KG_ROW_STAGE_SLOTS];
}

// ---------------------------------------------------------------- builder
struct KfPart {  // one workgroup's summary of a batch
  uint64_t lo[2], hi[2];
  double flo, fhi;
  uint32_t nan, nulls;
  uint64_t rows;
};

struct KfAdd {
  DCol key, inc;
  int32_t has_inc, cls, width, single;
  uint64_t *k0, *k1, *z;  // distinct set (k0 null: not collected)
  uint64_t mask;
  uint32_t *ctl;  // [0] keys inserted, [1] SET_* flags
  uint32_t max;
  uint64_t *hashes;  // null: not collected
  unsigned long long *nh;
  KfPart *parts;
};

__device__ __forceinline__ void set_new(const KfAdd &A) {
  const uint32_t c = atomicAdd(A.ctl, 1u) + 1;
  if (c > A.max) atomicOr(A.ctl + 1, (uint32_t)SET_OVER);
}

// Distinct::insert (filter.rs:125-140) into the device set.
__device__ void set_insert(const KfAdd &A, uint64_t lo, uint64_t hi, uint64_t h) {
  if (__atomic_load_n(A.ctl + 1, __ATOMIC_RELAXED) & SET_OVER) return;
  if (A.cls != KF_DEC) {
    if (lo == kSetEmpty) {
      if (!(atomicOr(A.ctl + 1, (uint32_t)SET_SENTINEL) & SET_SENTINEL)) set_new(A);
      return;
    }
    uint64_t s = h & A.mask;
    for (uint64_t p = 0; p <= A.mask; p++) {
      const uint64_t old = atomicCAS((unsigned long long *)&A.k0[s], kSetEmpty, lo);
      if (old == kSetEmpty) {
        set_new(A);
        return;
      }
      if (old == lo) return;
      s = (s + 1) & A.mask;
    }
    atomicOr(A.ctl + 1, (uint32_t)SET_OVER);
    return;
  }
  if (hi == kHiEmpty) {  // no decimal of 38 digits has this high word
    atomicOr(A.ctl + 1, (uint32_t)SET_WIDE);
    return;
  }
  uint64_t s = h & A.mask;
  if (lo == 0) {  // the low word cannot mark its own slot: a set of high words
    for (uint64_t p = 0; p <= A.mask; p++) {
      const uint64_t old = atomicCAS((unsigned long long *)&A.z[s], kHiEmpty, hi);
      if (old == kHiEmpty) {
        set_new(A);
        return;
      }
      if (old == hi) return;
      s = (s + 1) & A.mask;
    }
    atomicOr(A.ctl + 1, (uint32_t)SET_OVER);
    return;
  }
  // a slot is the high word (claimed first) and then the low word: a lane
  // whose high word matches installs or compares the low word; every slot's
  // words are written once, so lanes of one value stop at the same slot
  for (uint64_t p = 0; p <= A.mask; p++) {
    const uint64_t oh = atomicCAS((unsigned long long *)&A.k0[s], kHiEmpty, hi);
    if (oh == kHiEmpty || oh == hi) {
      const uint64_t ol = atomicCAS((unsigned long long *)&A.k1[s], 0ULL, lo);
      if (ol == 0) {
        set_new(A);
        return;
      }
      if (ol == lo) return;
    }
    s = (s + 1) & A.mask;
  }
  atomicOr(A.ctl + 1, (uint32_t)SET_OVER);
}

// add_hashed over one batch (filter.rs:225-284): block-uniform steps, so the
// hash appends take one atomic a wave.
__global__ void __launch_bounds__(kKfBlock) k_kf_add(KfAdd A, int64_t n) {
  __shared__ KfPart s_part[kKfBlock];
  i128 lo = 0, hi = 0;
  double flo = INFINITY, fhi = -INFINITY;
  uint32_t nan = 0, nulls = 0;
  uint64_t rows = 0;
  const int lane = threadIdx.x & 63;
  for (int64_t b0 = blockIdx.x * (int64_t)kKfBlock; b0 < n; b0 += (int64_t)gridDim.x * kKfBlock) {
    const int64_t i = b0 + threadIdx.x;
    const bool in = i < n;
    const bool kept = in && (!A.has_inc || (dev::valid(A.inc, i) && dev::load_i64(A.inc, i) != 0));
    const bool live = kept && dev::valid(A.key, i);
    if (kept && !live) nulls = 1;
    uint64_t vlo = 0, vhi = 0, h = 0;
    if (live) {
      if (A.cls == KF_FLOAT) {
        const double v = spark_f(dev::load_f64(A.key, i));
        if (v != v) nan = 1;
        flo = fmin(flo, v);
        fhi = fmax(fhi, v);
        vlo = float_key_bits(v, A.single);
        h = A.single ? dev::xxh64_fixed(vlo, 0, 4) : dev::xxh64_fixed(vlo, 0, 8);
      } else if (A.cls == KF_DEC) {
        int64_t shi;
        dev::load_i128(A.key, i, vlo, shi);
        vhi = (uint64_t)shi;
        const i128 v = mk128(vlo, vhi);
        if (rows == 0 || v < lo) lo = v;
        if (rows == 0 || v > hi) hi = v;
        h = dev::xxh64_fixed(vlo, vhi, 16);
      } else {
        const int64_t v = dev::load_i64(A.key, i);
        if (rows == 0 || v < lo) lo = v;
        if (rows == 0 || v > hi) hi = v;
        vlo = (uint64_t)v;
        vhi = (uint64_t)(v >> 63);
        h = int_hash(v, A.width);
      }
      rows++;
      if (A.k0) set_insert(A, vlo, vhi, h);
    }
    if (A.hashes) {
      const uint64_t m = __ballot(live);
      unsigned long long at = 0;
      if (lane == 0 && m) at = atomicAdd(A.nh, (unsigned long long)__popcll(m));
      at = __shfl(at, 0);
      if (live) A.hashes[at + __popcll(m & ((1ULL << lane) - 1))] = h;
    }
  }
  KfPart &mine = s_part[threadIdx.x];
  mine.lo[0] = (uint64_t)lo;
  mine.lo[1] = (uint64_t)((u128)lo >> 64);
  mine.hi[0] = (uint64_t)hi;
  mine.hi[1] = (uint64_t)((u128)hi >> 64);
  mine.flo = flo;
  mine.fhi = fhi;
  mine.nan = nan;
  mine.nulls = nulls;
  mine.rows = rows;
  __syncthreads();
  if (threadIdx.x == 0) {
    KfPart t = s_part[0];
    for (int k = 1; k < kKfBlock; k++) {
      const KfPart &o = s_part[k];
      if (o.rows) {
        const i128 olo = mk128(o.lo[0], o.lo[1]), ohi = mk128(o.hi[0], o.hi[1]);
        if (!t.rows || olo < mk128(t.lo[0], t.lo[1])) {
          t.lo[0] = o.lo[0];
          t.lo[1] = o.lo[1];
        }
        if (!t.rows || ohi > mk128(t.hi[0], t.hi[1])) {
          t.hi[0] = o.hi[0];
          t.hi[1] = o.hi[1];
        }
      }
      t.flo = fmin(t.flo, o.flo);
      t.fhi = fmax(t.fhi, o.fhi);
      t.nan |= o.nan;
      t.nulls |= o.nulls;
      t.rows += o.rows;
    }
    A.parts[blockIdx.x] = t;
  }
}

// The batch's summary: parts[0..np) folded into parts[np].
__global__ void k_kf_fold(KfPart *parts, int np) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  KfPart t = parts[0];
  for (int k = 1; k < np; k++) {
    const KfPart &o = parts[k];
    if (o.rows) {
      if (!t.rows || mk128(o.lo[0], o.lo[1]) < mk128(t.lo[0], t.lo[1])) {
        t.lo[0] = o.lo[0];
        t.lo[1] = o.lo[1];
      }
      if (!t.rows || mk128(o.hi[0], o.hi[1]) > mk128(t.hi[0], t.hi[1])) {
        t.hi[0] = o.hi[0];
        t.hi[1] = o.hi[1];
      }
    }
    t.flo = fmin(t.flo, o.flo);
    t.fhi = fmax(t.fhi, o.fhi);
    t.nan |= o.nan;
    t.nulls |= o.nulls;
    t.rows += o.rows;
  }
  parts[np] = t;
}

// Members of another builder's set inserted into this one (a builder merge).
__global__ void k_kf_set_merge(KfAdd A, const uint64_t *vals, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = vals[3 * i], hi = vals[3 * i + 1], h = vals[3 * i + 2];
    set_insert(A, lo, hi, h);
  }
}

__global__ void k_kf_bloom_put(const uint64_t *h, int64_t n, uint32_t *words, uint64_t nblocks) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dev::sbbf_put(words, nblocks, h[i]);
}

__global__ void k_kf_bloom_or(uint32_t *a, const uint32_t *b, int64_t nwords) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nwords;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] |= b[i];
}

// RuntimeBloom::pass_rate (bloom.rs:165-179) times 32^8 * blocks: each
// block's product of its words' popcounts, summed exactly.
__global__ void k_kf_pass_rate(const uint32_t *words, uint64_t nblocks, unsigned long long *sum) {
  unsigned long long mine = 0;
  for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < nblocks;
       b += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long p = 1;
    for (int w = 0; w < 8; w++) p *= (unsigned long long)__popc(words[b * 8 + w]);
    mine += p;
  }
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(sum, mine);
}

__global__ void k_kf_fill(uint64_t *p, int64_t n, uint64_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------- host helpers
static uint64_t host_int_hash(int64_t v, int width) {
  const uint64_t u = (uint64_t)v;
  if (width == 1) return dev::xxh64_fixed(u & 0xFF, 0, 1);
  if (width == 2) return dev::xxh64_fixed(u & 0xFFFF, 0, 2);
  if (width == 4) return dev::xxh64_fixed(u & 0xFFFFFFFFULL, 0, 4);
  return dev::xxh64_fixed(u, 0, 8);
}

static uint64_t lo_of(i128 v) { return (uint64_t)v; }
static uint64_t hi_of(i128 v) { return (uint64_t)((u128)v >> 64); }
static i128 mk(uint64_t lo, uint64_t hi) { return (i128)(((u128)hi << 64) | lo); }

}  // namespace kg

using namespace kg;

struct kg_key_filter {
  kg_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  int type = 0, precision = 0, scale = 0, cls = 0, width = 0, single = 0;
  int tz_len = -1;
  char tz[64] = {};
  kg_filter_options opt{};
  bool finished = false;
  // builder summary (host) and device collections
  bool has_range = false, nan = false, nulls = false;
  i128 lo = 0, hi = 0;
  double flo = 0, fhi = 0;
  uint64_t rows = 0;
  bool distinct_ok = false;
  DevBuf set0, set1, setz, setctl;
  uint64_t set_mask = 0;
  bool hashes_ok = false;
  DevBuf hashes, nh;
  uint64_t nhashes = 0;
  DevBuf parts;
  // the published filter
  bool values_ok = false;
  std::vector<i128> values;
  int form = KG_FORM_RANGE;
  std::vector<i128> members;  // set forms, ascending
  int64_t dense_lo = 0;
  std::vector<uint64_t> dense_host;
  DevBuf dense_bits, tab, tab_used;
  uint64_t tab_mask = 0;
  DevBuf bloom;
  uint64_t nblocks = 0;
  bool empty = false;
  DevBuf kdev;  // the KfDev the kernels read
  kg::KfDev hdev{};  // its host copy (a row stage's launch copies it)
  // evaluation scratch
  DevBuf bitmap, tiles, tile_off, temp, probe, err;
};

namespace kg {

struct KfDel {
  void operator()(kg_key_filter *f) const { kg_key_filter_destroy(f); }
};
using KfPtr = std::unique_ptr<kg_key_filter, KfDel>;

static kg_key_filter *kf_new(kg_ctx *ctx, int type, int precision, int scale,
                             const kg_filter_options *o) {
  int w, sg;
  const int cls = key_class(type, &w, &sg);
  if (cls < 0) fail(KG_ERR_UNSUPPORTED, "no runtime filter for this key type");
  auto f = std::make_unique<kg_key_filter>();
  f->ctx = ctx;
  f->type = type;
  f->precision = precision;
  f->scale = scale;
  f->cls = cls;
  f->width = w;
  f->single = sg;
  if (o) f->opt = *o;
  else kg_filter_default_options(&f->opt);
  f->s = make_stream(ctx->device);
  return f.release();
}

static int64_t distinct_max(const kg_key_filter *f) {
  return (int64_t)std::max(f->opt.max_values, f->opt.exact_max_values);
}

static KfAdd add_args(kg_key_filter *f) {
  KfAdd A{};
  A.cls = f->cls;
  A.width = f->width;
  A.single = f->single;
  if (f->distinct_ok) {
    A.k0 = f->set0.get<uint64_t>();
    A.k1 = f->cls == KF_DEC ? f->set1.get<uint64_t>() : nullptr;
    A.z = f->cls == KF_DEC ? f->setz.get<uint64_t>() : nullptr;
    A.mask = f->set_mask;
    A.ctl = f->setctl.get<uint32_t>();
    A.max = (uint32_t)distinct_max(f);
  }
  if (f->hashes_ok) {
    A.hashes = f->hashes.get<uint64_t>();
    A.nh = f->nh.get<unsigned long long>();
  }
  A.parts = f->parts.get<KfPart>();
  return A;
}

static void drop_set(kg_key_filter *f) {
  f->distinct_ok = false;
  f->set0.release();
  f->set1.release();
  f->setz.release();
}

static void drop_hashes(kg_key_filter *f) {
  f->hashes_ok = false;
  f->hashes.release();
  f->nhashes = 0;
}

static uint32_t read_set_flags(kg_key_filter *f, uint32_t *count) {
  uint32_t ctl[2] = {0, 0};
  KG_HIP(hipMemcpyAsync(ctl, f->setctl.get(), 8, hipMemcpyDeviceToHost, f->s));
  KG_HIP(hipStreamSynchronize(f->s));
  if (count) *count = ctl[0];
  return ctl[1];
}

static void check_set(kg_key_filter *f) {
  const uint32_t fl = read_set_flags(f, nullptr);
  if (fl & SET_WIDE) fail(KG_ERR_UNSUPPORTED, "a decimal key beyond 38 digits");
  if (fl & SET_OVER) drop_set(f);
}

// The set's members, ascending (the device set read back).
static std::vector<i128> set_members(kg_key_filter *f) {
  std::vector<i128> out;
  const uint64_t cap = f->set_mask + 1;
  std::vector<uint64_t> k0(cap), k1, z;
  KG_HIP(hipMemcpyAsync(k0.data(), f->set0.get(), cap * 8, hipMemcpyDeviceToHost, f->s));
  if (f->cls == KF_DEC) {
    k1.resize(cap);
    z.resize(cap);
    KG_HIP(hipMemcpyAsync(k1.data(), f->set1.get(), cap * 8, hipMemcpyDeviceToHost, f->s));
    KG_HIP(hipMemcpyAsync(z.data(), f->setz.get(), cap * 8, hipMemcpyDeviceToHost, f->s));
  }
  const uint32_t fl = read_set_flags(f, nullptr);
  if (f->cls == KF_DEC) {
    for (uint64_t s = 0; s < cap; s++) {
      if (k0[s] != kHiEmpty && k1[s] != 0) out.push_back(mk(k1[s], k0[s]));
      if (z[s] != kHiEmpty) out.push_back(mk(0, z[s]));
    }
  } else {
    for (uint64_t s = 0; s < cap; s++)
      if (k0[s] != kSetEmpty)
        out.push_back(f->cls == KF_FLOAT ? (i128)k0[s] : (i128)(int64_t)k0[s]);
    if (fl & SET_SENTINEL) out.push_back(f->cls == KF_FLOAT ? (i128)kSetEmpty : (i128)INT64_MIN);
  }
  std::sort(out.begin(), out.end());
  return out;
}

// Class::hash64 / decimal_hash of a member (member.rs:89-105).
static uint64_t member_hash(const kg_key_filter *f, i128 m) {
  if (f->cls == KF_DEC) return dev::xxh64_fixed(lo_of(m), hi_of(m), 16);
  if (f->cls == KF_FLOAT)
    return f->single ? dev::xxh64_fixed(lo_of(m) & 0xFFFFFFFFULL, 0, 4)
                     : dev::xxh64_fixed(lo_of(m), 0, 8);
  return host_int_hash((int64_t)m, f->width);
}

// Dense::of (member.rs:150-176) over ascending distinct integers.
static bool dense_of(kg_key_filter *f, const std::vector<i128> &v) {
  if (v.empty()) return false;
  const i128 span = v.back() - v.front() + 1;
  if (span > (i128)kDenseMaxBits) return false;
  const int64_t words = (int64_t)((span + 63) / 64);
  const int64_t allow = std::max<int64_t>(kDenseFreeBytes,
                                          (int64_t)v.size() * kHashSetBytesPerKey * kDenseSizeFactor);
  if (words * 8 > allow) return false;
  f->dense_lo = (int64_t)v.front();
  f->dense_host.assign((size_t)words, 0);
  for (i128 x : v) {
    const uint64_t d = (uint64_t)(int64_t)x - (uint64_t)f->dense_lo;
    f->dense_host[d >> 6] |= 1ULL << (d & 63);
  }
  return true;
}

// Member::exact (member.rs:320-375) with Exact::densify (210-219) over
// f->members.
static void set_exact_form(kg_key_filter *f) {
  if (f->members.size() <= (size_t)kSmallMax) {
    f->form = KG_FORM_SMALL;
    return;
  }
  f->form = KG_FORM_EXACT;
  if (f->cls == KF_INT && dense_of(f, f->members)) f->form = KG_FORM_DENSE;
}

static void bloom_alloc(kg_key_filter *f, uint64_t nblocks) {
  f->nblocks = nblocks;
  f->bloom.reset(f->ctx, nblocks * 32);
  KG_HIP(hipMemsetAsync(f->bloom.get(), 0, nblocks * 32, f->s));
}

static void bloom_put_hashes(kg_key_filter *f, const uint64_t *dev_hashes, int64_t n) {
  if (n <= 0) return;
  k_kf_bloom_put<<<stream_grid(n, 256), 256, 0, f->s>>>(dev_hashes, n, f->bloom.get<uint32_t>(),
                                                        f->nblocks);
  KG_HIP(hipGetLastError());
}

static void bloom_put_members(kg_key_filter *f, const kg_key_filter *from,
                              const std::vector<i128> &m) {
  if (m.empty()) return;
  std::vector<uint64_t> h(m.size());
  for (size_t k = 0; k < m.size(); k++) h[k] = member_hash(from, m[k]);
  DevBuf d(f->ctx, h.size() * 8);
  KG_HIP(hipMemcpyAsync(d.get(), h.data(), h.size() * 8, hipMemcpyHostToDevice, f->s));
  bloom_put_hashes(f, d.get<uint64_t>(), (int64_t)h.size());
  KG_HIP(hipStreamSynchronize(f->s));
}

static double pass_rate(kg_key_filter *f) {
  DevBuf sum(f->ctx, 8);
  KG_HIP(hipMemsetAsync(sum.get(), 0, 8, f->s));
  k_kf_pass_rate<<<stream_grid((int64_t)f->nblocks, 256), 256, 0, f->s>>>(
      f->bloom.get<uint32_t>(), f->nblocks, sum.get<unsigned long long>());
  KG_HIP(hipGetLastError());
  unsigned long long s = 0;
  KG_HIP(hipMemcpyAsync(&s, sum.get(), 8, hipMemcpyDeviceToHost, f->s));
  KG_HIP(hipStreamSynchronize(f->s));
  return (double)s / 1099511627776.0 / (double)f->nblocks;  // 32^8 = 2^40
}

// The device structures of the published form and its KfDev.
static void publish(kg_key_filter *f) {
  KfDev D{};
  D.cls = f->cls;
  D.key_width = f->width;
  D.single = f->single;
  D.form = f->form;
  D.has_range = f->has_range;
  D.empty = f->empty;
  D.nulls_pass = f->nulls;
  D.ilo = (int64_t)f->lo;
  D.ihi = (int64_t)f->hi;
  D.dlo[0] = lo_of(f->lo);
  D.dlo[1] = hi_of(f->lo);
  D.dhi[0] = lo_of(f->hi);
  D.dhi[1] = hi_of(f->hi);
  D.flo = f->flo;
  D.fhi = f->fhi;
  if (f->form == KG_FORM_SMALL) {
    D.nsmall = (int)f->members.size();
    for (size_t k = 0; k < f->members.size(); k++) {
      D.small[k][0] = lo_of(f->members[k]);
      D.small[k][1] = hi_of(f->members[k]);
    }
  } else if (f->form == KG_FORM_DENSE) {
    f->dense_bits.reset(f->ctx, f->dense_host.size() * 8);
    KG_HIP(hipMemcpyAsync(f->dense_bits.get(), f->dense_host.data(), f->dense_host.size() * 8,
                          hipMemcpyHostToDevice, f->s));
    D.dense_lo = f->dense_lo;
    D.dense_words = f->dense_host.size();
    D.dense_bits = f->dense_bits.get<uint64_t>();
  } else if (f->form == KG_FORM_EXACT) {
    uint64_t cap = 16;
    while (cap < 2 * f->members.size()) cap <<= 1;
    std::vector<uint64_t> tab(2 * cap, 0);
    std::vector<uint8_t> used(cap, 0);
    for (i128 m : f->members) {
      uint64_t s = member_hash(f, m) & (cap - 1);
      while (used[s]) s = (s + 1) & (cap - 1);
      used[s] = 1;
      tab[2 * s] = lo_of(m);
      tab[2 * s + 1] = hi_of(m);
    }
    f->tab.reset(f->ctx, tab.size() * 8);
    f->tab_used.reset(f->ctx, cap);
    KG_HIP(hipMemcpyAsync(f->tab.get(), tab.data(), tab.size() * 8, hipMemcpyHostToDevice, f->s));
    KG_HIP(hipMemcpyAsync(f->tab_used.get(), used.data(), cap, hipMemcpyHostToDevice, f->s));
    f->tab_mask = cap - 1;
    D.tab_mask = cap - 1;
    D.tab = f->tab.get<uint64_t>();
    D.tab_used = f->tab_used.get<uint8_t>();
  } else if (f->form == KG_FORM_BLOOM) {
    D.bloom = f->bloom.get<uint32_t>();
    D.nblocks = f->nblocks;
  }
  f->hdev = D;
  f->kdev.reset(f->ctx, sizeof(KfDev));
  KG_HIP(hipMemcpyAsync(f->kdev.get(), &D, sizeof D, hipMemcpyHostToDevice, f->s));
  KG_HIP(hipStreamSynchronize(f->s));
  f->finished = true;
}

// KfProbe of a column for filter f; false when the column does not widen.
static bool probe_of(const kg_key_filter *f, const DCol &c, KfProbe &P) {
  P = KfProbe{};
  P.col = c;
  if (!widens(c.type, c.precision, c.scale, f->type, f->precision, f->scale)) return false;
  if (f->cls == KF_DEC) {
    const int from_scale = is_dec_type(c.type) ? c.scale : 0;
    u128 fac = 1;
    for (int k = from_scale; k < f->scale; k++) fac *= 10;
    if (fac != 1) {
      P.rescale = 1;
      const u128 lim = ((u128)1 << 127) / fac;  // |v| * fac < 2^127 below this
      P.fac[0] = (uint64_t)fac;
      P.fac[1] = (uint64_t)(fac >> 64);
      P.lim[0] = (uint64_t)(lim - 1);
      P.lim[1] = (uint64_t)((lim - 1) >> 64);
    }
  }
  return true;
}

static double row_cost(int form) {  // Member::row_cost_ns (member.rs:410-418)
  switch (form) {
    case KG_FORM_RANGE: return 0.5;
    case KG_FORM_SMALL: return 3.0;
    case KG_FORM_BLOOM: return 4.5;
    case KG_FORM_DENSE: return 1.2;
    default: return 5.0;
  }
}

static kg_key_filter *clone_filter(const kg_key_filter *a) {
  KfPtr hold(kf_new(a->ctx, a->type, a->precision, a->scale, &a->opt));
  kg_key_filter *f = hold.get();
  StreamScope kg_scope(f->s);
  f->tz_len = a->tz_len;
  memcpy(f->tz, a->tz, sizeof f->tz);
  f->has_range = a->has_range;
  f->lo = a->lo;
  f->hi = a->hi;
  f->flo = a->flo;
  f->fhi = a->fhi;
  f->nulls = a->nulls;
  f->rows = a->rows;
  f->values_ok = a->values_ok;
  f->values = a->values;
  f->form = a->form;
  f->members = a->members;
  f->dense_lo = a->dense_lo;
  f->dense_host = a->dense_host;
  f->empty = a->empty;
  if (a->form == KG_FORM_BLOOM) {
    bloom_alloc(f, a->nblocks);
    KG_HIP(hipMemcpyAsync(f->bloom.get(), a->bloom.get(), a->nblocks * 32,
                          hipMemcpyDeviceToDevice, f->s));
  }
  return hold.release();
}

// A set form's members (a dense set written out), ascending.
static std::vector<i128> form_members(const kg_key_filter *f) {
  if (f->form != KG_FORM_DENSE) return f->members;
  std::vector<i128> out;
  for (size_t w = 0; w < f->dense_host.size(); w++)
    for (int b = 0; b < 64; b++)
      if ((f->dense_host[w] >> b) & 1)
        out.push_back((i128)(int64_t)((uint64_t)f->dense_lo + (uint64_t)(w * 64 + b)));
  return out;
}

static bool is_set_form(int form) {
  return form == KG_FORM_SMALL || form == KG_FORM_EXACT || form == KG_FORM_DENSE;
}

}  // namespace kg

// ---------------------------------------------------------------- row stage state
struct kg_row_stage {
  kg_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  uint64_t min_rows = 65536;
  double max_pass_rate = 0.9;
  int n = 0;
  const kg_key_filter *f[KG_ROW_STAGE_SLOTS] = {};
  int column[KG_ROW_STAGE_SLOTS] = {};
  uint64_t generation[KG_ROW_STAGE_SLOTS] = {};
  int64_t off_at[KG_ROW_STAGE_SLOTS] = {};
  uint64_t window_rows[KG_ROW_STAGE_SLOTS] = {}, window_passed[KG_ROW_STAGE_SLOTS] = {};
  uint64_t tested[KG_ROW_STAGE_SLOTS] = {}, dropped[KG_ROW_STAGE_SLOTS] = {};
  DevBuf rs, counts, err, bitmap, tiles, tile_off, temp;
};

extern "C" {

void kg_filter_default_options(kg_filter_options *o) {
  std::memset(o, 0, sizeof *o);
  o->max_values = 50;          // krishiv.runtimeFilter.inList.maxValues (filter.rs:92-107)
  o->exact_max_values = 4096;  // EXACT_SET_MAX (member.rs:39)
  o->bloom = 1;
  o->bloom_max_items = 4000000;
  o->bloom_max_bytes = 8u << 20;
  o->bloom_fpp = 0.01;
  o->bloom_max_fpp = 0.05;
  o->bloom_bytes = 0;
  o->side = KG_SIDE_BOTH;
}

int kg_key_filter_create(kg_ctx *ctx, const kg_column *key_type, const kg_filter_options *opts,
                         kg_key_filter **out) {
  return guarded([&] {
    *out = nullptr;
    if (!key_type) fail(KG_ERR_INTERNAL, "no key type");
    kg_filter_options o;
    if (opts) o = *opts;
    else kg_filter_default_options(&o);
    if (o.exact_max_values > kExactMaxCap || o.max_values > kExactMaxCap)
      fail(KG_ERR_UNSUPPORTED, "exact sets of more than 65536 keys");
    KfPtr f(kf_new(ctx, key_type->type, key_type->precision, key_type->scale, &o));
    StreamScope kg_scope(f->s);
    const bool members = o.side == KG_SIDE_BOTH;
    f->hashes_ok = o.bloom && members;
    f->distinct_ok = members && distinct_max(f.get()) > 0;
    if (f->distinct_ok) {
      uint64_t cap = 64;
      while (cap < 4 * (uint64_t)distinct_max(f.get())) cap <<= 1;
      f->set_mask = cap - 1;
      f->set0.reset(ctx, cap * 8);
      const uint64_t e = f->cls == KF_DEC ? kHiEmpty : kSetEmpty;
      k_kf_fill<<<stream_grid((int64_t)cap, 256), 256, 0, f->s>>>(f->set0.get<uint64_t>(),
                                                                  (int64_t)cap, e);
      if (f->cls == KF_DEC) {
        f->set1.reset(ctx, cap * 8);
        f->setz.reset(ctx, cap * 8);
        KG_HIP(hipMemsetAsync(f->set1.get(), 0, cap * 8, f->s));
        k_kf_fill<<<stream_grid((int64_t)cap, 256), 256, 0, f->s>>>(f->setz.get<uint64_t>(),
                                                                    (int64_t)cap, kHiEmpty);
      }
      KG_HIP(hipGetLastError());
      f->setctl.reset(ctx, 8);
      KG_HIP(hipMemsetAsync(f->setctl.get(), 0, 8, f->s));
    }
    f->nh.reset(ctx, 8);
    KG_HIP(hipMemsetAsync(f->nh.get(), 0, 8, f->s));
    f->parts.reset(ctx, (kKfParts + 1) * sizeof(KfPart));
    KG_HIP(hipStreamSynchronize(f->s));
    *out = f.release();
  });
}

int kg_key_filter_add(kg_key_filter *f, const kg_column *key, const kg_column *include,
                      int64_t nrows) {
  return guarded([&] {
    StreamScope kg_scope(f->s);
    if (f->finished) fail(KG_ERR_INTERNAL, "keys added to a published filter");
    if (nrows <= 0) return;
    check_column(*key, nrows, "filter key");
    if (key->type != f->type || (is_dec_type(f->type) && key->scale != f->scale))
      fail(KG_ERR_INTERNAL, "runtime filter keys of another type");
    if (include) check_column(*include, nrows, "include column");
    Staged st(f->ctx, f->s);
    if (f->hashes_ok) {
      const uint64_t need = (f->nhashes + (uint64_t)nrows) * 8;
      if (need > f->hashes.bytes()) {
        DevBuf bigger(f->ctx, std::max<uint64_t>(need, f->hashes.bytes() * 2));
        if (f->nhashes)
          KG_HIP(hipMemcpyAsync(bigger.get(), f->hashes.get(), f->nhashes * 8,
                                hipMemcpyDeviceToDevice, f->s));
        f->hashes = std::move(bigger);
      }
    }
    KfAdd A = add_args(f);
    A.key = st.add(*key, nrows);
    if (include) {
      A.inc = st.add(*include, nrows);
      A.has_inc = 1;
    }
    const int grid = (int)std::min<int64_t>(kKfParts, ceil_div(nrows, kKfBlock * 4));
    {
      ProfScope prof("filter_add", f->s);
      k_kf_add<<<grid, kKfBlock, 0, f->s>>>(A, nrows);
      KG_HIP(hipGetLastError());
      k_kf_fold<<<1, 1, 0, f->s>>>(A.parts, grid);
      KG_HIP(hipGetLastError());
    }
    KfPart t{};
    unsigned long long nh = 0;
    KG_HIP(hipMemcpyAsync(&t, A.parts + grid, sizeof t, hipMemcpyDeviceToHost, f->s));
    KG_HIP(hipMemcpyAsync(&nh, f->nh.get(), 8, hipMemcpyDeviceToHost, f->s));
    KG_HIP(hipStreamSynchronize(f->s));
    f->nulls |= t.nulls != 0;
    if (t.rows == 0) return;
    // add_range (filter.rs:286-336)
    if (f->cls == KF_FLOAT) {
      f->nan |= t.nan != 0;
      f->flo = f->has_range ? std::fmin(f->flo, t.flo) : t.flo;
      f->fhi = f->has_range ? std::fmax(f->fhi, t.fhi) : t.fhi;
    } else {
      const i128 lo = mk(t.lo[0], t.lo[1]), hi = mk(t.hi[0], t.hi[1]);
      f->lo = f->has_range ? std::min(f->lo, lo) : lo;
      f->hi = f->has_range ? std::max(f->hi, hi) : hi;
    }
    f->has_range = true;
    f->rows += t.rows;
    f->nhashes = nh;
    if (f->hashes_ok && f->rows > f->opt.bloom_max_items) drop_hashes(f);
    if (f->distinct_ok) check_set(f);
  });
}

int kg_key_filter_merge_builder(kg_key_filter *a, const kg_key_filter *b) {
  return guarded([&] {
    StreamScope kg_scope(a->s);
    if (a->finished || b->finished) fail(KG_ERR_INTERNAL, "a builder merge of published filters");
    if (a->type != b->type || a->scale != b->scale)
      fail(KG_ERR_INTERNAL, "runtime filter builders of different key types merged");
    KG_HIP(hipStreamSynchronize(b->s));
    if (b->has_range) {
      if (!a->has_range) {
        a->lo = b->lo;
        a->hi = b->hi;
        a->flo = b->flo;
        a->fhi = b->fhi;
      } else {
        a->lo = std::min(a->lo, b->lo);
        a->hi = std::max(a->hi, b->hi);
        a->flo = std::fmin(a->flo, b->flo);
        a->fhi = std::fmax(a->fhi, b->fhi);
      }
      a->has_range = true;
    }
    a->nan |= b->nan;
    a->nulls |= b->nulls;
    a->rows += b->rows;
    if (a->distinct_ok && b->distinct_ok) {
      std::vector<i128> m = set_members(const_cast<kg_key_filter *>(b));
      if (!m.empty()) {
        std::vector<uint64_t> v(3 * m.size());
        for (size_t k = 0; k < m.size(); k++) {
          v[3 * k] = lo_of(m[k]);
          v[3 * k + 1] = hi_of(m[k]);
          v[3 * k + 2] = member_hash(a, m[k]);
        }
        DevBuf d(a->ctx, v.size() * 8);
        KG_HIP(hipMemcpyAsync(d.get(), v.data(), v.size() * 8, hipMemcpyHostToDevice, a->s));
        k_kf_set_merge<<<stream_grid((int64_t)m.size(), 256), 256, 0, a->s>>>(
            add_args(a), d.get<uint64_t>(), (int64_t)m.size());
        KG_HIP(hipGetLastError());
        check_set(a);
      }
    } else if (a->distinct_ok) {
      drop_set(a);
    }
    const bool keep = a->rows <= a->opt.bloom_max_items;
    if (a->hashes_ok && b->hashes_ok && keep) {
      DevBuf both(a->ctx, (a->nhashes + b->nhashes) * 8 + 8);
      if (a->nhashes)
        KG_HIP(hipMemcpyAsync(both.get(), a->hashes.get(), a->nhashes * 8, hipMemcpyDeviceToDevice,
                              a->s));
      if (b->nhashes)
        KG_HIP(hipMemcpyAsync(both.get<uint64_t>() + a->nhashes, b->hashes.get(), b->nhashes * 8,
                              hipMemcpyDeviceToDevice, a->s));
      a->hashes = std::move(both);
      a->nhashes += b->nhashes;
      const unsigned long long nh = a->nhashes;
      KG_HIP(hipMemcpyAsync(a->nh.get(), &nh, 8, hipMemcpyHostToDevice, a->s));
      KG_HIP(hipStreamSynchronize(a->s));
    } else if (a->hashes_ok) {
      drop_hashes(a);
    }
  });
}

int kg_key_filter_finish(kg_key_filter *f) {
  return guarded([&] {
    StreamScope kg_scope(f->s);
    if (f->finished) fail(KG_ERR_INTERNAL, "filter already published");
    // the range, NaN-free, opened per RangeSide (key.rs:403-432)
    bool has_range = f->has_range && !f->nan;
    if (has_range && f->opt.side != KG_SIDE_BOTH) {
      const bool upper = f->opt.side == KG_SIDE_UPPER;
      if (f->cls == KF_INT) {
        i128 llo, lhi;
        switch (f->type) {
          case KG_T_INT8: llo = INT8_MIN; lhi = INT8_MAX; break;
          case KG_T_INT16: llo = INT16_MIN; lhi = INT16_MAX; break;
          case KG_T_INT32: case KG_T_DATE32: llo = INT32_MIN; lhi = INT32_MAX; break;
          case KG_T_INT64: case KG_T_TIMESTAMP: llo = INT64_MIN; lhi = INT64_MAX; break;
          default: has_range = false; llo = lhi = 0;
        }
        if (upper) f->lo = llo;
        else f->hi = lhi;
      } else if (f->cls == KF_DEC) {
        i128 lim = 1;
        for (int k = 0; k < f->precision; k++) lim *= 10;
        lim -= 1;
        if (upper) f->lo = -lim;
        else f->hi = lim;
      } else {
        if (upper) f->flo = -INFINITY;
        else has_range = false;
      }
    }
    f->has_range = has_range;
    std::vector<i128> m;
    if (f->distinct_ok) {
      m = set_members(f);
      if ((uint64_t)m.size() <= f->opt.max_values) {
        f->values_ok = true;
        f->values = m;
      }
    }
    const bool dense = f->distinct_ok && f->has_range && f->cls == KF_INT &&
                       f->hi - f->lo + 1 == (i128)m.size();
    if (dense) {
      f->form = KG_FORM_RANGE;
    } else if (f->distinct_ok && (uint64_t)m.size() <= f->opt.exact_max_values) {
      f->members = m;
      set_exact_form(f);
    } else if (f->hashes_ok && f->nhashes > 0) {
      // bloom() (filter.rs:606-624)
      uint64_t bytes = f->opt.bloom_bytes ? f->opt.bloom_bytes
                                          : kg_bloom_bytes_for(f->nhashes, f->opt.bloom_fpp);
      bytes = std::min<uint64_t>(bytes, f->opt.bloom_max_bytes);
      const uint64_t nb = std::max<uint64_t>(1, (bytes + 31) / 32);
      if (sbbf_expected_fpp(nb, f->nhashes) <= f->opt.bloom_max_fpp) {
        bloom_alloc(f, nb);
        bloom_put_hashes(f, f->hashes.get<uint64_t>(), (int64_t)f->nhashes);
        f->form = KG_FORM_BLOOM;
      } else {
        f->form = KG_FORM_RANGE;
      }
    } else {
      f->form = KG_FORM_RANGE;
    }
    if (f->form == KG_FORM_DENSE) f->members = m;
    if (!is_set_form(f->form)) f->members.clear();
    f->empty = f->rows == 0;
    drop_set(f);
    drop_hashes(f);
    f->parts.release();
    publish(f);
  });
}

int kg_key_filter_info_get(const kg_key_filter *f, kg_key_filter_info *o) {
  return guarded([&] {
    *o = kg_key_filter_info{};
    o->form = f->form;
    o->has_range = f->has_range;
    o->members = is_set_form(f->form) ? (int64_t)f->members.size() : 0;
    o->values = f->values_ok ? (int64_t)f->values.size() : -1;
    o->lo[0] = lo_of(f->lo);
    o->lo[1] = hi_of(f->lo);
    o->hi[0] = lo_of(f->hi);
    o->hi[1] = hi_of(f->hi);
    o->flo = f->flo;
    o->fhi = f->fhi;
    o->nulls_pass = f->nulls;
    o->empty = f->empty;
    o->keys = f->rows;
    o->bloom_bytes = f->form == KG_FORM_BLOOM ? f->nblocks * 32 : 0;
    o->row_cost_ns = row_cost(f->form);
  });
}

int kg_key_filter_members(const kg_key_filter *f, uint64_t *out, int64_t *n) {
  return guarded([&] {
    const std::vector<i128> m = form_members(f);
    if (out) {
      if (*n < (int64_t)m.size()) fail(KG_ERR_INTERNAL, "member buffer too small");
      for (size_t k = 0; k < m.size(); k++) {
        out[2 * k] = lo_of(m[k]);
        out[2 * k + 1] = hi_of(m[k]);
      }
    }
    *n = (int64_t)m.size();
  });
}

int kg_key_filter_evaluate(kg_key_filter *f, const kg_column *col, int64_t nrows,
                           int32_t *out_rows, int64_t *out_count) {
  return guarded([&] {
    StreamScope kg_scope(f->s);
    *out_count = 0;
    if (!f->finished) fail(KG_ERR_INTERNAL, "evaluate before finish");
    if (nrows <= 0) return;
    if (nrows > INT32_MAX) fail(KG_ERR_UNSUPPORTED, "a probe batch of 2^31 rows or more");
    check_column(*col, nrows, "probe column");
    Staged st(f->ctx, f->s);
    KfProbe P;
    if (!probe_of(f, st.add(*col, nrows), P))
      fail(KG_ERR_UNSUPPORTED, "the column does not widen to the filter's key type");
    const int64_t tiles = ceil_div(nrows, kKfTile);
    f->bitmap.ensure(f->ctx, (uint64_t)ceil_div(nrows, 64) * 8 + 8);
    f->tiles.ensure(f->ctx, (uint64_t)tiles * 4);
    f->tile_off.ensure(f->ctx, (uint64_t)(tiles + 1) * 8);
    f->temp.ensure(f->ctx, (uint64_t)(scan_partials(tiles) + 2) * 8);
    f->probe.ensure(f->ctx, sizeof(KfProbe));
    f->err.ensure(f->ctx, 4);
    uint64_t *total = f->temp.get<uint64_t>() + scan_partials(tiles);
    KG_HIP(hipMemcpyAsync(f->probe.get(), &P, sizeof P, hipMemcpyHostToDevice, f->s));
    KG_HIP(hipMemsetAsync(f->err.get(), 0, 4, f->s));
    {
      ProfScope prof("filter_eval", f->s);
      k_kf_eval<<<(unsigned)tiles, kKfBlock, 0, f->s>>>(
          f->kdev.get<KfDev>(), f->probe.get<KfProbe>(), nrows, f->bitmap.get<uint64_t>(),
          f->tiles.get<uint32_t>(), f->err.get<unsigned int>());
      KG_HIP(hipGetLastError());
    }
    exclusive_scan(f->tiles.get<uint32_t>(), f->tile_off.get<uint64_t>(), tiles,
                   f->temp.get<uint64_t>(), f->s, nullptr, total);
    k_bitmap_rows<kKfBlock><<<(unsigned)tiles, kKfBlock, 0, f->s>>>(
        f->bitmap.get<uint64_t>(), nrows, f->tile_off.get<uint64_t>(), 0, out_rows);
    KG_HIP(hipGetLastError());
    uint64_t count = 0;
    unsigned int e = 0;
    KG_HIP(hipMemcpyAsync(&count, total, 8, hipMemcpyDeviceToHost, f->s));
    KG_HIP(hipMemcpyAsync(&e, f->err.get(), 4, hipMemcpyDeviceToHost, f->s));
    KG_HIP(hipStreamSynchronize(f->s));
    if (e) fail(KG_ERR_UNSUPPORTED, "a decimal probe value beyond its precision");
    *out_count = (int64_t)count;
  });
}

// ---- codec.rs ----------------------------------------------------------------
static int type_tag(int t) {
  switch (t) {
    case KG_T_BOOL: return 1;
    case KG_T_INT8: return 2;
    case KG_T_INT16: return 3;
    case KG_T_INT32: return 4;
    case KG_T_INT64: return 5;
    case KG_T_FLOAT32: return 6;
    case KG_T_FLOAT64: return 7;
    case KG_T_DECIMAL128: return 8;
    case KG_T_DATE32: return 9;
    case KG_T_TIMESTAMP: return 10;
    case KG_T_DECIMAL64: return 17;
    default: return 0;
  }
}

int kg_key_filter_encode(const kg_key_filter *f, uint8_t *bytes, uint64_t *len) {
  return guarded([&] {
    if (!f->finished) fail(KG_ERR_INTERNAL, "encode before finish");
    std::vector<uint8_t> w;
    auto put = [&](const void *p, size_t n) {
      const uint8_t *b = static_cast<const uint8_t *>(p);
      w.insert(w.end(), b, b + n);
    };
    auto u8 = [&](uint8_t v) { w.push_back(v); };
    auto u32 = [&](uint32_t v) { put(&v, 4); };
    auto u64 = [&](uint64_t v) { put(&v, 8); };
    auto i128w = [&](i128 v) { put(&v, 16); };
    u8(1);  // FORMAT_VERSION
    u8(1);  // RowHash::Xxh64Seed0
    u8(1);  // NORMALIZATION_VERSION
    u8((uint8_t)type_tag(f->type));
    if (is_dec_type(f->type)) {
      u8((uint8_t)f->precision);
      u8((uint8_t)(int8_t)f->scale);
    } else if (f->type == KG_T_TIMESTAMP) {
      if (f->tz_len < 0) {
        u8(0);
      } else {
        u8(1);
        u32((uint32_t)f->tz_len);
        put(f->tz, (size_t)f->tz_len);
      }
    }
    u8((uint8_t)((f->nulls ? 1 : 0) | (f->empty ? 2 : 0)));
    u64(f->rows);
    if (!f->has_range) {
      u8(0);
    } else if (f->cls == KF_INT) {
      u8(1);
      u64((uint64_t)(int64_t)f->lo);
      u64((uint64_t)(int64_t)f->hi);
    } else if (f->cls == KF_DEC) {
      u8(2);
      i128w(f->lo);
      i128w(f->hi);
    } else {
      u8(3);
      put(&f->flo, 8);
      put(&f->fhi, 8);
    }
    if (!f->values_ok) {
      u8(0);
    } else {
      u8(1);
      u32((uint32_t)f->values.size());
      for (i128 v : f->values) {
        u8(0);
        i128w(v);
      }
    }
    if (f->form == KG_FORM_RANGE) {
      u8(0);
    } else if (is_set_form(f->form)) {
      u8(f->form == KG_FORM_SMALL ? 1 : 2);
      const std::vector<i128> m = form_members(f);
      u32((uint32_t)m.size());
      for (i128 v : m) {
        if (f->cls == KF_DEC) i128w(v);
        else u64(lo_of(v));
      }
    } else {
      u8(3);
      u32((uint32_t)(f->nblocks * 32));
      const size_t at = w.size();
      w.resize(at + f->nblocks * 32);
      KG_HIP(hipMemcpyAsync(w.data() + at, f->bloom.get(), f->nblocks * 32, hipMemcpyDeviceToHost,
                            f->s));
      KG_HIP(hipStreamSynchronize(f->s));
    }
    if (bytes) {
      if (*len < w.size()) fail(KG_ERR_INTERNAL, "encode buffer too small");
      std::memcpy(bytes, w.data(), w.size());
    }
    *len = w.size();
  });
}

int kg_key_filter_decode(kg_ctx *ctx, const uint8_t *bytes, uint64_t len, kg_key_filter **out) {
  return guarded([&] {
    *out = nullptr;
    uint64_t at = 0;
    auto bad = [](const char *what) { fail(KG_ERR_INTERNAL, std::string("runtime filter encoding: ") + what); };
    auto take = [&](uint64_t n) -> const uint8_t * {
      if (len - at < n) bad("truncated");
      const uint8_t *p = bytes + at;
      at += n;
      return p;
    };
    auto u8 = [&]() { return *take(1); };
    auto u32 = [&]() { uint32_t v; std::memcpy(&v, take(4), 4); return v; };
    auto u64 = [&]() { uint64_t v; std::memcpy(&v, take(8), 8); return v; };
    auto i128r = [&]() { i128 v; std::memcpy(&v, take(16), 16); return v; };
    auto count = [&]() {
      const uint64_t n = u32();
      if (n > len - at) bad("a count beyond the input");
      return n;
    };
    if (len < 3 || bytes[0] != 1 || bytes[1] != 1 || bytes[2] != 1) bad("an unsupported version");
    at = 3;
    int type = 0;
    switch (u8()) {
      case 1: type = KG_T_BOOL; break;
      case 2: type = KG_T_INT8; break;
      case 3: type = KG_T_INT16; break;
      case 4: type = KG_T_INT32; break;
      case 5: type = KG_T_INT64; break;
      case 6: type = KG_T_FLOAT32; break;
      case 7: type = KG_T_FLOAT64; break;
      case 8: type = KG_T_DECIMAL128; break;
      case 9: type = KG_T_DATE32; break;
      case 10: type = KG_T_TIMESTAMP; break;
      case 17: type = KG_T_DECIMAL64; break;
      default: bad("an unknown key type");
    }
    int p = 0, sc = 0, tzl = -1;
    char tz[64] = {};
    if (is_dec_type(type)) {
      p = u8();
      sc = (int8_t)u8();
    } else if (type == KG_T_TIMESTAMP) {
      const int z = u8();
      if (z == 1) {
        const uint64_t n = u32();
        const uint8_t *b = take(n);
        if (n > 63) fail(KG_ERR_UNSUPPORTED, "a time zone name longer than 63 bytes");
        std::memcpy(tz, b, n);
        tzl = (int)n;
      } else if (z != 0) {
        bad("an unknown time zone flag");
      }
    }
    KfPtr f(kf_new(ctx, type, p, sc, nullptr));
    StreamScope kg_scope(f->s);
    f->tz_len = tzl;
    std::memcpy(f->tz, tz, sizeof tz);
    const uint8_t flags = u8();
    if (flags > 7) bad("unknown flags");
    if (flags & 4) fail(KG_ERR_UNSUPPORTED, "a top-N bound in a join runtime filter");
    f->nulls = flags & 1;
    f->empty = (flags >> 1) & 1;
    f->rows = u64();
    const int rt = u8();
    if (rt == 1 && f->cls == KF_INT) {
      f->lo = (int64_t)u64();
      f->hi = (int64_t)u64();
      f->has_range = true;
    } else if (rt == 2 && f->cls == KF_DEC) {
      f->lo = i128r();
      f->hi = i128r();
      f->has_range = true;
    } else if (rt == 3 && f->cls == KF_FLOAT) {
      const uint64_t a = u64(), b = u64();
      std::memcpy(&f->flo, &a, 8);
      std::memcpy(&f->fhi, &b, 8);
      f->has_range = true;
    } else if (rt != 0) {
      bad("a range of another type");
    }
    const int vt = u8();
    if (vt == 1) {
      const uint64_t n = count();
      f->values_ok = true;
      for (uint64_t k = 0; k < n; k++) {
        if (u8() != 0) bad("a value of another type");
        f->values.push_back(i128r());
      }
    } else if (vt != 0) {
      bad("an unknown value list");
    }
    const int mt = u8();
    if (mt == 1 || mt == 2) {
      const uint64_t n = count();
      if (mt == 1 && n > (uint64_t)kSmallMax) bad("a small set beyond its cap");
      for (uint64_t k = 0; k < n; k++)
        f->members.push_back(f->cls == KF_DEC ? i128r() : (i128)(int64_t)u64());
      if (mt == 1) {
        f->form = KG_FORM_SMALL;
      } else {
        // an exact set is re-placed by hash: a repeated member is held once
        std::sort(f->members.begin(), f->members.end());
        f->members.erase(std::unique(f->members.begin(), f->members.end()), f->members.end());
        f->form = KG_FORM_EXACT;
        if (f->cls == KF_INT && dense_of(f.get(), f->members)) f->form = KG_FORM_DENSE;
      }
      if ((f->form == KG_FORM_EXACT && f->members.size() > kExactMaxCap))
        fail(KG_ERR_UNSUPPORTED, "an exact set of more than 65536 keys");
    } else if (mt == 3) {
      const uint64_t n = u32();
      const uint8_t *b = take(n);
      if (n == 0 || n % 32 != 0) bad("a malformed Bloom");
      bloom_alloc(f.get(), n / 32);
      KG_HIP(hipMemcpyAsync(f->bloom.get(), b, n, hipMemcpyHostToDevice, f->s));
      KG_HIP(hipStreamSynchronize(f->s));
      f->form = KG_FORM_BLOOM;
    } else if (mt == 0) {
      f->form = KG_FORM_RANGE;
    } else {
      bad("an unknown row form");
    }
    if (at != len) bad("trailing bytes");
    f->distinct_ok = f->hashes_ok = false;
    publish(f.get());
    *out = f.release();
  });
}

int kg_key_filter_merge(const kg_key_filter *a, const kg_key_filter *b,
                        const kg_filter_options *opts, kg_key_filter **out) {
  return guarded([&] {
    *out = nullptr;
    if (!a->finished || !b->finished) fail(KG_ERR_INTERNAL, "merge of unpublished filters");
    if (a->type != b->type || a->precision != b->precision || a->scale != b->scale ||
        a->tz_len != b->tz_len || std::memcmp(a->tz, b->tz, sizeof a->tz) != 0)
      fail(KG_ERR_INTERNAL, "runtime filters of different key types merged");
    kg_filter_options o;
    if (opts) o = *opts;
    else kg_filter_default_options(&o);
    const bool nulls = a->nulls || b->nulls;
    uint64_t keys = a->rows + b->rows;
    if (keys < a->rows) keys = UINT64_MAX;
    if (a->empty || b->empty) {
      kg_key_filter *f = clone_filter(b->empty ? a : b);
      KfPtr hold(f);
      StreamScope kg_scope(f->s);
      f->nulls = nulls;
      f->rows = keys;
      publish(f);
      *out = hold.release();
      return;
    }
    KfPtr f(clone_filter(a));
    StreamScope kg_scope(f->s);
    f->opt = o;
    f->nulls = nulls;
    f->rows = keys;
    f->empty = false;
    if (a->has_range && b->has_range) {
      f->lo = std::min(a->lo, b->lo);
      f->hi = std::max(a->hi, b->hi);
      f->flo = std::fmin(a->flo, b->flo);
      f->fhi = std::fmax(a->fhi, b->fhi);
    } else {
      f->has_range = false;
    }
    f->values_ok = false;
    f->values.clear();
    if (a->values_ok && b->values_ok) {
      std::vector<i128> v = a->values;
      v.insert(v.end(), b->values.begin(), b->values.end());
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      if ((uint64_t)v.size() <= o.max_values) {
        f->values_ok = true;
        f->values = v;
      }
    }
    // merge_members (merge.rs:119-169)
    uint64_t planned = 0;
    if (o.bloom_bytes)
      planned = std::max<uint64_t>(1, (std::min(o.bloom_bytes, o.bloom_max_bytes) + 31) / 32) * 32;
    const bool ua = a->form == KG_FORM_BLOOM && a->nblocks * 32 != planned;
    const bool ub = b->form == KG_FORM_BLOOM && b->nblocks * 32 != planned;
    f->members.clear();
    f->dense_host.clear();
    if (ua || ub || a->form == KG_FORM_RANGE || b->form == KG_FORM_RANGE) {
      f->bloom.release();
      f->form = KG_FORM_RANGE;
    } else if (a->form == KG_FORM_BLOOM && b->form == KG_FORM_BLOOM) {
      // equal planned sizes: OR on the device (union_with, bloom.rs:153-163)
      k_kf_bloom_or<<<stream_grid((int64_t)f->nblocks * 8, 256), 256, 0, f->s>>>(
          f->bloom.get<uint32_t>(), b->bloom.get<uint32_t>(), (int64_t)f->nblocks * 8);
      KG_HIP(hipGetLastError());
      f->form = KG_FORM_BLOOM;
    } else if (a->form == KG_FORM_BLOOM || b->form == KG_FORM_BLOOM) {
      const kg_key_filter *bl = a->form == KG_FORM_BLOOM ? a : b;
      const kg_key_filter *st = a->form == KG_FORM_BLOOM ? b : a;
      if (bl != a) {
        bloom_alloc(f.get(), bl->nblocks);
        KG_HIP(hipMemcpyAsync(f->bloom.get(), bl->bloom.get(), bl->nblocks * 32,
                              hipMemcpyDeviceToDevice, f->s));
      }
      bloom_put_members(f.get(), st, form_members(st));
      f->form = KG_FORM_BLOOM;
    } else {
      std::vector<i128> all = form_members(a), mb = form_members(b);
      all.insert(all.end(), mb.begin(), mb.end());
      std::sort(all.begin(), all.end());
      all.erase(std::unique(all.begin(), all.end()), all.end());
      if ((uint64_t)all.size() <= o.exact_max_values) {
        f->bloom.release();
        f->members = all;
        set_exact_form(f.get());
      } else if (planned) {
        bloom_alloc(f.get(), planned / 32);
        bloom_put_members(f.get(), a, all);
        f->form = KG_FORM_BLOOM;
      } else {
        f->form = KG_FORM_RANGE;
      }
    }
    if (f->form == KG_FORM_BLOOM && pass_rate(f.get()) > o.bloom_max_fpp) {
      f->bloom.release();
      f->form = KG_FORM_RANGE;
    }
    if (f->form != KG_FORM_BLOOM) {
      f->bloom.release();
      f->nblocks = 0;
    }
    if (!is_set_form(f->form)) f->members.clear();
    publish(f.get());
    *out = f.release();
  });
}

void kg_key_filter_destroy(kg_key_filter *f) {
  if (!f) return;
  if (f->s) (void)hipStreamSynchronize(f->s);
  hipStream_t s = f->s;
  delete f;
  destroy_stream(s);
}

// ---- the row stage (operators/runtime_filter.rs:240-395) ---------------------
int kg_row_stage_create(kg_ctx *ctx, uint64_t min_rows, uint32_t max_pass_percent,
                        kg_row_stage **out) {
  return guarded([&] {
    *out = nullptr;
    auto s = std::make_unique<kg_row_stage>();
    s->ctx = ctx;
    s->min_rows = min_rows;
    s->max_pass_rate = (double)max_pass_percent / 100.0;
    for (int k = 0; k < KG_ROW_STAGE_SLOTS; k++) s->off_at[k] = -1;
    s->s = make_stream(ctx->device);
    *out = s.release();
  });
}

int kg_row_stage_set(kg_row_stage *s, int32_t slot, const kg_key_filter *f, int32_t column) {
  return guarded([&] {
    if (slot < 0 || slot >= KG_ROW_STAGE_SLOTS) fail(KG_ERR_INTERNAL, "row stage slot");
    if (f && !f->finished) fail(KG_ERR_INTERNAL, "an unpublished filter in a row stage");
    while (s->n <= slot) {
      s->f[s->n] = nullptr;
      s->off_at[s->n] = -1;
      s->n++;
    }
    s->f[slot] = f;
    s->column[slot] = column;
    if (f) s->generation[slot]++;
  });
}

int kg_row_stage_start_partition(kg_row_stage *s) {
  return guarded([&] {
    for (int k = 0; k < s->n; k++) {
      s->window_rows[k] = s->window_passed[k] = 0;
      s->off_at[k] = -1;
    }
  });
}

int kg_row_stage_apply(kg_row_stage *s, int32_t ncols, const kg_column *cols, int64_t nrows,
                       int32_t *out_rows, int64_t *out_count) {
  return guarded([&] {
    StreamScope kg_scope(s->s);
    *out_count = -1;
    if (nrows <= 0) return;
    if (nrows > INT32_MAX) fail(KG_ERR_UNSUPPORTED, "a batch of 2^31 rows or more");
    // the applicable filters, cheapest first (stable)
    int order[KG_ROW_STAGE_SLOTS], m = 0;
    for (int k = 0; k < s->n; k++) {
      if (s->off_at[k] >= 0 && (uint64_t)s->off_at[k] == s->generation[k]) continue;
      s->off_at[k] = -1;
      if (s->f[k]) order[m++] = k;
    }
    std::stable_sort(order, order + m, [&](int x, int y) {
      return row_cost(s->f[x]->form) < row_cost(s->f[y]->form);
    });
    Staged st(s->ctx, s->s);
    RsDev R{};
    int slot_of[KG_ROW_STAGE_SLOTS];
    std::vector<DCol> staged(ncols);
    std::vector<bool> have_col(ncols, false);
    for (int q = 0; q < m; q++) {
      const int k = order[q];
      const int c = s->column[k];
      if (c < 0 || c >= ncols) fail(KG_ERR_INTERNAL, "a runtime filter's column is missing");
      if (!have_col[c]) {
        check_column(cols[c], nrows, "filtered column");
        staged[c] = st.add(cols[c], nrows);
        have_col[c] = true;
      }
      KfProbe P;
      if (!probe_of(s->f[k], staged[c], P)) continue;  // the filter does not apply
      R.f[R.k] = s->f[k]->hdev;
      R.p[R.k] = P;
      slot_of[R.k] = k;
      R.k++;
    }
    if (R.k == 0) return;
    const int64_t tiles = ceil_div(nrows, kKfTile);
    s->rs.ensure(s->ctx, sizeof(RsDev));
    s->counts.ensure(s->ctx, 2 * KG_ROW_STAGE_SLOTS * 8);
    s->err.ensure(s->ctx, 4);
    s->bitmap.ensure(s->ctx, (uint64_t)ceil_div(nrows, 64) * 8 + 8);
    s->tiles.ensure(s->ctx, (uint64_t)tiles * 4);
    s->tile_off.ensure(s->ctx, (uint64_t)(tiles + 1) * 8);
    s->temp.ensure(s->ctx, (uint64_t)(scan_partials(tiles) + 2) * 8);
    KG_HIP(hipMemcpyAsync(s->rs.get(), &R, sizeof R, hipMemcpyHostToDevice, s->s));
    KG_HIP(hipMemsetAsync(s->counts.get(), 0, 2 * KG_ROW_STAGE_SLOTS * 8, s->s));
    KG_HIP(hipMemsetAsync(s->err.get(), 0, 4, s->s));
    {
      ProfScope prof("row_stage", s->s);
      k_rs_eval<<<(unsigned)tiles, kKfBlock, 0, s->s>>>(
          s->rs.get<RsDev>(), nrows, s->bitmap.get<uint64_t>(), s->tiles.get<uint32_t>(),
          s->counts.get<unsigned long long>(), s->err.get<unsigned int>());
      KG_HIP(hipGetLastError());
    }
    uint64_t *total = s->temp.get<uint64_t>() + scan_partials(tiles);
    exclusive_scan(s->tiles.get<uint32_t>(), s->tile_off.get<uint64_t>(), tiles,
                   s->temp.get<uint64_t>(), s->s, nullptr, total);
    uint64_t counts[2 * KG_ROW_STAGE_SLOTS], alive_total = 0;
    unsigned int e = 0;
    KG_HIP(hipMemcpyAsync(counts, s->counts.get(), sizeof counts, hipMemcpyDeviceToHost, s->s));
    KG_HIP(hipMemcpyAsync(&alive_total, total, 8, hipMemcpyDeviceToHost, s->s));
    KG_HIP(hipMemcpyAsync(&e, s->err.get(), 4, hipMemcpyDeviceToHost, s->s));
    KG_HIP(hipStreamSynchronize(s->s));
    if (e) fail(KG_ERR_UNSUPPORTED, "a decimal probe value beyond its precision");
    // replay the stage's accounting (runtime_filter.rs:357-393): a filter
    // after one that left fewer than half the rows was tested on the
    // survivors only
    const uint64_t n = (uint64_t)nrows;
    uint64_t alive = n;
    bool have = false;
    for (int q = 0; q < R.k; q++) {
      const int k = slot_of[q];
      const kg_key_filter *f = s->f[k];
      const bool narrowed = have && alive * 2 < n;
      const uint64_t tested = narrowed ? alive : n;
      const uint64_t passed = narrowed ? counts[2 * q + 1] : counts[2 * q];
      s->tested[k] += tested;
      s->dropped[k] += tested - passed;
      s->window_rows[k] += tested;
      s->window_passed[k] += passed;
      if (s->window_rows[k] >= s->min_rows) {
        const double cut = std::min(s->max_pass_rate, 1.0 - row_cost(f->form) / kDropSavesNs);
        const double rate = (double)s->window_passed[k] / (double)s->window_rows[k];
        if (rate > cut) s->off_at[k] = (int64_t)s->generation[k];
        s->window_rows[k] = s->window_passed[k] = 0;
      }
      alive = counts[2 * q + 1];
      have = true;
    }
    if (alive_total == n) return;  // nothing filtered
    k_bitmap_rows<kKfBlock><<<(unsigned)tiles, kKfBlock, 0, s->s>>>(
        s->bitmap.get<uint64_t>(), nrows, s->tile_off.get<uint64_t>(), 0, out_rows);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(s->s));
    *out_count = (int64_t)alive_total;
  });
}

int kg_row_stage_stats(const kg_row_stage *s, int32_t slot, uint64_t *tested, uint64_t *dropped,
                       int32_t *off) {
  return guarded([&] {
    if (slot < 0 || slot >= KG_ROW_STAGE_SLOTS) fail(KG_ERR_INTERNAL, "row stage slot");
    *tested = s->tested[slot];
    *dropped = s->dropped[slot];
    *off = s->off_at[slot] >= 0 && (uint64_t)s->off_at[slot] == s->generation[slot];
  });
}

void kg_row_stage_destroy(kg_row_stage *s) {
  if (!s) return;
  if (s->s) (void)hipStreamSynchronize(s->s);
  hipStream_t st = s->s;
  delete s;
  destroy_stream(st);
}

}  // extern "C"
