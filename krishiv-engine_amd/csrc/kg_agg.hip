// kg_agg.hip -- hash aggregation (HashAggregate / RowTable,
// crates/krishiv-exec/src/aggregate/sink.rs:1459-1585, rowtable.rs:816,1185)
// with Spark's sum/count/min/max/avg semantics (aggregate/mod.rs:588-775,
// accumulator/dense.rs:55-345, accumulator/helpers.rs:120-330,595-700).
//
// Device layout
//   keys[cap]        int64 key of each open-addressing slot (EMPTY = INT64_MIN);
//                    slot cap holds the key INT64_MIN itself, slot cap+1 the
//                    NULL key.
//   rec[(cap+2)*W]   W 64-bit state words per slot (array of structs: a
//                    row's atomics all land in one 128-byte line for W <= 16).
// Word 0 of every record counts the group's rows.  A column used by an
// aggregate owns a NULL-count word (valid rows = rows - NULLs, so non-NULL
// rows never touch it).  Integer and decimal sums are exact two's-complement
// 128-bit (decimal128: 192-bit) accumulators added with carries; float sums
// and averages are float64 atomics; min/max are 64-bit atomic min/max on
// order-preserving encodings.
//
// Kernel k_agg_update: each workgroup (8 waves) aggregates a chunk of rows
// into an LDS table of SLOTS records first (linear probe of 8 slots, LDS
// CAS insert); rows whose key finds no LDS slot go straight to the global
// table.  At the end of the chunk the LDS records are merged into the global
// table (one set of atomics per distinct key per chunk), which takes the
// contention of heavy (Zipf) keys off the global atomics.  One-byte keys
// (the reference's direct path D1, aggregate/rowtable/direct.rs) index both
// tables directly.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <memory>

#include "kg_aggops.hpp"

namespace kg {
template <int SLOTS>
__global__ void __launch_bounds__(kAggBlock) k_agg_update(AggProg P, Table T, int64_t nrows,
                                                          int64_t chunk) {
  extern __shared__ uint64_t smem[];
  uint64_t *s_keys = smem;
  uint64_t *s_rec = smem + SLOTS;
  const int W = P.W;
  for (int64_t base = blockIdx.x * chunk; base < nrows; base += (int64_t)gridDim.x * chunk) {
    for (int t = threadIdx.x; t < SLOTS; t += blockDim.x) s_keys[t] = kEmpty;
    for (int t = threadIdx.x; t < SLOTS * W; t += blockDim.x) s_rec[t] = P.init[t % W];
    __syncthreads();
    int64_t end = base + chunk < nrows ? base + chunk : nrows;
    for (int64_t i = base + threadIdx.x; i < end; i += blockDim.x) {
      uint64_t key;
      int cls = row_key(P, i, key);
      int ls = -1;
      if (P.direct) {
        ls = cls == 2 ? 256 : (int)(key & 0xFF);
      } else if (cls == 0) {
        uint64_t s = (P.has_key ? slot_hash(T, key) : 0) & (SLOTS - 1);
        for (int p = 0; p < kLdsProbe; p++) {
          uint64_t k = s_keys[s];
          if (k == key) {
            ls = (int)s;
            break;
          }
          if (k == kEmpty) {
            unsigned long long prev = atomicCAS((unsigned long long *)&s_keys[s], kEmpty, key);
            if (prev == kEmpty || prev == key) {
              ls = (int)s;
              break;
            }
          }
          s = (s + 1) & (SLOTS - 1);
        }
      }
      if (ls >= 0) {
        if (P.direct) s_keys[ls] = cls == 2 ? 1 : 0;  // marks the direct slot used
        apply_row<true>(P, s_rec + (uint64_t)ls * W, i);
      } else {
        apply_row<false>(P, global_record(P, T, cls, key), i);
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < SLOTS; t += blockDim.x) {
      uint64_t k = s_keys[t];
      if (k == kEmpty) continue;
      const uint64_t *l = s_rec + (uint64_t)t * W;
      if (l[0] == 0) continue;  // no row
      uint64_t *g;
      if (P.direct) {
        g = global_record(P, T, t == 256 ? 2 : 0, (uint64_t)t);
      } else {
        g = global_record(P, T, 0, k);
      }
      merge_record(P, g, l);
    }
    __syncthreads();
  }
}

// Rows applied straight to the global table, writing each row's record
// index (used by batches whose checked sums are replayed in order).
__global__ void k_agg_rows_global(AggProg P, Table T, int64_t nrows, uint64_t *slots) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrows;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t key;
    int cls = row_key(P, i, key);
    uint64_t *r = global_record(P, T, cls, key);
    apply_row<false>(P, r, i);
    slots[i] = (uint64_t)(r - T.rec) / P.W;
  }
}

__global__ void k_agg_replay(AggProg P, CheckedOps C, const uint64_t *slots, int64_t nrows,
                             uint64_t *rec) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int64_t i = 0; i < nrows; i++) {
    uint64_t *r = rec + slots[i] * (uint64_t)P.W;
    for (int k = 0; k < C.n; k++) {
      const DCol &c = P.in[C.col[k]];
      if (!dev::valid(c, i)) continue;
      uint64_t *st = r + C.st_word[k];
      if (*st == 2 && !C.avg[k]) continue;
      uint64_t vlo;
      int64_t vhi;
      dev::load_i128(c, i, vlo, vhi);
      uint64_t *w = r + C.word[k];
      // current value as 192 bits (wide) or 128 bits
      uint64_t a0 = w[0], a1 = w[1], a2 = C.wide[k] ? w[2] : (uint64_t)((int64_t)w[1] >> 63);
      uint64_t b0 = vlo, b1 = (uint64_t)vhi, b2 = (uint64_t)(vhi >> 63);
      uint64_t s0 = a0 + b0, c0 = s0 < a0;
      uint64_t s1 = a1 + b1, c1 = s1 < a1;
      s1 += c0;
      c1 |= (c0 && s1 == 0);
      uint64_t s2 = a2 + b2 + c1;
      // magnitude of (s0, s1, s2) must be below the bound (which is < 2^128)
      bool negative = (int64_t)s2 < 0;
      uint64_t m0 = s0, m1 = s1, m2 = s2;
      if (negative) {
        m0 = ~m0 + 1;
        uint64_t cc = m0 == 0;
        m1 = ~m1 + cc;
        cc = cc && m1 == 0;
        m2 = ~m2 + cc;
      }
      bool fits = m2 == 0 && lt_mag(m0, m1, C.bound_lo[k], C.bound_hi[k]);
      if (fits) {
        w[0] = s0;
        w[1] = s1;
        if (C.wide[k]) w[2] = s2;
        if (!C.avg[k]) *st = 1;
      } else {
        *st = 2;
        if (!C.avg[k]) {
          w[0] = 0;
          w[1] = 0;
          if (C.wide[k]) w[2] = 0;
        }
      }
    }
  }
}

// Σ|v| of a column's valid rows, as a double rounded up (an upper bound).
__global__ void k_abs_sum(DCol c, int64_t nrows, double *out) {
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrows;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!dev::valid(c, i)) continue;
    uint64_t lo;
    int64_t hi;
    dev::load_i128(c, i, lo, hi);
    uint64_t ulo = lo, uhi = (uint64_t)hi;
    if (hi < 0) negate128(ulo, uhi);
    acc += (double)uhi * 18446744073709551616.0 + (double)ulo;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(out, acc);
}

__global__ void k_agg_output(OutProg O, const uint64_t *keys, int key_stride, const uint64_t *rec,
                             uint64_t nslots, const uint64_t *pos, void *out_key,
                             uint8_t *out_key_valid, int key_width, unsigned int *err) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < nslots;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t *r = rec + s * (uint64_t)O.W;
    uint64_t rows = r[0];
    if (rows == 0 && !(O.global && s == 0)) continue;
    int64_t o = (int64_t)pos[s];
    // key
    bool null_key = O.direct ? s == 256 : s == O.cap + 1;
    if (out_key) {
      uint64_t k = O.direct ? (uint64_t)(int64_t)(int8_t)(s & 0xFF)
                            : (s >= O.cap ? kEmpty : keys[s * (uint64_t)key_stride]);
      if (O.direct && O.key_type != KG_T_INT8) k = s & 0xFF;
      if (null_key) k = 0;
      switch (key_width) {
        case 1: static_cast<uint8_t *>(out_key)[o] = (uint8_t)k; break;
        case 2: static_cast<uint16_t *>(out_key)[o] = (uint16_t)k; break;
        case 4: static_cast<uint32_t *>(out_key)[o] = (uint32_t)k; break;
        default: static_cast<uint64_t *>(out_key)[o] = k;
      }
      if (O.key_type == KG_T_FLOAT32 && !null_key) {
        static_cast<uint32_t *>(out_key)[o] = (uint32_t)k;
      }
      if (!null_key && out_key_valid) set_bit(out_key_valid, o);
    }
    for (int j = 0; j < O.n; j++) {
      const OutAgg &a = O.a[j];
      const uint64_t *w = r + a.word;
      uint64_t valid_rows = a.nulls_word >= 0 ? rows - r[a.nulls_word] : rows;
      bool ok = true;
      switch (a.kind) {
        case R_COUNT_STAR: static_cast<int64_t *>(a.values)[o] = (int64_t)rows; break;
        case R_COUNT: static_cast<int64_t *>(a.values)[o] = (int64_t)valid_rows; break;
        case R_SUM_INT: {
          ok = valid_rows > 0;
          // exact 128-bit sum; the result wraps (non-ANSI) or must fit (ANSI)
          bool fits = (int64_t)w[1] == ((int64_t)w[0] >> 63);
          bool ovf = (a.st_word >= 0 && r[a.st_word] == 2) || (O.ansi && !fits);
          if (ok && ovf) atomicOr(err, 1u);
          static_cast<int64_t *>(a.values)[o] = (int64_t)w[0];
          break;
        }
        case R_SUM_F:
          ok = valid_rows > 0;
          static_cast<double *>(a.values)[o] = __longlong_as_double((long long)w[0]);
          break;
        case R_SUM_DEC: {
          ok = valid_rows > 0;
          if (!ok) break;
          if (a.st_word >= 0 && r[a.st_word] == 2) {
            ok = false;
            if (O.ansi) atomicOr(err, 1u);
            break;
          }
          uint64_t lo = w[0], hi = w[1];
          bool wide_ok = a.in_type != KG_T_DECIMAL128 || (int64_t)w[2] == ((int64_t)w[1] >> 63);
          uint64_t mlo = lo, mhi = hi;
          if (neg128(lo, hi)) negate128(mlo, mhi);
          bool fits = wide_ok && lt_mag(mlo, mhi, a.pow_lo, a.pow_hi);
          if (!fits) {
            ok = false;
            if (O.ansi) atomicOr(err, 1u);
            break;
          }
          store_i128(a, o, lo, hi);
          break;
        }
        case R_MINMAX_I: {
          ok = valid_rows > 0;
          int64_t v = (int64_t)w[0];
          switch (a.out_type == KG_T_INT8 || a.out_type == KG_T_UINT8 || a.out_type == KG_T_BOOL
                      ? 1
                      : a.out_type == KG_T_INT16 || a.out_type == KG_T_UINT16
                            ? 2
                            : a.out_type == KG_T_INT32 || a.out_type == KG_T_UINT32 ||
                                      a.out_type == KG_T_DATE32
                                  ? 4
                                  : 8) {
            case 1: static_cast<int8_t *>(a.values)[o] = (int8_t)v; break;
            case 2: static_cast<int16_t *>(a.values)[o] = (int16_t)v; break;
            case 4: static_cast<int32_t *>(a.values)[o] = (int32_t)v; break;
            default: static_cast<int64_t *>(a.values)[o] = v;
          }
          break;
        }
        case R_MINMAX_F: {
          ok = valid_rows > 0;
          double d = f64_unorder(w[0]);
          if (a.out_type == KG_T_FLOAT32) static_cast<float *>(a.values)[o] = (float)d;
          else static_cast<double *>(a.values)[o] = d;
          break;
        }
        case R_AVG_INT: {
          ok = valid_rows > 0;
          if (!ok) break;
          // the exact sum, rounded once to double
          uint64_t lo = w[0], hi = w[1];
          bool n = neg128(lo, hi);
          if (n) negate128(lo, hi);
          // one correctly rounded conversion of the 128-bit magnitude
          double s = (double)(((unsigned __int128)hi << 64) | lo);
          if (n) s = -s;
          static_cast<double *>(a.values)[o] = s / (double)valid_rows;
          break;
        }
        case R_AVG_F:
          ok = valid_rows > 0;
          static_cast<double *>(a.values)[o] =
              __longlong_as_double((long long)w[0]) / (double)valid_rows;
          break;
        case R_AVG_DEC: {
          ok = valid_rows > 0;
          if (!ok) break;
          if (a.st_word >= 0 && r[a.st_word] == 2) {
            ok = false;
            if (O.ansi) atomicOr(err, 1u);
            break;
          }
          // avg_fast (accumulator/helpers.rs:595-609): |sum| * 10^(s - ss)
          // / count, HALF_UP; the digit loop's value where that overflows.
          uint64_t lo = w[0], hi = w[1];
          bool negative = neg128(lo, hi);
          if (negative) negate128(lo, hi);
          int k = a.out_s - a.in_s;
          bool ovf = false;
          for (int t = 0; t < k && !ovf; t++) ovf = mul128(lo, hi, 10);
          uint64_t qlo, qhi, rem;
          if (!ovf) {
            divmod128(lo, hi, valid_rows, qlo, qhi, rem);
            if (rem >= valid_rows - rem) {
              qlo += 1;
              if (qlo == 0) qhi += 1;
            }
          } else {
            // |sum| >= 2^128 / 10^k: the scale-39 digit loop rounds the same
            // quotient; compute it as (|sum| / count) * 10^k + digits.
            uint64_t slo = w[0], shi = w[1];
            if (negative) negate128(slo, shi);
            divmod128(slo, shi, valid_rows, qlo, qhi, rem);
            for (int t = 0; t < k; t++) {
              if (mul128(qlo, qhi, 10)) {
                ovf = true;
                break;
              }
              unsigned __int128 r10 = (unsigned __int128)rem * 10;
              uint64_t dgt = (uint64_t)(r10 / valid_rows);
              rem = (uint64_t)(r10 % valid_rows);
              uint64_t old = qlo;
              qlo += dgt;
              if (qlo < old) qhi += 1;
            }
            if (!ovf) {
              unsigned __int128 r10 = (unsigned __int128)rem * 10;
              if ((uint64_t)(r10 / valid_rows) >= 5) {
                qlo += 1;
                if (qlo == 0) qhi += 1;
              }
              ovf = false;
            } else {
              ovf = true;
            }
          }
          if (ovf || !lt_mag(qlo, qhi, a.pow_lo, a.pow_hi)) {
            ok = false;
            if (O.ansi) atomicOr(err, 2u);
            break;
          }
          if (negative) negate128(qlo, qhi);
          store_i128(a, o, qlo, qhi);
          break;
        }
      }
      if (ok) set_bit(a.validity, o);
    }
  }
}

}  // namespace kg

// ======================================================================
// host side
// ======================================================================
namespace kg {

__global__ void k_fill_words(uint64_t *rec, uint64_t nrec, int W, const uint64_t *init_dev) {
  uint64_t total = nrec * (uint64_t)W;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x)
    rec[t] = init_dev[t % W];
}

__global__ void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
       t += (uint64_t)gridDim.x * blockDim.x)
    p[t] = v;
}

__global__ void k_rehash(AggProg P, const uint64_t *okeys, const uint64_t *orec, uint64_t ocap,
                         Table T) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < ocap + 2;
       s += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t d;
    if (s >= ocap) {
      d = T.cap + (s - ocap);
    } else {
      uint64_t k = okeys[s];
      if (k == kEmpty) continue;
      d = global_slot(T, k);
    }
    for (int w = 0; w < P.W; w++) T.rec[d * P.W + w] = orec[s * P.W + w];
  }
}

__global__ void k_group_flags(const uint64_t *rec, uint64_t nslots, int W, int global,
                              uint32_t *flags) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < nslots;
       s += (uint64_t)gridDim.x * blockDim.x)
    flags[s] = (rec[s * (uint64_t)W] > 0 || (global && s == 0)) ? 1u : 0u;
}

static unsigned __int128 pow10_u128(int p) {
  unsigned __int128 v = 1;
  for (int i = 0; i < p; i++) v *= 10;
  return v;
}

}  // namespace kg

using namespace kg;

struct kg_hash_agg {
  kg_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  bool has_key = false, direct = false, ansi = false;
  kg_column key_type{};
  std::vector<kg_column> in_types;
  AggProg prog{};
  OutProg out{};
  int slots = 512;  // LDS records per workgroup
  size_t lds_bytes = 0;
  // table
  DevBuf keys, rec, occ;
  uint64_t cap = 0;
  int log2cap = 0;
  uint64_t seed = 0x9E3779B97F4A7C15ULL;
  uint64_t occ_bound = 0;
  // ordered sums
  std::vector<CheckedSum> checked;
  uint64_t rows_seen = 0;
  DevBuf scratch_slots, scratch_abs;
  bool finalized = false;
  int64_t ngroups = 0;
  DevBuf pos;
};

namespace kg {

static uint64_t nslots(const kg_hash_agg *a) { return a->direct ? 257 : a->cap + 2; }

static Table table_of(kg_hash_agg *a) {
  Table t{};
  t.keys = a->keys.get<uint64_t>();
  t.rec = a->rec.get<uint64_t>();
  t.occupied = a->occ.get<unsigned long long>();
  t.cap = a->cap;
  t.log2cap = a->log2cap;
  t.mask = a->cap - 1;
  t.seed = a->seed;
  return t;
}

static void alloc_table(kg_hash_agg *a, uint64_t cap, DevBuf &keys, DevBuf &rec) {
  uint64_t ns = a->direct ? 257 : cap + 2;
  keys.reset(a->ctx, (a->direct ? 257 : cap) * 8);
  rec.reset(a->ctx, ns * (uint64_t)a->prog.W * 8);
  k_fill_u64<<<stream_grid(keys.bytes() / 8, 256), 256, 0, a->s>>>(keys.get<uint64_t>(),
                                                                   keys.bytes() / 8, kEmpty);
  KG_HIP(hipGetLastError());
  init_records(a->ctx, rec.get<uint64_t>(), ns, a->prog, a->s);
}

static int log2_ceil(uint64_t v) {
  int b = 0;
  while ((1ull << b) < v) b++;
  return b;
}

// Doubles the table (rehash of every key and record).
static void grow(kg_hash_agg *a, uint64_t min_cap) {
  uint64_t ncap = a->cap * 2;
  while (ncap < min_cap) ncap *= 2;
  DevBuf nkeys, nrec;
  alloc_table(a, ncap, nkeys, nrec);
  uint64_t ocap = a->cap;
  DevBuf okeys = std::move(a->keys), orec = std::move(a->rec);
  a->keys = std::move(nkeys);
  a->rec = std::move(nrec);
  a->cap = ncap;
  a->log2cap = log2_ceil(ncap);
  KG_HIP(hipMemsetAsync(a->occ.get(), 0, 8, a->s));
  Table t = table_of(a);
  k_rehash<<<stream_grid(ocap + 2, 256), 256, 0, a->s>>>(a->prog, okeys.get<uint64_t>(),
                                                          orec.get<uint64_t>(), ocap, t);
  KG_HIP(hipGetLastError());
  KG_HIP(hipStreamSynchronize(a->s));
}

static uint64_t read_occupied(kg_hash_agg *a) {
  uint64_t v = 0;
  KG_HIP(hipMemcpyAsync(&v, a->occ.get(), 8, hipMemcpyDeviceToHost, a->s));
  KG_HIP(hipStreamSynchronize(a->s));
  return v;
}

// Rows the next launch may take so that the table stays below 3/4 full
// whatever keys they hold.
static int64_t room(kg_hash_agg *a, int64_t want) {
  if (a->direct) return want;
  uint64_t limit = a->cap / 4 * 3;
  if (a->occ_bound + (uint64_t)want <= limit) return want;
  uint64_t occ = read_occupied(a);
  a->occ_bound = occ;
  if (occ > limit / 2) {
    grow(a, a->cap * 2);
    occ = read_occupied(a);
    a->occ_bound = occ;
    limit = a->cap / 4 * 3;
  }
  uint64_t r = limit - a->occ_bound;
  return want < (int64_t)r ? want : (int64_t)r;
}

// Guarantees room for n new keys (the replay path, which does not split).
static void reserve(kg_hash_agg *a, int64_t n) {
  if (a->direct) return;
  if (a->occ_bound + (uint64_t)n <= a->cap / 4 * 3) return;
  uint64_t occ = read_occupied(a);
  a->occ_bound = occ;
  while (occ + (uint64_t)n > a->cap / 4 * 3) grow(a, a->cap * 2);
}

static void launch_update(kg_hash_agg *a, const AggProg &P, int64_t lo, int64_t n) {
  if (n <= 0) return;
  AggProg Q = P;
  // offset the columns to this piece of the batch
  for (int c = 0; c < kAggInputs; c++) Q.in[c].offset += lo;
  Q.key.offset += lo;
  Table t = table_of(a);
  int64_t chunk = 16384;
  int64_t blocks = ceil_div(n, chunk);
  int64_t maxb = (int64_t)(a->ctx->num_cus > 0 ? a->ctx->num_cus : 256) * 4;
  if (blocks > maxb) {
    blocks = maxb;
    chunk = ceil_div(n, blocks);
  }
  ProfScope prof("agg_update", a->s);
  switch (a->slots) {
    case 512: k_agg_update<512><<<(unsigned)blocks, kAggBlock, a->lds_bytes, a->s>>>(Q, t, n, chunk); break;
    case 256: k_agg_update<256><<<(unsigned)blocks, kAggBlock, a->lds_bytes, a->s>>>(Q, t, n, chunk); break;
    default: k_agg_update<128><<<(unsigned)blocks, kAggBlock, a->lds_bytes, a->s>>>(Q, t, n, chunk); break;
  }
  KG_HIP(hipGetLastError());
  a->occ_bound += (uint64_t)n;
}

// Whether no running sum of the checked op can leave its bound during this
// batch, in any row order (the per-batch proof of accumulator/helpers.rs:40-75).
static bool proven(kg_hash_agg *a, CheckedSum &c, const DCol &col, int64_t n) {
  // 1: rows so far x the type's largest magnitude
  unsigned __int128 rows = (unsigned __int128)(a->rows_seen + (uint64_t)n);
  unsigned __int128 tb = c.type_bound;
  if (tb == 0 || rows <= (((unsigned __int128)0) - 1) / tb) {
    if (rows * tb < c.bound) {
      return true;
    }
  }
  // 2: the stored sums' bound plus this batch's exact sum of magnitudes
  a->scratch_abs.ensure(a->ctx, 8);
  KG_HIP(hipMemsetAsync(a->scratch_abs.get(), 0, 8, a->s));
  k_abs_sum<<<stream_grid(n, 256), 256, 0, a->s>>>(col, n, a->scratch_abs.get<double>());
  KG_HIP(hipGetLastError());
  double batch = 0.0;
  KG_HIP(hipMemcpyAsync(&batch, a->scratch_abs.get(), 8, hipMemcpyDeviceToHost, a->s));
  KG_HIP(hipStreamSynchronize(a->s));
  double total = (c.cum_abs + batch) * (1.0 + 1e-9) + 1.0;
  double bound = (double)c.bound;
  if (total < bound * (1.0 - 1e-9)) {
    c.cum_abs += batch;
    return true;
  }
  c.cum_abs += batch;
  return false;
}

static void sink(kg_hash_agg *a, const kg_column *key, const kg_column *inputs, int64_t nrows) {
  if (a->finalized) fail(KG_ERR_INTERNAL, "sink after finalize");
  if (nrows <= 0) return;
  Staged st(a->ctx, a->s);
  AggProg P = a->prog;
  if (a->has_key) {
    if (!key) fail(KG_ERR_INTERNAL, "grouping column missing");
    check_column(*key, nrows, "grouping column");
    if (key->type != a->key_type.type) fail(KG_ERR_INTERNAL, "grouping column type changed");
    P.key = st.add(*key, nrows);
  }
  for (int c = 0; c < (int)a->in_types.size(); c++) {
    check_column(inputs[c], nrows, "aggregate input");
    if (inputs[c].type != a->in_types[c].type) fail(KG_ERR_INTERNAL, "input column type changed");
    P.in[c] = st.add(inputs[c], nrows);
  }
  bool replay = false;
  for (auto &c : a->checked)
    if (!proven(a, c, P.in[c.col], nrows)) replay = true;
  if (!replay) {
    P.skip_checked = 0;
    int64_t done = 0;
    while (done < nrows) {
      int64_t n = room(a, nrows - done);
      launch_update(a, P, done, n);
      done += n;
    }
  } else {
    // rows straight to the global table (their record index kept), then the
    // checked sums in row order
    P.skip_checked = 1;
    reserve(a, nrows);
    a->scratch_slots.ensure(a->ctx, (uint64_t)nrows * 8);
    Table t = table_of(a);
    k_agg_rows_global<<<stream_grid(nrows, 256), 256, 0, a->s>>>(P, t, nrows,
                                                                  a->scratch_slots.get<uint64_t>());
    KG_HIP(hipGetLastError());
    a->occ_bound += (uint64_t)nrows;
    CheckedOps C{};
    C.n = (int)a->checked.size();
    for (int k = 0; k < C.n; k++) {
      auto &c = a->checked[k];
      C.col[k] = (uint16_t)c.col;
      C.word[k] = (uint16_t)c.word;
      C.st_word[k] = (uint16_t)c.st_word;
      C.avg[k] = c.avg;
      C.wide[k] = c.wide;
      C.bound_lo[k] = (uint64_t)c.bound;
      C.bound_hi[k] = (uint64_t)(c.bound >> 64);
    }
    ProfScope prof("agg_replay", a->s);
    k_agg_replay<<<1, 64, 0, a->s>>>(P, C, a->scratch_slots.get<uint64_t>(), nrows, t.rec);
    KG_HIP(hipGetLastError());
  }
  a->rows_seen += (uint64_t)nrows;
  // host inputs were staged into buffers freed on return: finish with them
  if (st.staged_host()) KG_HIP(hipStreamSynchronize(a->s));
}

void compile_aggs(const std::vector<kg_column> &in_types, int32_t naggs, const kg_agg_spec *aggs,
                  bool ansi, AggProg &P, OutProg &O, std::vector<CheckedSum> &checked) {
  int W = 1;  // word 0: rows
  std::vector<AggOp> ops;
  ops.push_back(AggOp{OP_COUNT_STAR, 0, 0, 0, 0, 0});
  std::vector<int> nulls_word(in_types.size(), -1);
  std::vector<uint64_t> init(1, 0);
  auto words = [&](int n, uint64_t v) {
    int w = W;
    for (int k = 0; k < n; k++) init.push_back(v);
    W += n;
    return w;
  };
  struct Key {
    int kind, col, checked, avg;
  };
  std::vector<std::pair<Key, int>> shared;
  auto find_or_add = [&](int kind, int col, bool checked, bool avg, int nwords, uint64_t iv,
                         int &st_word) {
    for (auto &kv : shared)
      if (kv.first.kind == kind && kv.first.col == col && kv.first.checked == checked &&
          kv.first.avg == avg && !checked)
        return kv.second;
    int w = words(nwords, iv);
    st_word = checked ? words(1, 0) : 0;
    ops.push_back(AggOp{(uint8_t)kind, (uint8_t)col, (uint8_t)checked, 0, (uint16_t)w,
                        (uint16_t)st_word});
    shared.push_back({Key{kind, col, checked, avg}, w});
    return w;
  };
  O.n = naggs;
  for (int i = 0; i < naggs; i++) {
    OutAgg &o = O.a[i];
    o = OutAgg{};
    o.func = aggs[i].func;
    o.col = aggs[i].input;
    o.st_word = -1;
    o.nulls_word = -1;
    if (aggs[i].func == KG_AGG_COUNT_STAR) {
      o.kind = R_COUNT_STAR;
      o.out_type = KG_T_INT64;
      continue;
    }
    int c = aggs[i].input;
    if (c < 0 || c >= (int)in_types.size()) fail(KG_ERR_INTERNAL, "aggregate input out of range");
    const kg_column &t = in_types[c];
    if (nulls_word[c] < 0) {
      nulls_word[c] = words(1, 0);
      ops.push_back(AggOp{OP_NULLS, (uint8_t)c, 0, 0, (uint16_t)nulls_word[c], 0});
    }
    o.nulls_word = nulls_word[c];
    o.in_type = t.type;
    o.in_s = t.scale;
    int st = 0;
    switch (aggs[i].func) {
      case KG_AGG_COUNT:
        o.kind = R_COUNT;
        o.out_type = KG_T_INT64;
        break;
      case KG_AGG_SUM:
      case KG_AGG_AVG: {
        bool avg = aggs[i].func == KG_AGG_AVG;
        if (is_float(t.type)) {
          o.kind = avg ? R_AVG_F : R_SUM_F;
          o.out_type = KG_T_FLOAT64;
          o.word = find_or_add(OP_ADDF64, c, false, false, 1, 0, st);
        } else if (is_integral(t.type)) {
          // avg of integers sums in double in the reference (avg_sum_type);
          // the exact 128-bit sum rounded once gives that double whenever
          // the row-order double sum is exact (|sum| < 2^53)
          bool chk = !avg && ansi;
          o.kind = avg ? R_AVG_INT : R_SUM_INT;
          o.out_type = avg ? KG_T_FLOAT64 : KG_T_INT64;
          o.word = find_or_add(OP_ADD128, c, chk, false, 2, 0, st);
          if (chk) {
            o.st_word = st;
            int tw = type_width(t.type);
            checked.push_back({c, o.word, st, false, false,
                                  (unsigned __int128)1 << 63,
                                  (unsigned __int128)1 << (8 * tw - 1)});
          }
        } else if (is_decimal(t.type)) {
          if (t.scale < 0) fail(KG_ERR_UNSUPPORTED, "negative decimal scale");
          int sum_p = t.precision + 10 > 38 ? 38 : t.precision + 10;
          bool wide = t.type == KG_T_DECIMAL128;
          o.word = find_or_add(wide ? OP_ADD192 : OP_ADD128, c, true, avg, wide ? 3 : 2, 0, st);
          o.st_word = st;
          checked.push_back({c, o.word, st, avg, wide, pow10_u128(sum_p),
                                pow10_u128(t.precision)});
          if (avg) {
            o.kind = R_AVG_DEC;
            o.out_p = t.precision + 4 > 38 ? 38 : t.precision + 4;
            o.out_s = t.scale + 4 > 38 ? 38 : t.scale + 4;
          } else {
            o.kind = R_SUM_DEC;
            o.out_p = sum_p;
            o.out_s = t.scale;
          }
          o.out_type = (t.type == KG_T_DECIMAL64 && o.out_p <= 18) ? KG_T_DECIMAL64 : KG_T_DECIMAL128;
          unsigned __int128 pw = pow10_u128(o.out_p);
          o.pow_lo = (uint64_t)pw;
          o.pow_hi = (uint64_t)(pw >> 64);
        } else {
          fail(KG_ERR_UNSUPPORTED, "sum/avg of this type");
        }
        break;
      }
      case KG_AGG_MIN:
      case KG_AGG_MAX: {
        bool mx = aggs[i].func == KG_AGG_MAX;
        o.out_type = t.type;
        o.out_p = t.precision;
        o.out_s = t.scale;
        if (is_float(t.type)) {
          o.kind = R_MINMAX_F;
          o.word = find_or_add(mx ? OP_MAX_F : OP_MIN_F, c, false, false, 1, mx ? 0 : ~0ull, st);
        } else if (is_int_like(t.type) && t.type != KG_T_UINT64) {
          o.kind = R_MINMAX_I;
          o.word = find_or_add(mx ? OP_MAX_I64 : OP_MIN_I64, c, false, false, 1,
                               mx ? 0x8000000000000000ULL : 0x7FFFFFFFFFFFFFFFULL, st);
        } else {
          fail(KG_ERR_UNSUPPORTED, "min/max of this type on the device");
        }
        break;
      }
      default: fail(KG_ERR_UNSUPPORTED, "aggregate function");
    }
  }
  // columns' ops adjacent, so a row loads each value once
  std::stable_sort(ops.begin() + 1, ops.end(),
                   [](const AggOp &x, const AggOp &y) { return x.col < y.col; });
  if ((int)ops.size() > kAggOps || W > kAggWords) fail(KG_ERR_UNSUPPORTED, "too many aggregates");
  P.nops = (int)ops.size();
  for (int k = 0; k < P.nops; k++) P.ops[k] = ops[k];
  P.W = W;
  for (int k = 0; k < W; k++) P.init[k] = init[k];
  O.W = W;
  O.ansi = ansi;
}

void init_records(kg_ctx *ctx, uint64_t *rec, uint64_t nrec, const AggProg &P, hipStream_t s) {
  DevBuf init(ctx, (uint64_t)P.W * 8);
  KG_HIP(hipMemcpyAsync(init.get(), P.init, (size_t)P.W * 8, hipMemcpyHostToDevice, s));
  k_fill_words<<<stream_grid(nrec * P.W, 256), 256, 0, s>>>(rec, nrec, P.W, init.get<uint64_t>());
  KG_HIP(hipGetLastError());
  KG_HIP(hipStreamSynchronize(s));
}

int64_t group_positions(kg_ctx *ctx, const uint64_t *rec, uint64_t ns, int W, bool global,
                        DevBuf &pos, hipStream_t s) {
  DevBuf flags(ctx, ns * 4);
  pos.reset(ctx, (ns + 1) * 8);
  k_group_flags<<<stream_grid(ns, 256), 256, 0, s>>>(rec, ns, W, global, flags.get<uint32_t>());
  KG_HIP(hipGetLastError());
  size_t tb = 0;
  KG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flags.get<uint32_t>(), pos.get<uint64_t>(),
                                          (int)ns, s));
  DevBuf temp(ctx, tb + 16);
  KG_HIP(hipcub::DeviceScan::ExclusiveSum(temp.get(), tb, flags.get<uint32_t>(),
                                          pos.get<uint64_t>(), (int)ns, s));
  uint64_t last_pos = 0;
  uint32_t last_flag = 0;
  KG_HIP(hipMemcpyAsync(&last_pos, pos.get<uint64_t>() + ns - 1, 8, hipMemcpyDeviceToHost, s));
  KG_HIP(hipMemcpyAsync(&last_flag, flags.get<uint32_t>() + ns - 1, 4, hipMemcpyDeviceToHost, s));
  KG_HIP(hipStreamSynchronize(s));
  return (int64_t)(last_pos + last_flag);
}

void launch_agg_output(const OutProg &O, const uint64_t *keys, int key_stride,
                       const uint64_t *rec, uint64_t nslots, const uint64_t *pos, void *out_key,
                       uint8_t *out_key_valid, int key_width, unsigned int *err, hipStream_t s) {
  ProfScope prof("agg_output", s);
  k_agg_output<<<stream_grid(nslots, 256), 256, 0, s>>>(O, keys, key_stride, rec, nslots, pos,
                                                         out_key, out_key_valid, key_width, err);
  KG_HIP(hipGetLastError());
}

}  // namespace kg

extern "C" {

int kg_agg_create(kg_ctx *ctx, const kg_column *key_type, int32_t ninputs,
                  const kg_column *input_types, int32_t naggs, const kg_agg_spec *aggs,
                  uint32_t flags, uint64_t expected_groups, kg_hash_agg **out) {
  return guarded([&] {
    *out = nullptr;
    if (ninputs < 0 || ninputs > kAggInputs) fail(KG_ERR_UNSUPPORTED, "at most 8 input columns");
    auto a = std::make_unique<kg_hash_agg>();
    a->ctx = ctx;
    a->ansi = (flags & KG_AGG_ANSI) != 0;
    a->has_key = key_type != nullptr;
    if (a->has_key) {
      a->key_type = *key_type;
      int t = key_type->type;
      if (!(is_int_like(t) || is_float(t)))
        fail(KG_ERR_UNSUPPORTED, "grouping key type on the device");
      a->direct = type_width(t) <= 1;  // one-byte keys and booleans
    }
    for (int c = 0; c < ninputs; c++) a->in_types.push_back(input_types[c]);
    compile_aggs(a->in_types, naggs, aggs, a->ansi, a->prog, a->out, a->checked);
    a->prog.has_key = a->has_key;
    a->prog.direct = a->direct;
    a->prog.key.type = a->has_key ? key_type->type : 0;
    a->out.key_type = a->prog.key.type;
    a->out.direct = a->direct;
    a->out.global = !a->has_key;
    a->s = make_stream(ctx->device);
    int W = a->prog.W;
    // the largest LDS table (records + keys) the device grants a workgroup:
    // 64 KiB always, more where the dynamic-LDS attribute is accepted
    a->slots = 0;
    for (int slots : {512, 256, 128}) {
      if (a->direct && slots < 512) break;
      size_t bytes = (size_t)slots * (W + 1) * 8;
      if (bytes > 96 * 1024) continue;
      hipError_t e = hipSuccess;
      if (bytes > 64 * 1024) {
        const void *fn = slots == 512 ? (const void *)k_agg_update<512>
                         : slots == 256 ? (const void *)k_agg_update<256>
                                        : (const void *)k_agg_update<128>;
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) (void)hipGetLastError();
      }
      if (e == hipSuccess) {
        a->slots = slots;
        a->lds_bytes = bytes;
        break;
      }
    }
    if (a->slots == 0) fail(KG_ERR_UNSUPPORTED, "aggregate state too wide for the LDS table");
    uint64_t eg = expected_groups < 512 ? 512 : expected_groups;
    a->cap = a->direct ? 256 : (1ull << log2_ceil(eg * 2));
    a->log2cap = log2_ceil(a->cap);
    a->occ.reset(ctx, 8);
    KG_HIP(hipMemsetAsync(a->occ.get(), 0, 8, a->s));
    alloc_table(a.get(), a->cap, a->keys, a->rec);
    *out = a.release();
  });
}

int kg_agg_sink(kg_hash_agg *a, const kg_column *key, const kg_column *inputs, int64_t nrows) {
  return guarded([&] { sink(a, key, inputs, nrows); });
}

int kg_agg_finalize(kg_hash_agg *a) {
  return guarded([&] {
    if (a->finalized) return;
    a->ngroups = group_positions(a->ctx, a->rec.get<uint64_t>(), nslots(a), a->prog.W,
                                 !a->has_key, a->pos, a->s);
    a->finalized = true;
  });
}

int kg_agg_result_type(const kg_hash_agg *a, int32_t i, kg_column *out_type) {
  return guarded([&] {
    if (i < 0 || i >= a->out.n) fail(KG_ERR_INTERNAL, "aggregate index");
    *out_type = kg_column{};
    out_type->type = a->out.a[i].out_type;
    out_type->precision = a->out.a[i].out_p;
    out_type->scale = a->out.a[i].out_s;
  });
}

int kg_agg_result(kg_hash_agg *a, kg_batch **out) {
  return guarded([&] {
    *out = nullptr;
    if (!a->finalized) fail(KG_ERR_INTERNAL, "result before finalize");
    auto b = std::make_unique<kg_batch>();
    b->ctx = a->ctx;
    int64_t n = a->ngroups;
    b->nrows = n;
    uint64_t vb = (uint64_t)ceil_div(n, 32) * 4 + 4;
    OutProg O = a->out;
    O.cap = a->cap;
    int ncols = a->out.n + (a->has_key ? 1 : 0);
    b->cols.resize(ncols);
    int kw = 8;
    void *key_vals = nullptr;
    uint8_t *key_valid = nullptr;
    int base = 0;
    if (a->has_key) {
      auto &kc = b->cols[0];
      kc.type = a->key_type.type;
      kc.precision = a->key_type.precision;
      kc.scale = a->key_type.scale;
      kw = type_width(kc.type);
      uint64_t bytes = kw == 0 ? (uint64_t)ceil_div(n, 32) * 4 + 4 : (uint64_t)n * kw + 8;
      if (kw == 0) {
        // booleans leave the table as one byte per group; repack as bits
        kw = 1;
      }
      kc.values.reset(a->ctx, bytes);
      kc.validity.reset(a->ctx, vb);
      KG_HIP(hipMemsetAsync(kc.validity.get(), 0, vb, a->s));
      key_vals = kc.values.get();
      key_valid = kc.validity.get<uint8_t>();
      base = 1;
    }
    for (int j = 0; j < a->out.n; j++) {
      auto &c = b->cols[base + j];
      OutAgg &o = O.a[j];
      c.type = o.out_type;
      c.precision = o.out_p;
      c.scale = o.out_s;
      int w = type_width(o.out_type);
      c.values.reset(a->ctx, (uint64_t)n * (w > 0 ? w : 1) + 16);
      c.validity.reset(a->ctx, vb);
      KG_HIP(hipMemsetAsync(c.validity.get(), 0, vb, a->s));
      o.values = c.values.get();
      o.validity = c.validity.get<uint8_t>();
    }
    DevBuf err(a->ctx, 4);
    KG_HIP(hipMemsetAsync(err.get(), 0, 4, a->s));
    uint64_t ns = nslots(a);
    launch_agg_output(O, a->keys.get<uint64_t>(), 1, a->rec.get<uint64_t>(), ns,
                      a->pos.get<uint64_t>(), key_vals, key_valid, kw, err.get<unsigned int>(),
                      a->s);
    unsigned int e = 0;
    KG_HIP(hipMemcpyAsync(&e, err.get(), 4, hipMemcpyDeviceToHost, a->s));
    KG_HIP(hipStreamSynchronize(a->s));
    if (e & 1u) fail(KG_ERR_OVERFLOW, "Overflow in sum (use try_sum/try_avg to tolerate overflow)");
    if (e & 2u) fail(KG_ERR_OUT_OF_RANGE, "average out of the result decimal's range");
    if (a->has_key && a->key_type.type == KG_T_BOOL) {
      // repack the byte-per-group booleans as bits
      std::vector<uint8_t> bytes((size_t)n), bits((size_t)ceil_div(n, 32) * 4 + 4, 0);
      if (n) KG_HIP(hipMemcpy(bytes.data(), b->cols[0].values.get(), (size_t)n, hipMemcpyDeviceToHost));
      for (int64_t r = 0; r < n; r++)
        if (bytes[r]) bits[r >> 3] |= (uint8_t)(1u << (r & 7));
      KG_HIP(hipMemcpy(b->cols[0].values.get(), bits.data(), bits.size() < b->cols[0].values.bytes() ? bits.size() : b->cols[0].values.bytes(), hipMemcpyHostToDevice));
    }
    *out = b.release();
  });
}

void kg_agg_destroy(kg_hash_agg *a) {
  if (!a) return;
  if (a->s) (void)hipStreamSynchronize(a->s);
  hipStream_t s = a->s;
  delete a;
  destroy_stream(s);
}

}  // extern "C"
