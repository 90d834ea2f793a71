// kg_aggops.hpp -- the aggregate state machine shared by the hash
// aggregation (kg_agg.hip) and the join fused with GROUP BY the build key
// (kg_join.hip): the per-row op program, the state-word atomics (global and
// LDS), the global open-addressing table, and the final-value program.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "kg_common.hpp"
#include "kg_device.hpp"

namespace kg {

constexpr uint64_t kEmpty = 0x8000000000000000ULL;
constexpr int kAggInputs = 32;
constexpr int kAggOps = 48;
constexpr int kAggWords = 64;
constexpr int kAggBlock = 512;
constexpr int kLdsProbe = 8;

enum OpKind : uint8_t {
  OP_COUNT_STAR = 1,  // word += 1
  OP_NULLS,           // word += !valid
  OP_ADD128,          // words[w, w+1] += i128(v)
  OP_ADD192,          // words[w, w+2] += i192(v)  (decimal128)
  OP_ADDF64,          // word (double) += v
  OP_MIN_I64,
  OP_MAX_I64,
  OP_MIN_F,  // order-preserving u64 of a double
  OP_MAX_F,
  // merge of partial states (the final phase of a two-phase aggregate)
  OP_ADDV,        // word += int64 value (counts)
  OP_COUNT_TRUE,  // word += value != 0 (a `nonempty` flag)
  OP_STICKY,      // st = 2 when the sum state is NULL and col2 is nonzero
  // SUM of an integral column without overflow checks (not ANSI): only the
  // result's low 64 bits are kept, so one wrapping word is the whole state
  OP_ADD64,  // word += v (mod 2^64)
};

struct AggOp {
  uint8_t kind;
  uint8_t col;
  uint8_t checked;  // sum whose running value is bounded (ANSI int, decimals)
  uint8_t col2;     // OP_STICKY: the `nonempty` / `count` state column
  uint16_t word;
  uint16_t st_word;  // overflow-state word of a checked sum
};

struct AggProg {
  DCol key;
  DCol in[kAggInputs];
  AggOp ops[kAggOps];
  int nops;
  int W;
  int has_key;
  int direct;    // one-byte key: slot = (uint8)key, NULL -> 256
  int skip_checked;  // ordered-replay batches: checked sums are applied later
  uint64_t init[kAggWords];
  // the spill row layout read back by SpillSrc (words per row, whether word 1
  // holds validity bits, each input column's word offset)
  int spill_rw;
  int spill_vword;
  uint8_t spill_off[kAggInputs];
  uint8_t spill_packed;  // 1 + the 32-bit column packed into word 1 (0: none)
  uint8_t spill_vshift;  // the first validity bit of word 1
};

// Op o of the program through a uniform index: the op fields then come
// from the kernel argument segment with scalar loads, not one vector load
// and wait per field and row tile.
__device__ __forceinline__ AggOp op_at(const AggProg &P, int o) {
  return P.ops[__builtin_amdgcn_readfirstlane(o)];
}

// The hashed table's slots are cut into windows of wmask + 1 slots (a power
// of two): a key's home slot is the top log2cap bits of its slot hash, and it
// probes linearly within the window of its home slot, so each window holds
// exactly the keys homed in it (one workgroup can own a window, and a window
// fits in LDS).  A window with no free slot left makes an insert fail: it
// sets the window-full flag (occupied[1]) and the host reports it.
struct Table {
  uint64_t *keys;
  uint64_t *rec;
  unsigned long long *occupied;  // [0] keys inserted with atomics, [1] window-full flag
  uint64_t mask;
  int log2cap;
  uint64_t seed;
  uint64_t cap;
  uint64_t wmask;
};

// The next slot of a probe: the slot after s within s's window.
__host__ __device__ __forceinline__ uint64_t next_in_window(uint64_t s, uint64_t wmask) {
  return (s & ~wmask) | ((s + 1) & wmask);
}

// ---------------------------------------------------------------- encodings
__device__ __forceinline__ uint64_t f64_order(double f) {
  uint64_t b = dev::norm_f64_bits(f) == 0x7FF8000000000000ULL ? 0x7FF8000000000000ULL
                                                              : (uint64_t)__double_as_longlong(f);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__host__ __device__ __forceinline__ double f64_unorder(uint64_t k) {
  uint64_t b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k;
  double d;
  memcpy(&d, &b, 8);
  return d;
}

// ---------------------------------------------------------------- atomics
template <bool LDS>
__device__ __forceinline__ unsigned long long add_u64(uint64_t *p, uint64_t v) {
  if (LDS)
    return __hip_atomic_fetch_add((unsigned long long *)p, (unsigned long long)v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WORKGROUP);
  return __hip_atomic_fetch_add((unsigned long long *)p, (unsigned long long)v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}
template <bool LDS>
__device__ __forceinline__ void add_u64_nr(uint64_t *p, uint64_t v) {
  (void)add_u64<LDS>(p, v);
}

// Exact 128-bit add of (lo, hi): the carry out of the low word is taken from
// the value the atomic returns, so the pair sums exactly under any order.
template <bool LDS>
__device__ __forceinline__ void add128(uint64_t *w, uint64_t lo, uint64_t hi) {
  uint64_t hi_add = hi;
  if (lo) {
    uint64_t old = add_u64<LDS>(w, lo);
    hi_add += (old + lo) < old ? 1 : 0;
  }
  if (hi_add) add_u64_nr<LDS>(w + 1, hi_add);
}

template <bool LDS>
__device__ __forceinline__ void add192(uint64_t *w, uint64_t lo, uint64_t mid, uint64_t hi) {
  uint64_t c0 = 0;
  if (lo) {
    uint64_t old = add_u64<LDS>(w, lo);
    c0 = (old + lo) < old ? 1 : 0;
  }
  uint64_t t = mid + c0;
  uint64_t ct = (t < mid) ? 1 : 0;  // mid = ~0 with a carry in
  uint64_t c1 = 0;
  if (t) {
    uint64_t old = add_u64<LDS>(w + 1, t);
    c1 = (old + t) < old ? 1 : 0;
  }
  uint64_t h = hi + ct + c1;
  if (h) add_u64_nr<LDS>(w + 2, h);
}

template <bool LDS>
__device__ __forceinline__ void add_f64(uint64_t *w, double v) {
  if (LDS)
    __hip_atomic_fetch_add((double *)w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else
    unsafeAtomicAdd((double *)w, v);
}

// Min/max skip the atomic when the stored value already wins: a stale read
// is an older, weaker value, so the test never drops an update.
template <bool LDS>
__device__ __forceinline__ uint64_t peek(const uint64_t *w) {
  return __hip_atomic_load(w, __ATOMIC_RELAXED,
                           LDS ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT);
}
template <bool LDS>
__device__ __forceinline__ void min_i64(uint64_t *w, int64_t v) {
  if (v >= (int64_t)peek<LDS>(w)) return;
  __hip_atomic_fetch_min((long long *)w, (long long)v, __ATOMIC_RELAXED,
                         LDS ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT);
}
template <bool LDS>
__device__ __forceinline__ void max_i64(uint64_t *w, int64_t v) {
  if (v <= (int64_t)peek<LDS>(w)) return;
  __hip_atomic_fetch_max((long long *)w, (long long)v, __ATOMIC_RELAXED,
                         LDS ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT);
}
template <bool LDS>
__device__ __forceinline__ void min_u64(uint64_t *w, uint64_t v) {
  if (v >= peek<LDS>(w)) return;
  __hip_atomic_fetch_min((unsigned long long *)w, (unsigned long long)v, __ATOMIC_RELAXED,
                         LDS ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT);
}
template <bool LDS>
__device__ __forceinline__ void max_u64(uint64_t *w, uint64_t v) {
  if (v <= peek<LDS>(w)) return;
  __hip_atomic_fetch_max((unsigned long long *)w, (unsigned long long)v, __ATOMIC_RELAXED,
                         LDS ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT);
}

// Applies every op of the program to record `r` for input row i.
template <bool LDS>
__device__ __forceinline__ void apply_row(const AggProg &P, uint64_t *r, int64_t i) {
  int last_col = -1;
  bool ok = true;
  uint64_t lo = 0, hi = 0, top = 0;
  double f = 0.0;
  for (int o = 0; o < P.nops; o++) {
    const AggOp op = op_at(P, o);
    if (op.kind == OP_COUNT_STAR) {
      add_u64_nr<LDS>(r, 1);
      continue;
    }
    if (op.col != last_col) {
      last_col = op.col;
      const DCol &c = P.in[op.col];
      ok = dev::valid(c, i);
      if (dev::is_f(c.type)) {
        f = dev::load_f64(c, i);
      } else {
        int64_t h;
        dev::load_i128(c, i, lo, h);
        hi = (uint64_t)h;
        top = (uint64_t)(h >> 63);
      }
    }
    if (op.kind == OP_NULLS) {
      if (!ok) add_u64_nr<LDS>(r + op.word, 1);
      continue;
    }
    if (op.kind == OP_STICKY) {
      const DCol &c2 = P.in[op.col2];
      if (!ok && dev::valid(c2, i) && dev::load_i64(c2, i) != 0) max_u64<LDS>(r + op.st_word, 2);
      continue;
    }
    if (!ok || (op.checked && P.skip_checked)) continue;
    uint64_t *w = r + op.word;
    switch (op.kind) {
      case OP_ADDV:
      case OP_ADD64: add_u64_nr<LDS>(w, lo); break;
      case OP_COUNT_TRUE:
        if (lo) add_u64_nr<LDS>(w, 1);
        break;
      case OP_ADD128: add128<LDS>(w, lo, hi); break;
      case OP_ADD192: add192<LDS>(w, lo, hi, top); break;
      case OP_ADDF64: add_f64<LDS>(w, f); break;
      case OP_MIN_I64: min_i64<LDS>(w, (int64_t)lo); break;
      case OP_MAX_I64: max_i64<LDS>(w, (int64_t)lo); break;
      case OP_MIN_F: min_u64<LDS>(w, f64_order(f)); break;
      case OP_MAX_F: max_u64<LDS>(w, f64_order(f)); break;
      default: break;
    }
  }
}

// ---------------------------------------------------------------- wave reduction
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ int64_t wave_min_i64(int64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    int64_t o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ int64_t wave_max_i64(int64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    int64_t o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}

// The rows of the lanes in `peers` (all of one group) reduced across the
// wave, then applied once to record `r` by the lowest peer lane.  Every lane
// of the wave must call it (lanes outside `peers` contribute neutral values).
// Integer sums are reduced as four 32-bit limbs (exact), then added with the
// carry-taking atomics; float sums add in a tree order.
template <bool LDS>
__device__ __forceinline__ void apply_wave(const AggProg &P, uint64_t *r, int64_t i,
                                           uint64_t peers) {
  const int lane = __lane_id();
  const bool me = (peers >> lane) & 1;
  const bool leader = me && (peers & ((1ull << lane) - 1)) == 0;
  int last_col = -1;
  bool ok = false;
  uint64_t lo = 0, hi = 0, top = 0;
  double f = 0.0;
  for (int o = 0; o < P.nops; o++) {
    const AggOp op = op_at(P, o);
    if (op.kind == OP_COUNT_STAR) {
      if (leader) add_u64_nr<LDS>(r, (uint64_t)__popcll(peers));
      continue;
    }
    if (op.col != last_col) {
      last_col = op.col;
      const DCol &c = P.in[op.col];
      ok = me && dev::valid(c, i);
      lo = hi = top = 0;
      f = 0.0;
      if (me) {
        if (dev::is_f(c.type)) {
          f = dev::load_f64(c, i);
        } else {
          int64_t h;
          dev::load_i128(c, i, lo, h);
          hi = (uint64_t)h;
          top = (uint64_t)(h >> 63);
        }
      }
    }
    uint64_t *w = r + op.word;
    switch (op.kind) {
      case OP_NULLS: {
        uint64_t n = (uint64_t)__popcll(__ballot(me && !ok));
        if (leader && n) add_u64_nr<LDS>(w, n);
        break;
      }
      case OP_STICKY: {
        const DCol &c2 = P.in[op.col2];
        bool hit = me && !ok && dev::valid(c2, i) && dev::load_i64(c2, i) != 0;
        uint64_t any = __ballot(hit);  // every lane takes part in the ballot
        if (leader && any) max_u64<LDS>(r + op.st_word, 2);
        break;
      }
      case OP_COUNT_TRUE: {
        uint64_t n = (uint64_t)__popcll(__ballot(ok && lo != 0));
        if (leader && n) add_u64_nr<LDS>(w, n);
        break;
      }
      case OP_ADDV:
      case OP_ADD64: {
        uint64_t s = wave_sum_u64(ok ? lo : 0);
        if (leader && s) add_u64_nr<LDS>(w, s);
        break;
      }
      case OP_ADD128:
      case OP_ADD192: {
        if (op.checked && P.skip_checked) break;
        bool v = ok;
        // limbs of the (sign-extended 192-bit) value
        uint64_t l0 = v ? (lo & 0xFFFFFFFFull) : 0, l1 = v ? (lo >> 32) : 0;
        uint64_t l2 = v ? (hi & 0xFFFFFFFFull) : 0;
        int64_t l3 = v ? (int64_t)hi >> 32 : 0;
        uint64_t s0 = wave_sum_u64(l0), s1 = wave_sum_u64(l1), s2 = wave_sum_u64(l2);
        int64_t s3 = wave_sum_i64(l3);
        if (leader) {
          // S = s0 + s1*2^32 + s2*2^64 + s3*2^96 (s3 signed) as 192 bits:
          // w0 the low word, U = the rest, sign-extended
          unsigned __int128 a = (unsigned __int128)s0 + ((unsigned __int128)s1 << 32);
          uint64_t w0 = (uint64_t)a;
          __int128 U = (__int128)(a >> 64) + (__int128)s2 + ((__int128)s3 << 32);
          uint64_t w1 = (uint64_t)U;
          uint64_t w2 = (uint64_t)(U >> 64);
          if (op.kind == OP_ADD128) add128<LDS>(w, w0, w1);
          else add192<LDS>(w, w0, w1, w2);
        }
        break;
      }
      case OP_ADDF64: {
        double s = wave_sum_f64(ok ? f : 0.0);
        uint64_t any = __ballot(ok);
        if (leader && any) add_f64<LDS>(w, s);
        break;
      }
      case OP_MIN_I64: {
        // the neutral values equal the initial state: applying them is harmless
        int64_t m = wave_min_i64(ok ? (int64_t)lo : INT64_MAX);
        if (leader) min_i64<LDS>(w, m);
        break;
      }
      case OP_MAX_I64: {
        int64_t m = wave_max_i64(ok ? (int64_t)lo : INT64_MIN);
        if (leader) max_i64<LDS>(w, m);
        break;
      }
      case OP_MIN_F: {
        uint64_t m = wave_min_u64(ok ? f64_order(f) : ~0ull);
        if (leader) min_u64<LDS>(w, m);
        break;
      }
      case OP_MAX_F: {
        uint64_t m = wave_max_u64(ok ? f64_order(f) : 0ull);
        if (leader) max_u64<LDS>(w, m);
        break;
      }
      default: break;
    }
  }
}

// Applies the ops for row i as if it occurred `m` times (a probe row that
// matches m build rows contributes m join pairs to its group).  Checked
// sums and 192-bit sums are not used with multiplicities.
__device__ __forceinline__ void apply_row_mult(const AggProg &P, uint64_t *r, int64_t i,
                                               uint64_t m) {
  int last_col = -1;
  bool ok = true;
  uint64_t lo = 0, hi = 0;
  double f = 0.0;
  for (int o = 0; o < P.nops; o++) {
    const AggOp op = op_at(P, o);
    if (op.kind == OP_COUNT_STAR) {
      add_u64_nr<false>(r, m);
      continue;
    }
    if (op.col != last_col) {
      last_col = op.col;
      const DCol &c = P.in[op.col];
      ok = dev::valid(c, i);
      if (dev::is_f(c.type)) {
        f = dev::load_f64(c, i);
      } else {
        int64_t h;
        dev::load_i128(c, i, lo, h);
        hi = (uint64_t)h;
      }
    }
    if (op.kind == OP_NULLS) {
      if (!ok) add_u64_nr<false>(r + op.word, m);
      continue;
    }
    if (!ok) continue;
    uint64_t *w = r + op.word;
    switch (op.kind) {
      case OP_ADD128: {
        // (hi:lo) * m modulo 2^128 (two's complement)
        uint64_t plo = lo * m;
        uint64_t phi = __umul64hi(lo, m) + hi * m;
        add128<false>(w, plo, phi);
        break;
      }
      case OP_ADD64: add_u64_nr<false>(w, lo * m); break;
      case OP_ADDF64: add_f64<false>(w, f * (double)m); break;
      case OP_MIN_I64: min_i64<false>(w, (int64_t)lo); break;
      case OP_MAX_I64: max_i64<false>(w, (int64_t)lo); break;
      case OP_MIN_F: min_u64<false>(w, f64_order(f)); break;
      case OP_MAX_F: max_u64<false>(w, f64_order(f)); break;
      default: break;
    }
  }
}

// Merges an LDS record into a global record.
// Merges record l into g (global atomics; with LDS = true, LDS atomics).
template <bool LDS = false>
__device__ __forceinline__ void merge_record(const AggProg &P, uint64_t *g, const uint64_t *l) {
  for (int o = 0; o < P.nops; o++) {
    const AggOp op = op_at(P, o);
    const uint64_t *s = l + op.word;
    uint64_t *d = g + op.word;
    switch (op.kind) {
      case OP_COUNT_STAR:
      case OP_NULLS:
      case OP_ADDV:
      case OP_ADD64:
      case OP_COUNT_TRUE:
        if (s[0]) add_u64_nr<LDS>(d, s[0]);
        break;
      case OP_STICKY:
        if (l[op.st_word] == 2) max_u64<LDS>(g + op.st_word, 2);
        break;
      case OP_ADD128:
        if (!(op.checked && P.skip_checked)) add128<LDS>(d, s[0], s[1]);
        break;
      case OP_ADD192:
        if (!(op.checked && P.skip_checked)) add192<LDS>(d, s[0], s[1], s[2]);
        break;
      case OP_ADDF64: add_f64<LDS>(d, __longlong_as_double((long long)s[0])); break;
      case OP_MIN_I64:
        if (s[0] != P.init[op.word]) min_i64<LDS>(d, (int64_t)s[0]);
        break;
      case OP_MAX_I64:
        if (s[0] != P.init[op.word]) max_i64<LDS>(d, (int64_t)s[0]);
        break;
      case OP_MIN_F:
        if (s[0] != P.init[op.word]) min_u64<LDS>(d, s[0]);
        break;
      case OP_MAX_F:
        if (s[0] != P.init[op.word]) max_u64<LDS>(d, s[0]);
        break;
      default: break;
    }
  }
}

__device__ __forceinline__ uint64_t slot_hash(const Table &T, uint64_t key) {
  return dev::table_hash(key, T.seed);
}

// Find-or-insert of `key` in the global table (the host keeps the load below
// 3/4).  A full window (which that load makes all but impossible) flags the
// table and returns the NULL key's slot: the host then fails the operator.
__device__ __forceinline__ uint64_t global_slot(const Table &T, uint64_t key) {
  uint64_t s = slot_hash(T, key) >> (64 - T.log2cap);
  for (uint64_t p = 0; p <= T.wmask; p++) {
    uint64_t k = __hip_atomic_load(&T.keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == key) return s;
    if (k == kEmpty) {
      unsigned long long prev = atomicCAS((unsigned long long *)&T.keys[s], kEmpty, key);
      if (prev == kEmpty) {
        atomicAdd(T.occupied, 1ull);
        return s;
      }
      if (prev == key) return s;
    }
    s = next_in_window(s, T.wmask);
  }
  atomicOr(T.occupied + 1, 1ull);
  return T.cap + 1;
}

// Key of row i as the 64-bit table key, and its slot class:
// 0 = ordinary, 1 = the INT64_MIN key, 2 = NULL.
__device__ __forceinline__ int row_key(const AggProg &P, int64_t i, uint64_t &key) {
  if (!P.has_key) {
    key = 0;
    return 0;
  }
  if (!dev::valid(P.key, i)) return 2;
  const DCol &c = P.key;
  if (dev::is_f(c.type)) {
    key = c.type == KG_T_FLOAT64 ? dev::norm_f64_bits(dev::load_f64(c, i))
                                 : (uint64_t)dev::norm_f32_bits((float)dev::load_f64(c, i));
  } else {
    key = (uint64_t)dev::load_i64(c, i);
  }
  return key == kEmpty ? 1 : 0;
}

__device__ __forceinline__ uint64_t *global_record(const AggProg &P, const Table &T, int cls,
                                                   uint64_t key) {
  uint64_t s;
  if (P.direct) {
    s = cls == 2 ? 256 : (key & 0xFF);
  } else if (!P.has_key) {
    s = 0;
  } else if (cls == 0) {
    s = global_slot(T, key);
  } else {
    s = T.cap + (cls == 2 ? 1 : 0);
  }
  return T.rec + s * (uint64_t)P.W;
}

// ---------------------------------------------------------------- 128-bit helpers
struct U128 {
  uint64_t lo, hi;
};
__host__ __device__ __forceinline__ bool neg128(uint64_t lo, uint64_t hi) { return (int64_t)hi < 0; }
__host__ __device__ __forceinline__ void negate128(uint64_t &lo, uint64_t &hi) {
  lo = ~lo + 1;
  hi = ~hi + (lo == 0 ? 1 : 0);
}

// The ordered replay of checked sums: one lane walks the batch's rows in
// order and applies each checked add with Spark's per-row semantics
// (DecFold::sum / DecFold::avg, accumulator/helpers.rs:168-287; sum_ints'
// ANSI check, accumulator/dense.rs:55-125).  Bounds are 128-bit magnitudes.
struct CheckedOps {
  int n;
  uint16_t col[16];
  uint16_t word[16];
  uint16_t st_word[16];
  uint8_t avg[16];    // avg: an overflow keeps the sum
  uint8_t wide[16];   // 192-bit accumulator
  uint64_t bound_lo[16], bound_hi[16];  // |running| must stay below
};

__device__ __forceinline__ bool lt_mag(uint64_t lo, uint64_t hi, uint64_t blo, uint64_t bhi) {
  return hi < bhi || (hi == bhi && lo < blo);
}

// ---------------------------------------------------------------- output
struct OutAgg {
  int func;
  int kind;     // result kind (below)
  int col;      // input column
  int word;     // value word
  int st_word;  // checked-sum state word, -1 if none
  int nulls_word;
  int valid_word;  // >= 0: the count that decides validity (merged states)
  int cnt_word;    // >= 0: the row count an average divides by (merged states)
  int out_type, out_p, out_s, in_s, in_type;
  int wide;  // the sum accumulator has three words (192 bits)
  void *values;
  uint8_t *validity;  // LSB bitmap (32-bit words, atomically set)
  uint64_t pow_lo, pow_hi;  // 10^out_p (decimal result bound)
  // partial output: the state columns (state_fields, aggregate/mod.rs:747-775)
  int st_type, st_p, st_s;       // type of the first state column
  int st2;                       // second state column: 0 none, 1 nonempty (bool), 2 count
  void *values2;
  uint8_t *validity2;
};

enum OutKind {
  R_COUNT_STAR,
  R_COUNT,
  R_SUM_INT,
  R_SUM_F,
  R_SUM_DEC,
  R_MINMAX_I,
  R_MINMAX_F,
  R_AVG_INT,
  R_AVG_F,
  R_AVG_DEC
};

struct OutProg {
  OutAgg a[kAggOps];
  int n;
  int W;
  int ansi;
  int partial;  // emit partial states instead of final values
  int key_type;
  uint64_t cap;
  int direct;
  int global;
};

__device__ __forceinline__ void set_bit(uint8_t *bm, int64_t i) {
  atomicOr((unsigned int *)(bm + ((i >> 5) << 2)), 1u << (i & 31));
}

// u128 x u64 -> u128 with overflow flag
__device__ __forceinline__ bool mul128(uint64_t &lo, uint64_t &hi, uint64_t m) {
  uint64_t plo = lo * m, pmid = __umul64hi(lo, m);
  uint64_t hlo = hi * m, hhi = __umul64hi(hi, m);
  uint64_t nhi = pmid + hlo;
  bool ovf = hhi != 0 || nhi < pmid;
  lo = plo;
  hi = nhi;
  return ovf;
}

// (hi:lo) / d and remainder, d > 0 (shift-subtract over 128 bits).
__device__ __forceinline__ void divmod128(uint64_t lo, uint64_t hi, uint64_t d, uint64_t &qlo,
                                          uint64_t &qhi, uint64_t &rem) {
  qhi = hi / d;
  uint64_t r = hi % d;
  qlo = 0;
  for (int b = 63; b >= 0; b--) {
    uint64_t top = r >> 63;
    r = (r << 1) | ((lo >> b) & 1);
    if (top || r >= d) {
      r -= d;
      qlo |= 1ull << b;
    }
  }
  rem = r;
}

__device__ __forceinline__ void store_i128(const OutAgg &a, int64_t o, uint64_t lo, uint64_t hi) {
  if (a.out_type == KG_T_DECIMAL128) {
    uint64_t *v = static_cast<uint64_t *>(a.values) + 2 * o;
    v[0] = lo;
    v[1] = hi;
  } else {
    static_cast<uint64_t *>(a.values)[o] = lo;
  }
}

// A sum whose running value is bounded (ANSI integer sums, decimal sums and
// averages): checked by a per-batch proof or replayed in row order.
struct CheckedSum {
  int col, word, st_word;
  bool avg, wide;
  unsigned __int128 bound;       // |running sum| < bound
  unsigned __int128 type_bound;  // |value| <= type_bound
  double cum_abs = 0.0;          // upper bound of every stored |sum|
};

// Whether no running sum of the checked op can leave its bound during a
// batch of n rows of col, each counted up to `mult` times, in any row order
// (the per-batch proof of accumulator/helpers.rs:40-75): rows_before + n
// rows at the type's largest magnitude, or the stored sums' bound (c.cum_abs)
// plus the batch's exact sum of magnitudes.  batch_abs: what the batch adds
// to c.cum_abs.  scratch: 8 bytes of device memory (grown if smaller).
bool checked_batch_proven(kg_ctx *ctx, hipStream_t s, DevBuf &scratch, const CheckedSum &c,
                          const DCol &col, int64_t n, uint64_t rows_before, uint64_t mult,
                          double &batch_abs);

// Builds the update program (prog) and the final-value program (out) of an
// aggregate list over input columns of the given types.
// merge = the inputs are the partial-state columns of a first phase (in
// state_fields order), not raw rows; in_types are the aggregates' original
// argument types either way.  Returns the number of input columns a sink
// call takes.
int compile_aggs(const std::vector<kg_column> &in_types, int32_t naggs, const kg_agg_spec *aggs,
                 bool ansi, bool merge, bool partial, AggProg &prog, OutProg &out,
                 std::vector<CheckedSum> &checked);

// Fills nrec records with the program's initial state words.
void init_records(kg_ctx *ctx, uint64_t *rec, uint64_t nrec, const AggProg &P, hipStream_t s);

// The occupied records (rows > 0, or slot 0 of a global aggregate) in slot
// order: list[0..n) holds their slot indexes.  Returns n, the group count.
// A hashed table (keys given) flags its slots [0, nslots - 2) by their keys
// and the two fixed slots by their rows.
int64_t group_positions(kg_ctx *ctx, const uint64_t *rec, uint64_t nslots, int W, bool global,
                        DevBuf &list, hipStream_t s, const uint64_t *keys = nullptr);

// Writes the final values of the n records named by list (kernel k_agg_output).
void launch_agg_output(const OutProg &O, const uint64_t *keys, int key_stride,
                       const uint64_t *rec, const uint64_t *list, int64_t n, void *out_key,
                       uint8_t *out_key_valid, int key_width, unsigned int *err, hipStream_t s);

}  // namespace kg
