// kg_join.hip -- hash join: the build Sink (join/build.rs:1225-1695) and the
// probe Operator (join/probe.rs:784-948) of HashJoin (join/mod.rs:1014-1211),
// inner / left outer / left semi / left anti on `=` or `<=>` keys with Spark's
// NULL semantics (join/mod.rs:21-60), plus the join fused with GROUP BY the
// build key.
//
// Layout
//   build store     the build batches' columns, appended in HBM as they are
//                   sunk and made contiguous at finalize (the row arena).
//   slots[cap]      16-byte open-addressing slots {key, head, count}: the
//                   packed key (one integer-like column, or two whose widths
//                   sum to 8 bytes), the newest build row with that key, and
//                   the key's row count.  Slot cap holds the key equal to the
//                   EMPTY sentinel, slot cap+1 the NULL of a `<=>` key.
//   next[rows]      chains of duplicate keys, newest first (rows/index.rs:237-279).
// A probe is one 16-byte read for a unique key (key, head and count in one
// slot).  Probe output is sized by a count pass and an exclusive scan, then
// the pairs are written and the output columns gathered.
#include <hip/hip_runtime.h>


#include "kg_aggops.hpp"
#include "kg_scan.hpp"
#include "kg_aggtile.hpp"
#include "kg_gather.hpp"

namespace kg {

constexpr int kJoinCols = 32;
constexpr uint32_t kNoRow = 0xFFFFFFFFu;

struct Slot {
  uint64_t key;
  uint32_t head;
  uint32_t count;
};

struct KeySpec {
  DCol c[2];
  int n;
  int shift;  // bits of the second column (two-column keys)
  int null_safe;
};

struct JTable {
  Slot *slots;
  uint32_t *next;
  uint64_t mask;
  int log2cap;
  uint64_t cap;
  uint64_t seed;
  // direct index of a tiny build (join/direct.rs): slot of key dmin + d at
  // dslot[d] (-1: none), d < dspan; null when the build does not qualify
  const int32_t *dslot;
  int64_t dmin;
  uint64_t dspan;
};

// The residual filter as kernels evaluate it: the postfix program with each
// column operand re-indexed into `cols` (side 0 probe, 1 build).
constexpr int kFilterOps = 32, kFilterCols = 8, kFilterStack = 16;
struct FilterProg {
  int n;  // ops; 0: no filter
  int ncols;
  kg_expr op[kFilterOps];
  DCol cols[kFilterCols];
  int side[kFilterCols];
};

// Whether the pair (probe row prow, build row brow) passes the filter: the
// program yields TRUE (Spark's three-valued logic; NULL does not pass).
__device__ __noinline__ bool filter_pass(const FilterProg *__restrict__ Fp, int64_t prow,
                                         int64_t brow) {
  const FilterProg &F = *Fp;
  int64_t v[kFilterStack];
  uint32_t nul = 0;  // bit k: stack entry k is NULL
  int sp = 0;
  for (int k = 0; k < F.n; k++) {
    const kg_expr e = F.op[k];
    if (e.op == KG_X_PROBE || e.op == KG_X_BUILD) {
      const DCol &c = F.cols[e.arg];
      const int64_t r = F.side[e.arg] ? brow : prow;
      const bool ok = dev::valid(c, r);
      v[sp] = ok ? dev::load_i64(c, r) : 0;
      nul = ok ? nul & ~(1u << sp) : nul | (1u << sp);
      sp++;
      continue;
    }
    if (e.op == KG_X_CONST) {
      v[sp] = e.imm;
      nul &= ~(1u << sp);
      sp++;
      continue;
    }
    if (e.op == KG_X_NOT) {
      v[sp - 1] = !v[sp - 1];
      continue;
    }
    if (e.op == KG_X_IS_NULL) {
      v[sp - 1] = (nul >> (sp - 1)) & 1;
      nul &= ~(1u << (sp - 1));
      continue;
    }
    sp--;
    const int64_t a = v[sp - 1], b = v[sp];
    const bool an = (nul >> (sp - 1)) & 1, bn = (nul >> sp) & 1;
    int64_t r = 0;
    bool rn = false;
    if (e.op == KG_X_AND) {  // FALSE wins over NULL
      if ((!an && !a) || (!bn && !b)) r = 0;
      else { r = 1; rn = an || bn; }
    } else if (e.op == KG_X_OR) {  // TRUE wins over NULL
      if ((!an && a) || (!bn && b)) r = 1;
      else { r = 0; rn = an || bn; }
    } else if (an || bn) {
      rn = true;
    } else {
      const uint64_t ua = (uint64_t)a, ub = (uint64_t)b;
      switch (e.op) {
        case KG_X_ADD: r = (int64_t)(ua + ub); break;
        case KG_X_SUB: r = (int64_t)(ua - ub); break;
        case KG_X_MUL: r = (int64_t)(ua * ub); break;
        case KG_X_MOD:
          if (b == 0) rn = true;
          else r = b == -1 ? 0 : a % b;
          break;
        case KG_X_EQ: r = a == b; break;
        case KG_X_NE: r = a != b; break;
        case KG_X_LT: r = a < b; break;
        case KG_X_LE: r = a <= b; break;
        case KG_X_GT: r = a > b; break;
        default: r = a >= b; break;  // KG_X_GE
      }
    }
    v[sp - 1] = r;
    nul = rn ? nul | (1u << (sp - 1)) : nul & ~(1u << (sp - 1));
  }
  return !(nul & 1u) && v[0] != 0;
}

// Join types that return pairs, and those that mark the build rows they match.
__device__ __host__ __forceinline__ bool emits_pairs(int jt) {
  return jt == KG_JOIN_INNER || jt == KG_JOIN_LEFT_OUTER || jt == KG_JOIN_RIGHT_OUTER ||
         jt == KG_JOIN_FULL_OUTER;
}
__device__ __host__ __forceinline__ bool marks_build(int jt) { return jt >= KG_JOIN_RIGHT_SEMI; }

// Packed key of row i.  Returns 0 (ordinary), 1 (the EMPTY value), 2 (a
// `<=>` NULL) or 3 (an `=` NULL: matches nothing).
__device__ __forceinline__ int pack_key(const KeySpec &K, int64_t i, uint64_t &key) {
  uint64_t parts[2];
  for (int c = 0; c < K.n; c++) {
    const DCol &d = K.c[c];
    if (!dev::valid(d, i)) {
      key = 0;
      return K.null_safe ? 2 : 3;
    }
    if (d.type == KG_T_FLOAT64) parts[c] = dev::norm_f64_bits(dev::load_f64(d, i));
    else if (d.type == KG_T_FLOAT32) parts[c] = dev::norm_f32_bits((float)dev::load_f64(d, i));
    else parts[c] = (uint64_t)dev::load_i64(d, i);
  }
  if (K.n == 1) {
    key = parts[0];
  } else {
    uint64_t m = K.shift >= 64 ? ~0ull : ((1ull << K.shift) - 1);
    key = (parts[0] << K.shift) | (parts[1] & m);
  }
  return key == kEmpty ? 1 : 0;
}

__device__ __forceinline__ uint64_t jslot_start(const JTable &T, uint64_t key) {
  return dev::table_hash(key, T.seed) >> (64 - T.log2cap);
}

__global__ void k_join_init(Slot *slots, uint64_t n) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < n;
       s += (uint64_t)gridDim.x * blockDim.x)
    slots[s] = Slot{kEmpty, kNoRow, 0};
}

__global__ void k_join_build(KeySpec K, int64_t n, JTable T, unsigned long long *null_rows) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t key;
    int cls = pack_key(K, i, key);
    if (cls == 3) {
      T.next[i] = kNoRow;
      atomicAdd(null_rows, 1ull);  // a NOT IN build with a NULL keeps no probe row
      continue;  // an `=` key with a NULL is never indexed
    }
    uint64_t s;
    if (cls == 0) {
      s = jslot_start(T, key);
      while (true) {
        uint64_t k = __hip_atomic_load(&T.slots[s].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == key) break;
        if (k == kEmpty) {
          unsigned long long prev =
              atomicCAS((unsigned long long *)&T.slots[s].key, kEmpty, key);
          if (prev == kEmpty || prev == key) break;
        }
        s = (s + 1) & T.mask;
      }
    } else {
      s = T.cap + (cls == 2 ? 1 : 0);
    }
    uint32_t old = atomicExch(&T.slots[s].head, (uint32_t)i);
    T.next[i] = old;
    atomicAdd(&T.slots[s].count, 1u);
  }
}

// Direct-index lookup (join/direct.rs:199-230: key - min, one load): the
// slot of an integer key, or -1 outside the span or for an absent key.
__device__ __forceinline__ int64_t dfind(const JTable &T, uint64_t key) {
  const uint64_t d = key - (uint64_t)T.dmin;  // wraps outside the span
  return d < T.dspan ? (int64_t)T.dslot[d] : -1;
}

// Read-only lookup: slot index, or -1.
__device__ __forceinline__ int64_t jfind(const JTable &T, int cls, uint64_t key) {
  if (cls == 3) return -1;
  if (T.dslot) return dfind(T, key);  // no `<=>` NULL with a direct index
  if (cls != 0) return (int64_t)(T.cap + (cls == 2 ? 1 : 0));
  uint64_t s = jslot_start(T, key);
  while (true) {
    uint64_t k = T.slots[s].key;
    if (k == key) return (int64_t)s;
    if (k == kEmpty) return -1;
    s = (s + 1) & T.mask;
  }
}

// Pass 1: matches and output rows of each probe row.  kJCU rows a lane,
// in stages whose loads are all in flight together: the selection indexes,
// the keys (FAST: one 8-byte integer key column, validity bytes kept raw),
// then the first slot (or direct-index entry) of each key; a probe chain
// past the first slot is followed row by row (rare at the table's load).
constexpr int kJCU = 4;

// With a residual filter, or for the joins that return build rows by their
// matches (bmatch: one byte a build row, set by plain stores), a key's chain
// is walked: m becomes the number of pairs that pass, and head is cleared
// when none does.  naaj_all: NOT IN over an empty build keeps every row.
template <bool FAST, bool WALK>
__global__ void k_join_count(KeySpec K, int64_t n, const int32_t *sel, JTable T, int join_type,
                             const FilterProg *F, uint8_t *bmatch, int naaj_all, uint32_t *head,
                             uint32_t *nout) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * kJCU;
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x * kJCU + threadIdx.x; i0 < n; i0 += stride) {
    int64_t row[kJCU];
#pragma unroll
    for (int r = 0; r < kJCU; r++) {
      const int64_t i = i0 + (int64_t)r * blockDim.x, ic = i < n ? i : n - 1;
      row[r] = sel ? (int64_t)sel[ic] : ic;
    }
    uint64_t key[kJCU];
    int cls[kJCU];
    if (FAST) {
      const DCol &d = K.c[0];
      uint32_t kb[kJCU];
#pragma unroll
      for (int r = 0; r < kJCU; r++) {
        key[r] = dev::load_w64(d, row[r]);
        kb[r] = d.validity ? d.validity[(row[r] + d.offset) >> 3] : 0xFFu;
      }
#pragma unroll
      for (int r = 0; r < kJCU; r++) {
        const bool ok = (kb[r] >> ((row[r] + d.offset) & 7)) & 1;
        cls[r] = !ok ? (K.null_safe ? 2 : 3) : key[r] == kEmpty ? 1 : 0;
        if (!ok) key[r] = 0;
      }
    } else {
#pragma unroll
      for (int r = 0; r < kJCU; r++) cls[r] = pack_key(K, row[r], key[r]);
    }
    // first slot of each key (its fixed slot for classes 1 and 2, slot 0 as
    // a harmless stand-in where the row finds nothing)
    int64_t s[kJCU];
#pragma unroll
    for (int r = 0; r < kJCU; r++) {
      if (T.dslot) {
        const uint64_t dd = key[r] - (uint64_t)T.dmin;
        s[r] = cls[r] != 3 && dd < T.dspan ? (int64_t)T.dslot[dd] : -1;
      } else {
        s[r] = cls[r] == 3 ? -1
               : cls[r] != 0 ? (int64_t)(T.cap + (cls[r] == 2 ? 1 : 0))
                             : (int64_t)jslot_start(T, key[r]);
      }
    }
    uint64_t sk[kJCU];
    uint32_t sh[kJCU], sc[kJCU];
#pragma unroll
    for (int r = 0; r < kJCU; r++) {
      const Slot t = T.slots[s[r] >= 0 ? s[r] : 0];
      sk[r] = t.key;
      sh[r] = t.head;
      sc[r] = t.count;
    }
#pragma unroll
    for (int r = 0; r < kJCU; r++) {
      const int64_t i = i0 + (int64_t)r * blockDim.x;
      uint32_t h = kNoRow, m = 0;
      if (s[r] >= 0) {
        if (!T.dslot && cls[r] == 0) {
          // linear probing from the home slot
          uint64_t at = (uint64_t)s[r];
          while (sk[r] != key[r] && sk[r] != kEmpty) {
            at = (at + 1) & T.mask;
            const Slot t = T.slots[at];
            sk[r] = t.key;
            sh[r] = t.head;
            sc[r] = t.count;
          }
          if (sk[r] == key[r]) {
            h = sh[r];
            m = sc[r];
          }
        } else {
          h = sh[r];
          m = sc[r];
        }
      }
      if (WALK && h != kNoRow && i < n) {
        uint32_t mp = 0;
        uint32_t b = h;
        for (uint32_t c = 0; c < m && b != kNoRow; c++, b = T.next[b]) {
          if (F && !filter_pass(F, row[r], b)) continue;
          mp++;
          if (bmatch) bmatch[b] = 1;
          else if (!emits_pairs(join_type)) break;  // one match decides
        }
        m = mp;
        if (!mp) h = kNoRow;
      }
      uint32_t o;
      switch (join_type) {
        case KG_JOIN_INNER:
        case KG_JOIN_RIGHT_OUTER: o = m; break;
        case KG_JOIN_LEFT_OUTER:
        case KG_JOIN_FULL_OUTER: o = m ? m : 1; break;
        case KG_JOIN_LEFT_SEMI: o = m ? 1 : 0; break;
        case KG_JOIN_LEFT_ANTI: o = m ? 0 : 1; break;
        case KG_JOIN_NULL_AWARE_ANTI: o = naaj_all ? 1 : (cls[r] != 3 && !m) ? 1 : 0; break;
        case KG_JOIN_EXISTENCE: o = 1; break;
        default: o = 0; break;  // right semi / right anti: build rows only
      }
      if (i < n) {
        head[i] = h;
        nout[i] = o;
      }
    }
  }
}

// Pass 2: the (probe row, build row) pairs of output rows [w0, w1) of the
// batch, written at their index - w0; probe rows [ia, ib) hold them.  Build
// row -1 pads; an existence join's second column is its `exists` (1/0).
template <bool FILT>
__global__ void k_join_pairs(int64_t ia, int64_t ib, const int32_t *sel, const uint32_t *head,
                             const uint32_t *nout, const uint64_t *off, const uint32_t *next,
                             int join_type, const FilterProg *F, uint64_t w0, uint64_t w1,
                             int32_t *pidx, int32_t *bidx) {
  for (int64_t i = ia + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ib;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t o = nout[i];
    if (!o) continue;
    const uint64_t p = off[i];
    if (p >= w1 || p + o <= w0) continue;
    const int32_t row = sel ? sel[i] : (int32_t)i;
    uint32_t r = head[i];
    if (!emits_pairs(join_type)) {  // one output row, inside the window
      pidx[p - w0] = row;
      if (bidx) bidx[p - w0] = join_type == KG_JOIN_EXISTENCE ? (r != kNoRow) : -1;
      continue;
    }
    if (r == kNoRow) {  // an outer join's unmatched probe row
      pidx[p - w0] = row;
      bidx[p - w0] = -1;
      continue;
    }
    for (uint32_t k = 0; r != kNoRow && k < o; r = next[r]) {
      if (FILT && !filter_pass(F, row, r)) continue;
      const uint64_t q = p + k++;
      if (q >= w1) break;
      if (q >= w0) {
        pidx[q - w0] = row;
        bidx[q - w0] = (int32_t)r;
      }
    }
  }
}

// The probe rows holding output rows [w0, w1): out[0] = the first row whose
// outputs end past w0, out[1] = the first row whose outputs start at or past
// w1 (off is non-decreasing).
__global__ void k_join_window(const uint64_t *off, const uint32_t *nout, int64_t n, uint64_t w0,
                              uint64_t w1, int64_t *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t m = (lo + hi) / 2;
    if (off[m] + nout[m] > w0) hi = m;
    else lo = m + 1;
  }
  out[0] = lo;
  hi = n;
  while (lo < hi) {
    const int64_t m = (lo + hi) / 2;
    if (off[m] >= w1) hi = m;
    else lo = m + 1;
  }
  out[1] = lo;
}

// The tail of a join that returns build rows by their matches: flags of the
// rows it returns, then their ids in build order.
__global__ void k_tail_flags(const uint8_t *bmatch, int64_t n, int want_matched, uint32_t *flags) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x)
    flags[r] = (bmatch[r] != 0) == (want_matched != 0);
}
__global__ void k_tail_list(const uint32_t *flags, const uint64_t *pos, int64_t n, int32_t *list) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x)
    if (flags[r]) list[pos[r]] = (int32_t)r;
}

struct GatherCols {
  DCol src[kJoinCols];
  int from_build[kJoinCols];  // 0 probe column, 1 build column, 2 `exists` (from bidx)
  void *dst[kJoinCols];
  uint32_t *dst_valid[kJoinCols];  // NULL when the output column has no NULLs
  int n;
};

__device__ __forceinline__ void copy_elem(const DCol &in, int64_t j, void *out, int64_t pos) {
  j += in.offset;
  switch (in.width) {
    case 1: static_cast<uint8_t *>(out)[pos] = static_cast<const uint8_t *>(in.values)[j]; break;
    case 2: static_cast<uint16_t *>(out)[pos] = static_cast<const uint16_t *>(in.values)[j]; break;
    case 4: static_cast<uint32_t *>(out)[pos] = static_cast<const uint32_t *>(in.values)[j]; break;
    case 8: static_cast<uint64_t *>(out)[pos] = static_cast<const uint64_t *>(in.values)[j]; break;
    case 16: {
      const uint64_t *s = static_cast<const uint64_t *>(in.values) + 2 * j;
      uint64_t *d = static_cast<uint64_t *>(out) + 2 * pos;
      d[0] = s[0];
      d[1] = s[1];
      break;
    }
    default:
      if ((static_cast<const uint8_t *>(in.values)[j >> 3] >> (j & 7)) & 1)
        atomicOr(static_cast<uint32_t *>(out) + (pos >> 5), 1u << (pos & 31));
  }
}

// A wave takes 64 consecutive output rows (blocks of whole waves, wave-
// aligned strides), so each nullable column's validity bits of those rows
// are two whole 32-bit words: a ballot, stored by lanes 0 and 32 (no atomic
// per row).  Padded rows (build row -1) get a 0 bit: NULL.
__global__ void k_join_gather(GatherCols G, int64_t nout, const int32_t *pidx,
                              const int32_t *bidx) {
  const int lane = threadIdx.x & 63;
  for (int64_t b0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); b0 < nout;
       b0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = b0 + lane;
    const bool in = o < nout;
    const int32_t pi = in && pidx ? pidx[o] : -1;
    const int32_t bi = in && bidx ? bidx[o] : -1;
    for (int c = 0; c < G.n; c++) {
      if (G.from_build[c] == 2) {  // the existence join's BOOL column, never NULL
        const uint64_t m = __ballot(in && bi == 1);
        uint32_t *w = static_cast<uint32_t *>(G.dst[c]);
        if (lane == 0) w[b0 >> 5] = (uint32_t)m;
        if (lane == 32 && b0 + 32 < nout) w[(b0 >> 5) + 1] = (uint32_t)(m >> 32);
        continue;
      }
      const int32_t r = G.from_build[c] ? bi : pi;
      const DCol &s = G.src[c];
      if (r >= 0) copy_elem(s, r, G.dst[c], o);
      if (G.dst_valid[c]) {
        const uint64_t m = __ballot(r >= 0 && dev::valid(s, r));
        if (lane == 0) G.dst_valid[c][b0 >> 5] = (uint32_t)m;
        if (lane == 32 && b0 + 32 < nout) G.dst_valid[c][(b0 >> 5) + 1] = (uint32_t)(m >> 32);
      }
    }
  }
}

// The most build rows one key holds (slot counts), into *out.
__global__ void k_slot_max_count(const Slot *slots, uint64_t n, unsigned int *out) {
  unsigned int m = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    m = slots[i].count > m ? slots[i].count : m;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned int o = __shfl_xor(m, off);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

// The join fused with GROUP BY the build key: each probe row's aggregate
// ops go to its build key's record, once per matching build row.
__global__ void k_joinagg_probe(KeySpec K, int64_t n, JTable T, AggProg P, uint64_t *rec) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t key;
    int cls = pack_key(K, i, key);
    int64_t s = jfind(T, cls, key);
    if (s < 0) continue;
    uint32_t m = T.slots[s].count;
    if (!m) continue;
    apply_row_mult(P, rec + (uint64_t)s * P.W, i, m);
  }
}

// min/max of the keys in the table (as signed 64-bit values; the EMPTY
// value's fixed slot counts when it holds rows): out = {min, max}, preset
// to {INT64_MAX, INT64_MIN}.
__global__ void k_slot_minmax(JTable T, long long *out) {
  long long lo = LLONG_MAX, hi = LLONG_MIN;
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s <= T.cap;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = s < T.cap ? T.slots[s].key : kEmpty;
    if (s < T.cap ? key == kEmpty : T.slots[s].count == 0) continue;
    const long long v = (long long)key;
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int d = 32; d > 0; d >>= 1) {
    const long long a = __shfl_xor(lo, d, 64), b = __shfl_xor(hi, d, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & 63) == 0) {
    if (lo != LLONG_MAX) atomicMin(&out[0], lo);
    if (hi != LLONG_MIN) atomicMax(&out[1], hi);
  }
}

// The direct index: every occupied slot (and the EMPTY value's fixed slot)
// writes its slot number at key - min.  Keys in slots are distinct.
__global__ void k_join_direct(JTable T, int32_t *dslot) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s <= T.cap;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = s < T.cap ? T.slots[s].key : kEmpty;
    if (s < T.cap && key == kEmpty) continue;
    if (s == T.cap && T.slots[s].count == 0) continue;
    const uint64_t d = key - (uint64_t)T.dmin;
    if (d < T.dspan) dslot[d] = (int32_t)s;
  }
}

// The fused aggregate over a direct index whose records fit in LDS: each
// workgroup keeps a private record per key of the span (LDS atomics), reads
// its share of the probe rows once, and adds its records into the global
// ones at the end (global atomics: one set per key a workgroup touched).
constexpr int kJDBlock = 512;
constexpr int kJDTile = 4;

// FAST: one 8-byte integer key column; its rows' key and validity loads are
// issued together (clamped row index), as are the input loads (row 0 stands
// in for rows past n).
template <bool FAST>
__global__ void __launch_bounds__(kJDBlock) k_joinagg_direct(KeySpec K, int64_t n, JTable T,
                                                             AggProg P, uint64_t *rec) {
  extern __shared__ uint64_t smem[];
  const int W = P.W, WL = W | 1;
  const int span = (int)T.dspan;
  uint64_t *s_rec = smem;                                                // span x WL
  uint32_t *s_cnt = reinterpret_cast<uint32_t *>(s_rec + (size_t)span * WL);  // span
  fill_lds_records(P, s_rec, span, WL);
  for (int d = threadIdx.x; d < span; d += kJDBlock) {
    const int32_t sl = T.dslot[d];
    s_cnt[d] = sl >= 0 ? T.slots[sl].count : 0u;
  }
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kJDBlock * kJDTile;
  for (int64_t t0 = (int64_t)blockIdx.x * kJDBlock * kJDTile; t0 < n; t0 += stride) {
    int ls[kJDTile];
    uint64_t m[kJDTile];
    ColSrc<kJDTile> src;
    src.safe = 0;  // n >= 1
    uint64_t key[kJDTile];
    int cls[kJDTile];
    if (FAST) {
      const DCol &kc = K.c[0];
      uint32_t kb[kJDTile];
#pragma unroll
      for (int r = 0; r < kJDTile; r++) {
        const int64_t i = t0 + (int64_t)r * kJDBlock + threadIdx.x, ic = i < n ? i : n - 1;
        key[r] = dev::load_w64(kc, ic);
        kb[r] = kc.validity ? kc.validity[(ic + kc.offset) >> 3] : 0xFFu;
      }
#pragma unroll
      for (int r = 0; r < kJDTile; r++) {
        const int64_t i = t0 + (int64_t)r * kJDBlock + threadIdx.x, ic = i < n ? i : n - 1;
        // a NULL key matches nothing (no `<=>` join has a direct index)
        cls[r] = i < n && ((kb[r] >> ((ic + kc.offset) & 7)) & 1) ? 0 : 3;
      }
    } else {
#pragma unroll
      for (int r = 0; r < kJDTile; r++) {
        const int64_t i = t0 + (int64_t)r * kJDBlock + threadIdx.x;
        cls[r] = i < n ? pack_key(K, i, key[r]) : 3;
      }
    }
#pragma unroll
    for (int r = 0; r < kJDTile; r++) {
      const int64_t i = t0 + (int64_t)r * kJDBlock + threadIdx.x;
      src.idx[r] = i < n ? i : 0;
      ls[r] = -1;
      m[r] = 0;
      if (cls[r] != 3) {
        const uint64_t d = key[r] - (uint64_t)T.dmin;
        if (d < T.dspan && s_cnt[d]) {
          ls[r] = (int)d;
          m[r] = s_cnt[d];
        }
      }
    }
    tile_apply_mult<kJDTile>(P, s_rec, WL, ls, m, src);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < span; d += kJDBlock) {
    const uint64_t *l = s_rec + (size_t)d * WL;
    if (l[0] == 0) continue;  // no row of this workgroup reached the key
    merge_record(P, rec + (uint64_t)T.dslot[d] * W, l);
  }
}

// ---------------------------------------------------------------- partitioned
// The join fused with GROUP BY the build key, for large probes: the probe
// rows are radix-partitioned by the range of their key's home slot
// (k_jp_hist, k_jp_scatter: one read of the inputs, one write of compact
// rows [key, validity bits if any input has NULLs, aggregate inputs]).
// k_jp_merge gives a partition to one workgroup, which loads the window of
// the build slots it can reach (its C home slots plus kJMargin slots of
// probe overflow: keys and row counts) with delta records into LDS, probes
// and aggregates there (LDS atomics, times the key's build row count), then
// adds each delta into its key's record with plain loads and stores: only
// the workgroup of a key's home range has rows for it.  A probe that walks
// past the window falls back to the global table with atomics.
constexpr int kJPBlock = 256;
constexpr int kJMBlock = 1024;
constexpr int kJMargin = 128;
constexpr int kJTile = 4;

__device__ __forceinline__ uint64_t jhash(const JTable &T, uint64_t key) {
  return dev::table_hash(key, T.seed);
}

__device__ __forceinline__ uint32_t jpart(const JTable &T, uint64_t key, int log2p) {
  return log2p ? (uint32_t)(jhash(T, key) >> (64 - log2p)) : 0u;
}

// Key of probe row i for the partition kernels.  FAST: one 8-byte
// integer-like key column, whose value is the packed key.
template <bool FAST>
__device__ __forceinline__ int jp_key(const KeySpec &K, int64_t i, uint64_t &key) {
  if (FAST) {
    const DCol &d = K.c[0];
    if (!dev::valid(d, i)) {
      key = 0;
      return K.null_safe ? 2 : 3;
    }
    key = static_cast<const uint64_t *>(d.values)[i + d.offset];
    return key == kEmpty ? 1 : 0;
  }
  return pack_key(K, i, key);
}

constexpr int kJPre = 4;

// Values a lane keeps for each of its rows in the FAST staged scatter with R
// rows a lane (the register budget): 1 at 12 rows, 2 at 8, kJPre at 4.
__host__ __device__ constexpr int js_nv(int R) { return R >= 12 ? 1 : R >= 8 ? 2 : kJPre; }

// The FAST rows (one 8-byte key column, up to kJPre one-word inputs) i0 +
// r * stride, r < R, below hi: every key, value and validity load is issued
// before any is tested (clamped row index, no per-lane branch), then
// cls[r] as jp_key (3 also past hi) and, for cls 0, the values with their
// validity bits (0 and unset otherwise).  NV: whether values are wanted.
template <int R, int NV>
__device__ __forceinline__ void jp_rows_fast(const KeySpec &K, const DCol *in, int nv, int64_t i0,
                                             int64_t stride, int64_t hi, uint64_t (&keys)[R],
                                             int (&cls)[R], uint64_t (&val)[R][NV > 0 ? NV : 1],
                                             uint64_t (&vb)[R]) {
  constexpr int NA = NV > 0 ? NV : 1;
  const DCol &kc = K.c[0];
  // validity bytes stay raw until every load is issued (a compare on a
  // loaded value at once would wait for it)
  uint32_t kbyte[R], vbyte[R][NA];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int64_t i = i0 + r * stride;
    const int64_t ic = i < hi ? i : hi - 1;
    keys[r] = dev::load_w64(kc, ic);
    kbyte[r] = kc.validity ? kc.validity[(ic + kc.offset) >> 3] : 0xFFu;
#pragma unroll
    for (int c = 0; c < NA; c++) {
      const bool use = NV > 0 && c < nv;  // uniform
      val[r][c] = use ? dev::load_w64(in[c], ic) : 0;
      vbyte[r][c] = use ? (in[c].validity ? in[c].validity[(ic + in[c].offset) >> 3] : 0xFFu) : 0u;
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int64_t i = i0 + r * stride;
    const int64_t ic = i < hi ? i : hi - 1;
    const bool kok = (kbyte[r] >> ((ic + kc.offset) & 7)) & 1;
    cls[r] = i >= hi ? 3 : !kok ? (K.null_safe ? 2 : 3) : keys[r] == kEmpty ? 1 : 0;
    if (!kok) keys[r] = 0;
    vb[r] = 0;
#pragma unroll
    for (int c = 0; c < NA; c++) {
      const bool ok = NV > 0 && cls[r] == 0 && ((vbyte[r][c] >> ((ic + in[c].offset) & 7)) & 1);
      val[r][c] = ok ? val[r][c] : 0;
      vb[r] |= (uint64_t)ok << c;
    }
  }
}

template <bool FAST>
__global__ void k_jp_hist(KeySpec K, int64_t n, JTable T, int log2p, int64_t per_block,
                          uint32_t *hist) {
  extern __shared__ uint32_t s_h[];
  const uint32_t np = 1u << log2p;
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) s_h[p] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = lo + per_block < n ? lo + per_block : n;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += 8 * (int64_t)blockDim.x) {
    uint64_t key[8];
    bool live[8];
    if (FAST) {
      int cls[8];
      uint64_t val[8][1], vb[8];
      jp_rows_fast<8, 0>(K, nullptr, 0, i0, blockDim.x, hi, key, cls, val, vb);
#pragma unroll
      for (int r = 0; r < 8; r++) live[r] = cls[r] == 0;
    } else {
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const int64_t i = i0 + r * (int64_t)blockDim.x;
        live[r] = i < hi && jp_key<FAST>(K, i, key[r]) == 0;
      }
    }
#pragma unroll
    for (int r = 0; r < 8; r++)
      if (live[r]) atomicAdd(&s_h[jpart(T, key[r], log2p)], 1u);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x)
    hist[(uint64_t)p * gridDim.x + blockIdx.x] = s_h[p];
}

// Compact rows by partition.  Keys that are NULL under `=` match nothing;
// the key equal to the EMPTY sentinel and a `<=>` NULL have fixed slots and
// are applied here with atomics.  A thread takes four rows; their keys and
// (up to kJPre inputs of one word) values are loaded before the cursor
// atomics, so the loads are in flight together.

__device__ __forceinline__ void jp_value(const DCol &d, int64_t i, bool &v, uint64_t &lo,
                                         uint64_t &hi) {
  v = dev::valid(d, i);
  lo = hi = 0;
  if (!v) return;
  if (dev::is_f(d.type)) {
    lo = (uint64_t)__double_as_longlong(dev::load_f64(d, i));
  } else {
    int64_t h;
    dev::load_i128(d, i, lo, h);
    hi = (uint64_t)h;
  }
}

template <bool FAST>
__global__ void k_jp_scatter(KeySpec K, int64_t n, JTable T, AggProg P, SpillFmt F, int log2p,
                             int64_t per_block, const uint64_t *base, uint64_t *out,
                             uint64_t *rec) {
  extern __shared__ uint32_t s_c[];
  const uint32_t np = 1u << log2p;
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x)
    s_c[p] = (uint32_t)base[(uint64_t)p * gridDim.x + blockIdx.x];
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = lo + per_block < n ? lo + per_block : n;
  bool pre = F.ncols <= kJPre;
  for (int c = 0; c < F.ncols && c < kJPre; c++)
    if (P.in[c].type == KG_T_DECIMAL128) pre = false;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += 4 * (int64_t)blockDim.x) {
    uint64_t keys[4];
    int cls4[4];
    uint64_t val[4][kJPre];
    uint64_t vb[4];
    if (FAST) {  // FAST inputs are one-word, at most kJPre: pre holds
      jp_rows_fast<4, kJPre>(K, P.in, F.ncols, i0, blockDim.x, hi, keys, cls4, val, vb);
    } else {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t i = i0 + r * (int64_t)blockDim.x;
        cls4[r] = i < hi ? jp_key<FAST>(K, i, keys[r]) : 3;
        vb[r] = 0;
#pragma unroll
        for (int c = 0; c < kJPre; c++) {
          val[r][c] = 0;
          if (pre && c < F.ncols && cls4[r] == 0) {
            bool v;
            uint64_t l, h;
            jp_value(P.in[c], i, v, l, h);
            val[r][c] = l;
            if (v) vb[r] |= 1ull << c;
          }
        }
      }
    }
    uint32_t pos4[4];
#pragma unroll
    for (int r = 0; r < 4; r++)
      pos4[r] = cls4[r] == 0 ? atomicAdd(&s_c[jpart(T, keys[r], log2p)], 1u) : 0u;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int64_t i = i0 + r * (int64_t)blockDim.x;
      const int cls = cls4[r];
      if (cls == 3) continue;
      if (cls != 0) {
        const uint64_t s = T.cap + (cls == 2 ? 1 : 0);
        const uint32_t m = T.slots[s].count;
        if (m) apply_row_mult(P, rec + s * (uint64_t)P.W, i, m);
        continue;
      }
      uint64_t *row = out + (uint64_t)pos4[r] * F.rw;
      row[0] = keys[r];
      if (pre) {
#pragma unroll
        for (int c = 0; c < kJPre; c++)
          if (c < F.ncols) row[F.off[c]] = val[r][c];
        if (F.vword) row[1] = vb[r];
        continue;
      }
      uint64_t vbits = 0;
      for (int c = 0; c < F.ncols; c++) {
        bool v;
        uint64_t l, h;
        jp_value(P.in[c], i, v, l, h);
        if (v) vbits |= 1ull << c;
        uint64_t *w = row + F.off[c];
        w[0] = l;
        if (P.in[c].type == KG_T_DECIMAL128) w[1] = h;
      }
      if (F.vword) row[1] = vbits;
    }
  }
}

// Staged form of k_jp_scatter for up to 2048 partitions and narrow rows: a
// workgroup sorts a tile of kJSBlock * R rows by partition in LDS (counts,
// block scan, staged rows), then writes each partition's run of the tile
// with consecutive lanes on consecutive rows, so a wave's stores cover a few
// contiguous runs instead of 64 scattered rows.  Rows keep their input order
// within a partition.  FAST: one 8-byte key column and at most kJPre
// one-word inputs of 8 bytes, loaded before the count atomics.
constexpr int kJSBlock = 512;
constexpr uint32_t kJSMaxParts = 2048;

// Exclusive scan of c[0..np) into s (np <= 4 * kJSBlock); returns the total.
__device__ __forceinline__ uint32_t js_block_scan(const uint32_t *c, uint32_t *s, uint32_t np,
                                                  uint32_t *wsum) {
  constexpr int PER = 4;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  uint32_t x[PER], tot = 0;
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const uint32_t p = (uint32_t)t * PER + j;
    x[j] = p < np ? c[p] : 0;
    tot += x[j];
  }
  uint32_t inc = tot;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  uint32_t woff = 0, all = 0;
  for (int w = 0; w < kJSBlock / 64; w++) {
    const uint32_t ws = wsum[w];
    if (w < wv) woff += ws;
    all += ws;
  }
  uint32_t run = woff + inc - tot;
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const uint32_t p = (uint32_t)t * PER + j;
    if (p < np) s[p] = run;
    run += x[j];
  }
  __syncthreads();
  return all;
}

__host__ __device__ constexpr size_t js_lds_bytes(int R, int rw, uint32_t np) {
  return (size_t)kJSBlock * R * rw * 8 + (size_t)np * 16 + (size_t)kJSBlock * R * 2 +
         kJSBlock / 64 * 4;
}

template <int R, bool FAST>
__global__ void __launch_bounds__(kJSBlock)
    k_jp_scatter_staged(KeySpec K, int64_t n, JTable T, AggProg P, SpillFmt F, int log2p,
                        int64_t per_block, const uint64_t *base, uint64_t *out, uint64_t *rec) {
  constexpr int TT = kJSBlock * R;
  extern __shared__ uint64_t smem[];
  const uint32_t np = 1u << log2p;
  const int rw = F.rw;
  uint64_t *st = smem;                                        // word w of staged row s: st[w * TT + s]
  uint64_t *cur = st + (size_t)rw * TT;                       // np: next output row of the block
  uint32_t *cnt = reinterpret_cast<uint32_t *>(cur + np);     // np: rows of the tile
  uint32_t *start = cnt + np;                                 // np: first staged row
  uint16_t *st_p = reinterpret_cast<uint16_t *>(start + np);  // TT: partition of staged row s
  uint32_t *wsum = reinterpret_cast<uint32_t *>(st_p + TT);   // kJSBlock / 64
  for (uint32_t p = threadIdx.x; p < np; p += kJSBlock)
    cur[p] = base[(uint64_t)p * gridDim.x + blockIdx.x];
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = lo + per_block < n ? lo + per_block : n;
  const int nv = F.ncols;
  for (int64_t t0 = lo; t0 < hi; t0 += TT) {
    for (uint32_t p = threadIdx.x; p < np; p += kJSBlock) cnt[p] = 0;
    __syncthreads();
    constexpr int NV = js_nv(R);  // FAST launches have nv <= NV
    uint64_t keys[R];
    uint32_t pp[R], lp[R];
    uint64_t val[R][NV];
    uint64_t vb[R];
    int cls[R];
    if (FAST) {
      jp_rows_fast<R, NV>(K, P.in, nv, t0 + threadIdx.x, kJSBlock, hi, keys, cls, val, vb);
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t i = t0 + (int64_t)r * kJSBlock + threadIdx.x;
        cls[r] = i < hi ? jp_key<FAST>(K, i, keys[r]) : 3;
        vb[r] = 0;
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      pp[r] = cls[r] == 0 ? jpart(T, keys[r], log2p) : 0xffffffffu;
      if (cls[r] == 1 || cls[r] == 2) {  // fixed slots (rare): atomics
        const int64_t i = t0 + (int64_t)r * kJSBlock + threadIdx.x;
        const uint64_t s = T.cap + (cls[r] == 2 ? 1 : 0);
        const uint32_t m = T.slots[s].count;
        if (m) apply_row_mult(P, rec + s * (uint64_t)P.W, i, m);
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) lp[r] = pp[r] != 0xffffffffu ? atomicAdd(&cnt[pp[r]], 1u) : 0u;
    __syncthreads();
    const uint32_t tn = js_block_scan(cnt, start, np, wsum);
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (pp[r] == 0xffffffffu) continue;
      const uint32_t s = start[pp[r]] + lp[r];
      st_p[s] = (uint16_t)pp[r];
      st[s] = keys[r];
      if (FAST) {
#pragma unroll
        for (int c = 0; c < NV; c++)
          if (c < nv) st[(size_t)F.off[c] * TT + s] = val[r][c];
      } else {
        const int64_t i = t0 + (int64_t)r * kJSBlock + threadIdx.x;
        for (int c = 0; c < nv; c++) {
          bool v;
          uint64_t l, h;
          jp_value(P.in[c], i, v, l, h);
          if (v) vb[r] |= 1ull << c;
          st[(size_t)F.off[c] * TT + s] = l;
          if (P.in[c].type == KG_T_DECIMAL128) st[(size_t)(F.off[c] + 1) * TT + s] = h;
        }
      }
      if (F.vword) st[TT + s] = vb[r];
    }
    __syncthreads();
    if (rw == 2) {
      for (uint32_t s = threadIdx.x; s < tn; s += kJSBlock) {
        const uint32_t p = st_p[s];
        const uint64_t d = cur[p] + (s - start[p]);
        const uint64_t a = st[s], b = st[TT + s];
        reinterpret_cast<uint4 *>(out)[d] =
            make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
      }
    } else {
      for (uint32_t s = threadIdx.x; s < tn; s += kJSBlock) {
        const uint32_t p = st_p[s];
        uint64_t *row = out + (cur[p] + (s - start[p])) * (uint64_t)rw;
        for (int w = 0; w < rw; w++) row[w] = st[(size_t)w * TT + s];
      }
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < np; p += kJSBlock) cur[p] += cnt[p];
    __syncthreads();
  }
}

// Probe of the LDS window for `key`: (slot, build row count) on a match,
// (-1, 0) on none, (-2, 0) when the probe walks past the window.
__device__ __forceinline__ int2 window_probe(const JTable &T, const uint64_t *s_keys,
                                             const uint32_t *s_cnt, int Wn, bool whole,
                                             uint64_t start, uint64_t key) {
  int l = (int)((jhash(T, key) >> (64 - T.log2cap)) - start);
  while (true) {
    if (l >= Wn) {
      if (!whole) return make_int2(-2, 0);
      l = 0;
    }
    const uint64_t k = s_keys[l];
    if (k == key) return s_cnt[l] ? make_int2(l, (int)s_cnt[l]) : make_int2(-1, 0);
    if (k == kEmpty) return make_int2(-1, 0);
    l++;
  }
}

__global__ void __launch_bounds__(kJMBlock) k_jp_merge(JTable T, AggProg P, SpillFmt F,
                                                       const uint64_t *rows,
                                                       const uint32_t *hist, uint64_t hn,
                                                       const uint64_t *base, uint32_t nb,
                                                       int log2p, int C, int Wn, uint64_t *rec) {
  // rows written by the scatter: the exclusive scan's last offset + count
  const uint64_t nrows = base[hn - 1] + hist[hn - 1];
  extern __shared__ uint64_t smem[];
  const int W = P.W, WL = W | 1;
  uint64_t *s_keys = smem;
  uint64_t *s_rec = s_keys + Wn;
  uint32_t *s_cnt = reinterpret_cast<uint32_t *>(s_rec + (uint64_t)Wn * WL);
  const uint32_t np = 1u << log2p;
  const bool whole = (uint64_t)Wn == T.cap;  // the window is the whole table
  for (uint32_t p = blockIdx.x; p < np; p += gridDim.x) {
    const uint64_t start = (uint64_t)p * C;
    for (int l = threadIdx.x; l < Wn; l += blockDim.x) {
      const uint64_t s = (start + l) & T.mask;
      s_keys[l] = T.slots[s].key;
      s_cnt[l] = T.slots[s].count;
    }
    fill_lds_records(P, s_rec, Wn, WL);
    __syncthreads();
    const int64_t a = (int64_t)base[(uint64_t)p * nb];
    const int64_t b = p + 1 < np ? (int64_t)base[(uint64_t)(p + 1) * nb] : (int64_t)nrows;
    for (int64_t t0 = a; t0 < b; t0 += (int64_t)kJTile * blockDim.x) {
      int ls[kJTile];
      uint64_t m[kJTile];
      bool fb[kJTile];
      SpillSrc<kJTile> src;
      src.rows = rows;
      src.safe = a;  // a < b inside this loop
      // every row's key load is issued before the first probe (clamped index)
      uint64_t kk[kJTile];
#pragma unroll
      for (int r = 0; r < kJTile; r++) {
        const int64_t i = t0 + (int64_t)r * blockDim.x + threadIdx.x;
        kk[r] = rows[(i < b ? i : b - 1) * F.rw];
      }
#pragma unroll
      for (int r = 0; r < kJTile; r++) {
        const int64_t i = t0 + (int64_t)r * blockDim.x + threadIdx.x;
        src.idx[r] = i;
        const int2 hit = i < b ? window_probe(T, s_keys, s_cnt, Wn, whole, start, kk[r])
                               : make_int2(-1, 0);
        ls[r] = hit.x >= 0 ? hit.x : -1;
        m[r] = hit.x >= 0 ? (uint64_t)(uint32_t)hit.y : 0;
        fb[r] = hit.x == -2;
      }
      tile_apply_mult<kJTile>(P, s_rec, WL, ls, m, src);
#pragma unroll
      for (int r = 0; r < kJTile; r++) {
        if (!fb[r]) continue;  // (rare) the global table, with atomics
        const uint64_t key = rows[src.idx[r] * F.rw];
        const int64_t s = jfind(T, 0, key);
        if (s < 0 || T.slots[s].count == 0) continue;
        SpillSrc<1> one;
        one.rows = rows;
        one.idx[0] = src.idx[r];
        const int z[1] = {0};
        const uint64_t mm[1] = {T.slots[s].count};
        tile_apply_mult<1, false>(P, rec + (uint64_t)s * W, 0, z, mm, one);
      }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < Wn; l += blockDim.x) {
      const uint64_t *d = s_rec + (uint64_t)l * WL;
      if (d[0] == 0) continue;
      merge_plain(P, rec + ((start + l) & T.mask) * (uint64_t)W, d);
    }
    __syncthreads();
  }
}

}  // namespace kg

using namespace kg;

// A driver's probe state: its stream and buffers, and the batch it is
// returning (kg_join_probe_push / kg_join_probe_next).
struct kg_join_local {
  kg_hash_join *j = nullptr;
  hipStream_t s = nullptr;
  DevBuf head, nout, off, temp, pidx, bidx, win;
  bool active = false;
  std::vector<DevBuf> held;  // host input columns staged to HBM
  std::vector<DCol> d;
  bool device_cols = false;
  const int32_t *sel = nullptr;
  int64_t np = 0;  // probe rows (the selected ones)
  uint64_t total = 0, emitted = 0;
  FilterProg F{};
  DevBuf fprog;  // F in device memory (kernels read it through a pointer)
};

struct kg_hash_join {
  kg_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  kg_join_spec spec{};
  uint64_t expected = 0;
  // build store: per column, the appended chunks, then one contiguous buffer
  struct BCol {
    int32_t type = 0, precision = 0, scale = 0;
    bool nullable = false;
    std::vector<std::pair<DevBuf, int64_t>> chunks;      // values, rows
    std::vector<std::pair<DevBuf, int64_t>> vchunks;     // validity bitmaps (bytes), rows
    DevBuf values, validity;
  };
  std::vector<BCol> bcols;
  int64_t build_rows = 0;
  bool finalized = false;
  // table
  DevBuf slots, next;
  uint64_t cap = 0;
  int log2cap = 0;
  uint64_t seed = 0x2545F4914F6CDD1DULL;
  // direct index (tiny integer builds)
  DevBuf dslot;
  int64_t dmin = 0;
  uint64_t dspan = 0;
  // fused aggregate
  bool has_agg = false;
  TuneCfg shape;  // the size class's shape when the table was built (its load
                  // factor sized the table; the partitioned probe takes the rest)
  AggProg prog{};
  OutProg out{};
  std::vector<kg_column> agg_types;
  DevBuf rec;
  // checked sums (decimal, ANSI integer) of the fused aggregate: each probe
  // batch is proven unable to overflow them before it is applied
  std::vector<CheckedSum> checked;
  uint64_t agg_rows = 0;  // probe rows applied so far
  uint64_t max_dup = 1;   // the most build rows of one key (a probe row's pairs)
  DevBuf abs_scratch;
  // small host batches gathered in HBM: build batches until finalize, fused
  // aggregate probe batches (with their input columns) until the result
  std::unique_ptr<kg::HostGather> bgather, pgather;
  std::vector<int32_t> pgather_inputs;
  // scan scratch of the fused aggregate and the tail
  DevBuf temp;
  // residual filter (column indices of the probe and build batches)
  std::vector<kg_expr> filter;
  // joins that return build rows by their matches: one byte a build row
  DevBuf bmatch;
  uint64_t null_rows = 0;  // build rows whose `=` key is NULL
  bool probed = false;
  std::unique_ptr<kg_join_local> own;  // the local of kg_join_probe
  // the build-side tail
  DevBuf tail_list;
  int64_t tail_n = -1, tail_emitted = 0;
};

namespace kg {

static JTable jtable(kg_hash_join *j) {
  JTable t{};
  t.slots = j->slots.get<Slot>();
  t.next = j->next.get<uint32_t>();
  t.cap = j->cap;
  t.log2cap = j->log2cap;
  t.mask = j->cap - 1;
  t.seed = j->seed;
  t.dslot = j->dspan ? j->dslot.get<int32_t>() : nullptr;
  t.dmin = j->dmin;
  t.dspan = j->dspan;
  return t;
}

static int jlog2(uint64_t v) {
  int b = 0;
  while ((1ull << b) < v) b++;
  return b;
}

static KeySpec key_spec(kg_hash_join *j, const DCol *cols, const int32_t *which) {
  KeySpec K{};
  K.n = j->spec.num_keys;
  K.null_safe = j->spec.null_safe[0];
  for (int c = 0; c < K.n; c++) K.c[c] = cols[which[c]];
  K.shift = K.n == 2 ? 8 * type_width(K.c[1].type) : 0;
  return K;
}

static void check_keys(const kg_join_spec &sp, int32_t ncols, const kg_column *cols,
                       const int32_t *which) {
  int width = 0;
  for (int c = 0; c < sp.num_keys; c++) {
    if (which[c] < 0 || which[c] >= ncols) fail(KG_ERR_INTERNAL, "join key column index");
    int t = cols[which[c]].type;
    int w = type_width(t);
    if (!(is_int_like(t) || is_float(t)) || w <= 0)
      fail(KG_ERR_UNSUPPORTED, "join key type on the device");
    width += w;
  }
  if (sp.num_keys == 2 && width > 8)
    fail(KG_ERR_UNSUPPORTED, "two join keys wider than 8 bytes together");
}

static void stage_all(Staged &st, int32_t ncols, const kg_column *cols, int64_t nrows,
                      std::vector<DCol> &out) {
  out.resize(ncols);
  for (int c = 0; c < ncols; c++) {
    check_column(cols[c], nrows, "join column");
    out[c] = st.add(cols[c], nrows);
  }
}

__global__ void k_copy_bits(const uint8_t *src, int64_t src_off, int64_t n, uint32_t *dst,
                            int64_t dst_off) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t s = src_off + i, d = dst_off + i;
    if ((src[s >> 3] >> (s & 7)) & 1) atomicOr(dst + (d >> 5), 1u << (d & 31));
  }
}

// Appends a build batch.  Every buffer is allocated before anything is
// changed, so a refusal for memory (NeedMemory) leaves the build as it was
// and the same batch can be sunk again (operator.rs:314-329).
static void append_build(kg_hash_join *j, int32_t ncols, const kg_column *cols, int64_t nrows) {
  if (!j->bcols.empty() && (int)j->bcols.size() != ncols)
    fail(KG_ERR_INTERNAL, "build batches differ in columns");
  for (int c = 0; c < (int)j->bcols.size(); c++)
    if (cols[c].type != j->bcols[c].type) fail(KG_ERR_INTERNAL, "build column type changed");
  Staged st(j->ctx, j->s);
  std::vector<DCol> d;
  stage_all(st, ncols, cols, nrows, d);
  std::vector<DevBuf> vals(ncols), valids(ncols);
  for (int c = 0; c < ncols; c++) {
    const int w = type_width(cols[c].type);
    // values: bit-packed booleans re-based to bit 0 as 32-bit words
    const uint64_t vbytes = w == 0 ? (uint64_t)ceil_div(nrows, 32) * 4 : (uint64_t)nrows * w;
    vals[c].reset(j->ctx, vbytes ? vbytes : 4);
    if (d[c].validity) valids[c].reset(j->ctx, (uint64_t)ceil_div(nrows, 32) * 4 + 4);
  }
  // nothing below allocates: the batch is appended
  if (j->bcols.empty()) {
    j->bcols.resize(ncols);
    for (int c = 0; c < ncols; c++) {
      j->bcols[c].type = cols[c].type;
      j->bcols[c].precision = cols[c].precision;
      j->bcols[c].scale = cols[c].scale;
    }
  }
  for (int c = 0; c < ncols; c++) {
    auto &bc = j->bcols[c];
    const int w = type_width(bc.type);
    DevBuf &v = vals[c];
    if (w == 0) {
      KG_HIP(hipMemsetAsync(v.get(), 0, v.bytes(), j->s));
      k_copy_bits<<<stream_grid(nrows, 256), 256, 0, j->s>>>(
          static_cast<const uint8_t *>(d[c].values), d[c].offset, nrows, v.get<uint32_t>(), 0);
    } else if (nrows > 0) {
      KG_HIP(hipMemcpyAsync(v.get(), static_cast<const uint8_t *>(d[c].values) + d[c].offset * w,
                            (size_t)nrows * w, hipMemcpyDeviceToDevice, j->s));
    }
    bc.chunks.emplace_back(std::move(v), nrows);
    if (d[c].validity) {
      bc.nullable = true;
      DevBuf &vb = valids[c];
      KG_HIP(hipMemsetAsync(vb.get(), 0, vb.bytes(), j->s));
      k_copy_bits<<<stream_grid(nrows, 256), 256, 0, j->s>>>(d[c].validity, d[c].offset, nrows,
                                                             vb.get<uint32_t>(), 0);
      bc.vchunks.emplace_back(std::move(vb), nrows);
    } else {
      bc.vchunks.emplace_back(DevBuf(), nrows);  // all valid
    }
    KG_HIP(hipGetLastError());
  }
  j->build_rows += nrows;
  KG_HIP(hipStreamSynchronize(j->s));
}

__global__ void k_set_bits(uint32_t *dst, int64_t off, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t d = off + i;
    atomicOr(dst + (d >> 5), 1u << (d & 31));
  }
}

// The contiguous buffers of every build column, allocated before the chunks
// are copied into them (concat_build), so finalize can refuse for memory
// with the build unchanged.
struct Arena {
  std::vector<DevBuf> values, validity;
};
static Arena alloc_arena(kg_hash_join *j) {
  const int64_t n = j->build_rows;
  Arena A;
  A.values.resize(j->bcols.size());
  A.validity.resize(j->bcols.size());
  for (size_t c = 0; c < j->bcols.size(); c++) {
    const auto &bc = j->bcols[c];
    const int w = type_width(bc.type);
    const uint64_t vbytes = w == 0 ? (uint64_t)ceil_div(n, 32) * 4 : (uint64_t)n * w;
    A.values[c].reset(j->ctx, vbytes + 16);
    if (bc.nullable) A.validity[c].reset(j->ctx, (uint64_t)ceil_div(n, 32) * 4 + 4);
  }
  return A;
}

// Concatenates the chunks of every build column into the arena's buffers
// (the row arena made contiguous).
static void concat_build(kg_hash_join *j, Arena &A) {
  for (size_t ci = 0; ci < j->bcols.size(); ci++) {
    auto &bc = j->bcols[ci];
    int w = type_width(bc.type);
    bc.values = std::move(A.values[ci]);
    if (w == 0) KG_HIP(hipMemsetAsync(bc.values.get(), 0, bc.values.bytes(), j->s));
    if (bc.nullable) {
      bc.validity = std::move(A.validity[ci]);
      KG_HIP(hipMemsetAsync(bc.validity.get(), 0, bc.validity.bytes(), j->s));
    }
    int64_t at = 0;
    for (size_t k = 0; k < bc.chunks.size(); k++) {
      int64_t rows = bc.chunks[k].second;
      if (w == 0) {
        k_copy_bits<<<stream_grid(rows, 256), 256, 0, j->s>>>(bc.chunks[k].first.get<uint8_t>(), 0,
                                                             rows, bc.values.get<uint32_t>(), at);
      } else if (rows) {
        KG_HIP(hipMemcpyAsync(bc.values.get<uint8_t>() + at * w, bc.chunks[k].first.get(),
                              (size_t)rows * w, hipMemcpyDeviceToDevice, j->s));
      }
      if (bc.nullable && rows) {
        if (bc.vchunks[k].first.bytes())
          k_copy_bits<<<stream_grid(rows, 256), 256, 0, j->s>>>(
              bc.vchunks[k].first.get<uint8_t>(), 0, rows, bc.validity.get<uint32_t>(), at);
        else
          k_set_bits<<<stream_grid(rows, 256), 256, 0, j->s>>>(bc.validity.get<uint32_t>(), at, rows);
      }
      at += rows;
    }
    KG_HIP(hipGetLastError());
    bc.chunks.clear();
    bc.vchunks.clear();
  }
  KG_HIP(hipStreamSynchronize(j->s));
}

static DCol build_dcol(const kg_hash_join::BCol &bc) {
  DCol d{};
  d.type = bc.type;
  d.precision = bc.precision;
  d.scale = bc.scale;
  d.width = type_width(bc.type);
  d.offset = 0;
  d.values = bc.values.get();
  d.validity = bc.nullable ? bc.validity.get<uint8_t>() : nullptr;
  return d;
}

// The direct index of a tiny build (join/direct.rs:96-160): one integer
// key, `=` only, at most kDirectRows build rows whose keys span at most
// kDirectSlots values and at most kDirectPerRow slots a row (a floor of 16
// rows).  Probes then find a key's slot with one subtraction and one load;
// chains, counts and emitted rows stay those of the hashed table.
constexpr int64_t kDirectRows = 4096;
constexpr uint64_t kDirectSlots = 16 * 1024;
constexpr uint64_t kDirectPerRow = 4;

static void build_direct(kg_hash_join *j) {
  j->dspan = 0;
  const int64_t n = j->build_rows;
  if (j->spec.no_direct_index || j->spec.num_keys != 1 || j->spec.null_safe[0] || n <= 0 ||
      n > kDirectRows)
    return;
  const int t = j->bcols[j->spec.build_keys[0]].type;
  const int w = type_width(t);
  if (w <= 0 || w > 8 || is_float(t)) return;  // integers, dates, timestamps, decimal64
  DevBuf mm(j->ctx, 16);
  PinnedReads pin;
  long long *init = pin.take<long long>(2), *got = pin.take<long long>(2);
  init[0] = LLONG_MAX;
  init[1] = LLONG_MIN;
  KG_HIP(hipMemcpyAsync(mm.get(), init, 16, hipMemcpyHostToDevice, j->s));
  k_slot_minmax<<<stream_grid(j->cap + 1, 256), 256, 0, j->s>>>(jtable(j), mm.get<long long>());
  KG_HIP(hipGetLastError());
  KG_HIP(hipMemcpyAsync(got, mm.get(), 16, hipMemcpyDeviceToHost, j->s));
  KG_HIP(hipStreamSynchronize(j->s));
  if (got[0] > got[1]) return;  // every key NULL
  const uint64_t span = (uint64_t)got[1] - (uint64_t)got[0] + 1;
  const uint64_t rows = (uint64_t)(n > 16 ? n : 16);
  if (span == 0 || span > kDirectSlots || span > kDirectPerRow * rows) return;
  j->dslot.reset(j->ctx, span * 4);
  KG_HIP(hipMemsetAsync(j->dslot.get(), 0xFF, span * 4, j->s));  // -1: no key
  JTable T = jtable(j);
  T.dmin = got[0];
  T.dspan = span;
  k_join_direct<<<stream_grid(j->cap + 1, 256), 256, 0, j->s>>>(T, j->dslot.get<int32_t>());
  KG_HIP(hipGetLastError());
  j->dmin = got[0];
  j->dspan = span;
}

static void scan_u32(kg_hash_join *j, const uint32_t *in, uint64_t *out, int64_t n) {
  j->temp.ensure(j->ctx, (uint64_t)scan_partials(n) * 8);
  exclusive_scan(in, out, n, j->temp.get<uint64_t>(), j->s);
}

// Whether a batch of host columns is small enough to gather (every column
// in host memory, fewer rows than a slot of the largest gather).
static bool gathers_host(int32_t ncols, const kg_column *cols, int64_t nrows) {
  if (nrows <= 0 || ncols < 1 || ncols > kGatherCols) return false;
  int width[kGatherCols];
  for (int c = 0; c < ncols; c++) {
    if (cols[c].flags & KG_COL_DEVICE) return false;
    width[c] = type_width(cols[c].type);
  }
  return nrows < gather_max_rows(width, ncols) / 8;
}

// A host batch into gather g: the gathered rows are taken first (flush)
// when it does not fit or differs in types; g starts at 64 times the first
// batch's rows and is replaced by one four times larger when it fills, up to
// the largest for the row width.
template <class Flush>
static void gather_into(kg_hash_join *j, std::unique_ptr<HostGather> &g, int32_t ncols,
                        const kg_column *cols, int64_t nrows, Flush flush) {
  const kg_column *cp[kGatherCols];
  int width[kGatherCols];
  for (int c = 0; c < ncols; c++) {
    cp[c] = &cols[c];
    width[c] = type_width(cols[c].type);
  }
  const int64_t maxcap = gather_max_rows(width, ncols);
  const int64_t want = gather_rows_for(nrows, maxcap);
  int64_t cap = want;
  if (g) {
    bool widths = g->ncols() == ncols;
    for (int c = 0; c < ncols && widths; c++) widths = width[c] == g->width(c);
    const bool fits = g->rows() + nrows <= g->cap(), in_slot = nrows < g->slot_rows();
    if (g->rows() > 0 && (!widths || !g->same_types(cp) || !fits || !in_slot)) flush();
    cap = g->cap();
    if (!fits && cap < maxcap) cap = std::min(cap * 4, maxcap);
    if (!in_slot || !widths) cap = std::max(cap, want);
    if (cap != g->cap() || !widths) g.reset();  // empty by now
  }
  if (!g) g = std::make_unique<HostGather>(j->ctx, ncols, width, cap);
  g->append(j->s, cp, nrows);
}

// The gathered build rows appended as one device batch (kept gathered when
// that fails, so a NeedMemory refusal leaves the build as it was).
static void flush_build_gather(kg_hash_join *j) {
  if (!j->bgather || j->bgather->rows() == 0) return;
  HostGather &g = *j->bgather;
  kg_column cols[kGatherCols];
  g.submit(j->s, cols);
  try {
    append_build(j, g.ncols(), cols, g.rows());
  } catch (...) {
    g.keep();
    throw;
  }
  g.clear();
}

}  // namespace kg

extern "C" {

int kg_join_create(kg_ctx *ctx, const kg_join_spec *spec, uint64_t expected_build_rows,
                   kg_hash_join **out) {
  return guarded([&] {
    *out = nullptr;
    if (spec->num_keys < 1 || spec->num_keys > 2)
      fail(KG_ERR_UNSUPPORTED, "one or two join keys on the device");
    if (spec->join_type < KG_JOIN_INNER || spec->join_type > KG_JOIN_FULL_OUTER)
      fail(KG_ERR_UNSUPPORTED, "join type on the device");
    if (spec->join_type == KG_JOIN_NULL_AWARE_ANTI &&
        (spec->num_keys != 1 || spec->null_safe[0]))
      fail(KG_ERR_UNSUPPORTED, "NOT IN takes one `=` key");
    if (spec->num_keys == 2 && (spec->null_safe[0] || spec->null_safe[1]))
      fail(KG_ERR_UNSUPPORTED, "null-safe keys in a two-key device join");
    auto j = std::make_unique<kg_hash_join>();
    j->ctx = ctx;
    j->spec = *spec;
    j->expected = expected_build_rows;
    j->s = make_stream(ctx->device);
    *out = j.release();
  });
}

int kg_join_build_sink(kg_hash_join *j, int32_t ncols, const kg_column *cols, int64_t nrows) {
  return guarded([&] {
    StreamScope kg_scope(j->s);
    if (j->finalized) fail(KG_ERR_INTERNAL, "build sink after finalize");
    if (ncols < 1 || ncols > kJoinCols) fail(KG_ERR_UNSUPPORTED, "1 to 32 build columns");
    check_keys(j->spec, ncols, cols, j->spec.build_keys);
    if (j->build_rows + (j->bgather ? j->bgather->rows() : 0) + nrows > (int64_t)0x7FFFFFF0)
      fail(KG_ERR_UNSUPPORTED, "a build side of 2^31 rows or more (grace partitions first)");
    // a first batch of no rows still fixes the build schema (an empty build
    // side's output columns)
    if (gathers_host(ncols, cols, nrows)) {
      if (!j->bcols.empty() && (int)j->bcols.size() != ncols)
        fail(KG_ERR_INTERNAL, "build batches differ in columns");
      for (int c = 0; c < (int)j->bcols.size(); c++)
        if (cols[c].type != j->bcols[c].type) fail(KG_ERR_INTERNAL, "build column type changed");
      if (j->bcols.empty() && j->bgather && j->bgather->rows() > 0) {  // the gathered schema
        if (j->bgather->ncols() != ncols) fail(KG_ERR_INTERNAL, "build batches differ in columns");
        for (int c = 0; c < ncols; c++)
          if (cols[c].type != j->bgather->type(c))
            fail(KG_ERR_INTERNAL, "build column type changed");
      }
      for (int c = 0; c < ncols; c++) check_column(cols[c], nrows, "join column");
      gather_into(j, j->bgather, ncols, cols, nrows, [&] { flush_build_gather(j); });
      return;
    }
    flush_build_gather(j);
    if (nrows > 0 || j->bcols.empty()) append_build(j, ncols, cols, nrows);
  });
}

int kg_join_build_finalize(kg_hash_join *j) {
  return guarded([&] {
    StreamScope kg_scope(j->s);
    if (j->finalized) return;
    flush_build_gather(j);
    int64_t n = j->build_rows;
    uint64_t want = (uint64_t)(n > (int64_t)j->expected ? n : (int64_t)j->expected);
    if (want < 512) want = 512;
    // capacity for the tuned load factor (50 % unless tuned), a power of two
    const TuneCfg tc = tune_for(j->ctx, (int64_t)want);
    j->shape = tc;
    const uint64_t pct = (uint64_t)std::min(90, std::max(10, (int)tc.load_pct));
    const uint64_t cap = 1ull << jlog2((want * 100 + pct - 1) / pct);
    // every buffer first: a refusal for memory leaves the build unfinalized
    // with its batches intact
    Arena arena = alloc_arena(j);
    DevBuf slots(j->ctx, (cap + 2) * sizeof(Slot)), next(j->ctx, (uint64_t)(n > 0 ? n : 1) * 4);
    DevBuf nulls(j->ctx, 8), bmatch;
    if (marks_build(j->spec.join_type)) bmatch.reset(j->ctx, (uint64_t)(n > 0 ? n : 1));
    concat_build(j, arena);
    j->cap = cap;
    j->log2cap = jlog2(cap);
    j->slots = std::move(slots);
    j->next = std::move(next);
    j->bmatch = std::move(bmatch);
    k_join_init<<<stream_grid(j->cap + 2, 256), 256, 0, j->s>>>(j->slots.get<Slot>(), j->cap + 2);
    KG_HIP(hipGetLastError());
    KG_HIP(hipMemsetAsync(nulls.get(), 0, 8, j->s));
    if (n > 0) {
      std::vector<DCol> d(j->bcols.size());
      for (size_t c = 0; c < d.size(); c++) d[c] = build_dcol(j->bcols[c]);
      KeySpec K = key_spec(j, d.data(), j->spec.build_keys);
      ProfScope prof("join_build", j->s);
      k_join_build<<<stream_grid(n, 256), 256, 0, j->s>>>(K, n, jtable(j),
                                                          nulls.get<unsigned long long>());
      KG_HIP(hipGetLastError());
    }
    if (j->bmatch.bytes()) KG_HIP(hipMemsetAsync(j->bmatch.get(), 0, j->bmatch.bytes(), j->s));
    PinnedReads pin;
    uint64_t *null_rows = pin.take<uint64_t>(1);
    KG_HIP(hipMemcpyAsync(null_rows, nulls.get(), 8, hipMemcpyDeviceToHost, j->s));
    // the direct index is an accelerator: without the memory for it the
    // hashed table serves the probes
    try {
      build_direct(j);
    } catch (const Error &e) {
      if (e.code != KG_ERR_MEMORY) throw;
      j->dspan = 0;
    }
    KG_HIP(hipStreamSynchronize(j->s));
    j->null_rows = *null_rows;
    j->finalized = true;
  });
}

int64_t kg_join_build_rows(const kg_hash_join *j) {
  return j ? j->build_rows + (j->bgather ? j->bgather->rows() : 0) : 0;
}

}  // extern "C"

namespace kg {

// The stored filter with its column operands resolved against a probe
// batch's columns and the build store: each distinct column once in cols.
static FilterProg resolve_filter(kg_hash_join *j, const std::vector<DCol> &probe) {
  FilterProg F{};
  F.n = (int)j->filter.size();
  for (int k = 0; k < F.n; k++) {
    kg_expr e = j->filter[k];
    if (e.op == KG_X_PROBE || e.op == KG_X_BUILD) {
      const int side = e.op == KG_X_BUILD;
      const int ncol = side ? (int)j->bcols.size() : (int)probe.size();
      if (e.arg < 0 || e.arg >= ncol) fail(KG_ERR_INTERNAL, "filter column index out of range");
      const DCol c = side ? build_dcol(j->bcols[e.arg]) : probe[e.arg];
      if (!is_int_like(c.type) || c.width <= 0)
        fail(KG_ERR_UNSUPPORTED, "residual filter over a column that is not integer-like");
      int slot = -1;
      for (int q = 0; q < F.ncols; q++)
        if (F.side[q] == side && F.cols[q].values == c.values && F.cols[q].offset == c.offset)
          slot = q;
      if (slot < 0) {
        if (F.ncols == kFilterCols) fail(KG_ERR_UNSUPPORTED, "a filter over more than 8 columns");
        slot = F.ncols++;
        F.cols[slot] = c;
        F.side[slot] = side;
      }
      e.arg = slot;
    }
    F.op[k] = e;
  }
  return F;
}

// The output columns a probe batch (or the tail) returns, allocated for
// `rows` rows; G gets the matching gather program.  probe: the probe
// batch's columns (for the tail, typeless stand-ins whose rows are all NULL).
static std::unique_ptr<kg_batch> out_batch(kg_hash_join *j, const std::vector<DCol> &probe,
                                           bool tail, int64_t rows, GatherCols &G,
                                           hipStream_t s) {
  const int jt = j->spec.join_type;
  const bool with_probe = !(jt == KG_JOIN_RIGHT_SEMI || jt == KG_JOIN_RIGHT_ANTI);
  const bool with_build = emits_pairs(jt) || !with_probe;
  const bool with_exists = jt == KG_JOIN_EXISTENCE;
  const int np = with_probe ? (int)probe.size() : 0;
  const int nb = with_build ? (int)j->bcols.size() : 0;
  const int nc = np + nb + (with_exists ? 1 : 0);
  if (nc > kJoinCols) fail(KG_ERR_UNSUPPORTED, "more than 32 output columns");
  auto b = std::make_unique<kg_batch>();
  b->ctx = j->ctx;
  b->nrows = rows;
  b->cols.resize(nc);
  G = GatherCols{};
  G.n = nc;
  const uint64_t vb = (uint64_t)ceil_div(rows, 32) * 4 + 4;
  for (int c = 0; c < nc; c++) {
    const int kind = c < np ? 0 : c < np + nb ? 1 : 2;
    DCol src{};
    if (kind == 0) src = probe[c];
    else if (kind == 1) src = build_dcol(j->bcols[c - np]);
    else src.type = KG_T_BOOL;
    auto &oc = b->cols[c];
    oc.type = src.type;
    oc.precision = src.precision;
    oc.scale = src.scale;
    const int w = type_width(src.type);
    const uint64_t bytes = w == 0 ? vb : (uint64_t)rows * w + 16;
    oc.values.reset(j->ctx, bytes);
    if (w == 0) KG_HIP(hipMemsetAsync(oc.values.get(), 0, bytes, s));
    const bool pads_build = jt == KG_JOIN_LEFT_OUTER || jt == KG_JOIN_FULL_OUTER;
    const bool nullable = kind == 2 ? false
                          : kind == 0 ? (tail || src.validity != nullptr)
                                      : (src.validity != nullptr || pads_build);
    if (nullable) {
      oc.validity.reset(j->ctx, vb);
      KG_HIP(hipMemsetAsync(oc.validity.get(), 0, vb, s));
    }
    G.src[c] = src;
    G.from_build[c] = kind;
    G.dst[c] = oc.values.get();
    // padded rows (row -1) keep a 0 validity bit: NULL
    G.dst_valid[c] = nullable ? oc.validity.get<uint32_t>() : nullptr;
  }
  return b;
}

static void check_probe_columns(kg_hash_join *j, int32_t ncols, const kg_column *cols,
                                int64_t nrows) {
  if (!j->finalized) fail(KG_ERR_INTERNAL, "probe before the build finalized");
  if (ncols < 1 || ncols > kJoinCols) fail(KG_ERR_UNSUPPORTED, "1 to 32 probe columns");
  if (nrows > INT32_MAX) fail(KG_ERR_UNSUPPORTED, "a probe batch of 2^31 rows or more");
  check_keys(j->spec, ncols, cols, j->spec.probe_keys);
  for (int c = 0; c < j->spec.num_keys; c++) {
    int bt = j->bcols.empty() ? cols[j->spec.probe_keys[c]].type
                              : j->bcols[j->spec.build_keys[c]].type;
    if (cols[j->spec.probe_keys[c]].type != bt)
      fail(KG_ERR_INTERNAL, "probe and build keys differ in type");
  }
}

static void probe_push(kg_join_local *l, int32_t ncols, const kg_column *cols, int64_t nrows,
                       const int32_t *sel, int64_t nsel) {
  kg_hash_join *j = l->j;
  if (l->active) fail(KG_ERR_INTERNAL, "the previous probe batch is not done");
  check_probe_columns(j, ncols, cols, nrows);
  if (sel && nsel > nrows) fail(KG_ERR_INTERNAL, "more selected rows than rows");
  if (nsel > 0 && !sel) fail(KG_ERR_INTERNAL, "no selection vector");
  j->probed = true;
  l->held.clear();
  {
    Staged st(j->ctx, l->s);
    stage_all(st, ncols, cols, nrows, l->d);
    l->held = st.take_buffers();
  }
  l->device_cols = false;
  for (int c = 0; c < ncols; c++) l->device_cols |= (cols[c].flags & KG_COL_DEVICE) != 0;
  l->F = resolve_filter(j, l->d);
  if (l->F.n) {
    l->fprog.ensure(j->ctx, sizeof(FilterProg));
    KG_HIP(hipMemcpyAsync(l->fprog.get(), &l->F, sizeof(FilterProg), hipMemcpyHostToDevice, l->s));
  }
  const FilterProg *fp = l->F.n ? l->fprog.get<FilterProg>() : nullptr;
  const int jt = j->spec.join_type;
  const int64_t np = sel ? nsel : nrows;
  uint64_t total = 0;
  const bool none = jt == KG_JOIN_NULL_AWARE_ANTI && j->null_rows > 0;  // NOT IN (.., NULL)
  if (np > 0 && !none) {
    KeySpec K = key_spec(j, l->d.data(), j->spec.probe_keys);
    l->head.ensure(j->ctx, (uint64_t)np * 4);
    l->nout.ensure(j->ctx, (uint64_t)np * 4);
    l->off.ensure(j->ctx, (uint64_t)(np + 1) * 8);
    l->temp.ensure(j->ctx, (uint64_t)(scan_partials(np) + 1) * 8);
    const int naaj_all = j->build_rows == 0;
    uint8_t *bm = marks_build(jt) ? j->bmatch.get<uint8_t>() : nullptr;
    {
      ProfScope prof(j->dspan ? "join_probe_direct" : "join_probe", l->s);
      const bool fast = K.n == 1 && type_width(K.c[0].type) == 8 && !is_float(K.c[0].type);
      const bool walk = fp || bm;
      const unsigned grid = stream_grid(np, 256, 2 * kJCU);
      uint32_t *hd = l->head.get<uint32_t>(), *no = l->nout.get<uint32_t>();
      if (fast && walk)
        k_join_count<true, true><<<grid, 256, 0, l->s>>>(K, np, sel, jtable(j), jt, fp, bm,
                                                         naaj_all, hd, no);
      else if (fast)
        k_join_count<true, false><<<grid, 256, 0, l->s>>>(K, np, sel, jtable(j), jt, fp, bm,
                                                          naaj_all, hd, no);
      else if (walk)
        k_join_count<false, true><<<grid, 256, 0, l->s>>>(K, np, sel, jtable(j), jt, fp, bm,
                                                          naaj_all, hd, no);
      else
        k_join_count<false, false><<<grid, 256, 0, l->s>>>(K, np, sel, jtable(j), jt, fp, bm,
                                                           naaj_all, hd, no);
      KG_HIP(hipGetLastError());
    }
    uint64_t *tot = l->temp.get<uint64_t>() + scan_partials(np);
    exclusive_scan(l->nout.get<uint32_t>(), l->off.get<uint64_t>(), np, l->temp.get<uint64_t>(),
                   l->s, nullptr, tot);
    KG_HIP(hipMemcpyAsync(&total, tot, 8, hipMemcpyDeviceToHost, l->s));
    KG_HIP(hipStreamSynchronize(l->s));
  }
  l->sel = sel;
  l->np = np;
  l->total = total;
  l->emitted = 0;
  l->active = true;
}

static void probe_next(kg_join_local *l, int64_t max_rows, kg_batch **out, int32_t *done) {
  kg_hash_join *j = l->j;
  *out = nullptr;
  *done = 0;
  if (!l->active) fail(KG_ERR_INTERNAL, "no probe batch pushed");
  if (max_rows < 1) fail(KG_ERR_INTERNAL, "max_rows must be positive");
  const int jt = j->spec.join_type;
  const uint64_t w0 = l->emitted;
  const uint64_t w1 = std::min<uint64_t>(l->total, w0 + (uint64_t)max_rows);
  const int64_t rows = (int64_t)(w1 - w0);
  GatherCols G;
  auto b = out_batch(j, l->d, false, rows, G, l->s);
  if (rows > 0) {
    l->win.ensure(j->ctx, 16);
    k_join_window<<<1, 64, 0, l->s>>>(l->off.get<uint64_t>(), l->nout.get<uint32_t>(), l->np, w0,
                                      w1, l->win.get<int64_t>());
    KG_HIP(hipGetLastError());
    int64_t range[2] = {0, 0};
    KG_HIP(hipMemcpyAsync(range, l->win.get(), 16, hipMemcpyDeviceToHost, l->s));
    KG_HIP(hipStreamSynchronize(l->s));
    const bool two = emits_pairs(jt) || jt == KG_JOIN_EXISTENCE;
    l->pidx.ensure(j->ctx, (uint64_t)rows * 4);
    if (two) l->bidx.ensure(j->ctx, (uint64_t)rows * 4);
    const FilterProg *fp = l->F.n ? l->fprog.get<FilterProg>() : nullptr;
    const auto pairs = fp ? k_join_pairs<true> : k_join_pairs<false>;
    pairs<<<stream_grid(range[1] - range[0], 256), 256, 0, l->s>>>(
        range[0], range[1], l->sel, l->head.get<uint32_t>(), l->nout.get<uint32_t>(),
        l->off.get<uint64_t>(), j->next.get<uint32_t>(), jt, fp, w0, w1, l->pidx.get<int32_t>(),
        two ? l->bidx.get<int32_t>() : nullptr);
    KG_HIP(hipGetLastError());
    ProfScope prof("join_gather", l->s);
    k_join_gather<<<stream_grid(rows, 256), 256, 0, l->s>>>(
        G, rows, l->pidx.get<int32_t>(), two ? l->bidx.get<int32_t>() : nullptr);
    KG_HIP(hipGetLastError());
  }
  KG_HIP(hipStreamSynchronize(l->s));
  if (l->device_cols) note_column_use(j->ctx, l->s);
  l->emitted = w1;
  if (l->emitted >= l->total) {
    *done = 1;
    l->active = false;
    l->held.clear();
  }
  *out = b.release();
}

}  // namespace kg

extern "C" {

int kg_join_set_filter(kg_hash_join *j, int32_t nops, const kg_expr *ops) {
  return guarded([&] {
    if (j->probed) fail(KG_ERR_INTERNAL, "the filter is set before the first probe");
    if (j->has_agg) fail(KG_ERR_UNSUPPORTED, "a residual filter with the fused aggregate");
    if (nops > 0 && j->spec.join_type == KG_JOIN_NULL_AWARE_ANTI)
      fail(KG_ERR_UNSUPPORTED, "NOT IN with a residual filter");
    if (nops < 0 || nops > kFilterOps) fail(KG_ERR_UNSUPPORTED, "a filter of more than 32 ops");
    int depth = 0;
    for (int k = 0; k < nops; k++) {
      const int op = ops[k].op;
      if (op == KG_X_PROBE || op == KG_X_BUILD || op == KG_X_CONST) depth++;
      else if (op == KG_X_NOT || op == KG_X_IS_NULL) {
        if (depth < 1) fail(KG_ERR_INTERNAL, "filter op without an operand");
      } else if (op >= KG_X_ADD && op <= KG_X_OR) {
        if (depth < 2) fail(KG_ERR_INTERNAL, "filter op without two operands");
        depth--;
      } else {
        fail(KG_ERR_INTERNAL, "unknown filter op");
      }
      if (depth > kFilterStack) fail(KG_ERR_UNSUPPORTED, "a filter deeper than 16");
    }
    if (nops > 0 && depth != 1) fail(KG_ERR_INTERNAL, "a filter must leave one value");
    j->filter.assign(ops, ops + nops);
  });
}

int kg_join_local_create(kg_hash_join *j, kg_join_local **out) {
  return guarded([&] {
    *out = nullptr;
    auto l = std::make_unique<kg_join_local>();
    l->j = j;
    l->s = make_stream(j->ctx->device);
    *out = l.release();
  });
}

void kg_join_local_destroy(kg_join_local *l) {
  if (!l) return;
  if (l->s) (void)hipStreamSynchronize(l->s);
  hipStream_t s = l->s;
  {
    StreamScope kg_scope(s);
    delete l;
  }
  destroy_stream(s);
}

int kg_join_probe_push(kg_join_local *l, int32_t ncols, const kg_column *cols, int64_t nrows,
                       const int32_t *sel_rows, int64_t nsel) {
  return guarded([&] {
    StreamScope kg_scope(l->s);
    probe_push(l, ncols, cols, nrows, sel_rows, sel_rows ? nsel : 0);
  });
}

int kg_join_probe_next(kg_join_local *l, int64_t max_rows, kg_batch **out, int32_t *done) {
  return guarded([&] {
    StreamScope kg_scope(l->s);
    probe_next(l, max_rows, out, done);
  });
}

static kg_join_local *own_local(kg_hash_join *j) {
  if (!j->own) {
    j->own = std::make_unique<kg_join_local>();
    j->own->j = j;
    j->own->s = j->s;
  }
  return j->own.get();
}

int kg_join_probe(kg_hash_join *j, int32_t ncols, const kg_column *cols, int64_t nrows,
                  kg_batch **out) {
  return guarded([&] {
    StreamScope kg_scope(j->s);
    kg_join_local *l = own_local(j);
    l->active = false;
    probe_push(l, ncols, cols, nrows, nullptr, 0);
    int32_t done = 0;
    probe_next(l, INT64_MAX, out, &done);
  });
}

int kg_join_probe_selected(kg_hash_join *j, int32_t ncols, const kg_column *cols, int64_t nrows,
                           const int32_t *sel_rows, int64_t nsel, kg_batch **out) {
  return guarded([&] {
    StreamScope kg_scope(j->s);
    if (nsel > 0 && !sel_rows) fail(KG_ERR_INTERNAL, "no selection vector");
    kg_join_local *l = own_local(j);
    l->active = false;
    probe_push(l, ncols, cols, nrows, sel_rows ? sel_rows : nullptr, sel_rows ? nsel : 0);
    int32_t done = 0;
    probe_next(l, INT64_MAX, out, &done);
  });
}

int kg_join_tail_next(kg_hash_join *j, int32_t nprobe_cols, const kg_column *probe_types,
                      int64_t max_rows, kg_batch **out, int32_t *done) {
  return guarded([&] {
    StreamScope kg_scope(j->s);
    *out = nullptr;
    *done = 1;
    const int jt = j->spec.join_type;
    if (!marks_build(jt)) return;
    if (!j->finalized) fail(KG_ERR_INTERNAL, "tail before the build finalized");
    if (max_rows < 1) fail(KG_ERR_INTERNAL, "max_rows must be positive");
    if (nprobe_cols < 0 || nprobe_cols > kJoinCols) fail(KG_ERR_UNSUPPORTED, "probe columns");
    const int64_t n = j->build_rows;
    if (j->tail_n < 0) {  // the returned build rows, in build order
      j->tail_n = 0;
      if (n > 0) {
        DevBuf flags(j->ctx, (uint64_t)n * 4), pos(j->ctx, (uint64_t)n * 8);
        j->temp.ensure(j->ctx, (uint64_t)(scan_partials(n) + 1) * 8);
        uint64_t *tot = j->temp.get<uint64_t>() + scan_partials(n);
        k_tail_flags<<<stream_grid(n, 256), 256, 0, j->s>>>(j->bmatch.get<uint8_t>(), n,
                                                             jt == KG_JOIN_RIGHT_SEMI,
                                                             flags.get<uint32_t>());
        KG_HIP(hipGetLastError());
        exclusive_scan(flags.get<uint32_t>(), pos.get<uint64_t>(), n, j->temp.get<uint64_t>(), j->s,
                       nullptr, tot);
        uint64_t cnt = 0;
        KG_HIP(hipMemcpyAsync(&cnt, tot, 8, hipMemcpyDeviceToHost, j->s));
        KG_HIP(hipStreamSynchronize(j->s));
        j->tail_list.reset(j->ctx, cnt * 4 + 4);
        k_tail_list<<<stream_grid(n, 256), 256, 0, j->s>>>(flags.get<uint32_t>(),
                                                            pos.get<uint64_t>(), n,
                                                            j->tail_list.get<int32_t>());
        KG_HIP(hipGetLastError());
        j->tail_n = (int64_t)cnt;
      }
      j->tail_emitted = 0;
    }
    std::vector<DCol> probe(nprobe_cols);
    for (int c = 0; c < nprobe_cols; c++) {
      probe[c] = DCol{};
      probe[c].type = probe_types[c].type;
      probe[c].precision = probe_types[c].precision;
      probe[c].scale = probe_types[c].scale;
      probe[c].width = type_width(probe_types[c].type);
      if (probe[c].width < 0) fail(KG_ERR_UNSUPPORTED, "probe column type");
    }
    const int64_t w0 = j->tail_emitted;
    const int64_t rows = std::min<int64_t>(j->tail_n - w0, max_rows);
    GatherCols G;
    auto b = out_batch(j, probe, true, rows, G, j->s);
    if (rows > 0) {
      ProfScope prof("join_gather", j->s);
      k_join_gather<<<stream_grid(rows, 256), 256, 0, j->s>>>(
          G, rows, nullptr, j->tail_list.get<int32_t>() + w0);
      KG_HIP(hipGetLastError());
    }
    KG_HIP(hipStreamSynchronize(j->s));
    j->tail_emitted = w0 + rows;
    *done = j->tail_emitted >= j->tail_n;
    *out = b.release();
  });
}

int kg_joinagg_create(kg_hash_join *j, int32_t ninputs, const kg_column *input_types,
                      int32_t naggs, const kg_agg_spec *aggs, uint32_t flags) {
  return guarded([&] {
    StreamScope kg_scope(j->s);
    if (!j->finalized) fail(KG_ERR_INTERNAL, "aggregate before the build finalized");
    if (j->spec.join_type != KG_JOIN_INNER || j->spec.num_keys != 1)
      fail(KG_ERR_UNSUPPORTED, "the fused aggregate needs an inner join on one key");
    if (!j->filter.empty()) fail(KG_ERR_UNSUPPORTED, "the fused aggregate with a residual filter");
    if (ninputs > kAggInputs) fail(KG_ERR_UNSUPPORTED, "at most 32 aggregate inputs");
    j->agg_types.assign(input_types, input_types + ninputs);
    std::vector<CheckedSum> checked;
    compile_aggs(j->agg_types, naggs, aggs, (flags & KG_AGG_ANSI) != 0, false, false, j->prog,
                 j->out, checked);
    for (int o = 0; o < j->prog.nops; o++)
      if (j->prog.ops[o].kind == OP_ADD192)
        fail(KG_ERR_UNSUPPORTED, "decimal128 sums in the fused join aggregate");
    j->checked = checked;
    j->agg_rows = 0;
    j->max_dup = 1;
    if (!checked.empty()) {  // a probe row counts once per build row of its key
      DevBuf mx(j->ctx, 8);
      KG_HIP(hipMemsetAsync(mx.get(), 0, 8, j->s));
      k_slot_max_count<<<stream_grid(j->cap + 2, 256), 256, 0, j->s>>>(
          j->slots.get<Slot>(), j->cap + 2, mx.get<unsigned int>());
      KG_HIP(hipGetLastError());
      PinnedReads pin;
      unsigned int *m = pin.take<unsigned int>(1);
      KG_HIP(hipMemcpyAsync(m, mx.get(), 4, hipMemcpyDeviceToHost, j->s));
      KG_HIP(hipStreamSynchronize(j->s));
      j->max_dup = *m > 1 ? *m : 1;
    }
    j->rec.reset(j->ctx, (j->cap + 2) * (uint64_t)j->prog.W * 8);
    init_records(j->ctx, j->rec.get<uint64_t>(), j->cap + 2, j->prog, j->s);
    j->has_agg = true;
  });
}

namespace kg {

// The partitioned probe of the fused join aggregate (k_jp_hist, k_jp_scatter,
// k_jp_merge).  Returns false (nothing launched) when the table needs more
// partitions than a workgroup can count in LDS; the caller then probes with
// global atomics.
static bool joinagg_partitioned(kg_hash_join *j, const KeySpec &K, const AggProg &P,
                                int64_t nrows) {
  const int W = P.W, WL = W | 1;
  const int ninputs = (int)j->agg_types.size();
  SpillFmt F{};
  F.ncols = ninputs;
  F.vword = 0;
  for (int c = 0; c < ninputs; c++)
    if (P.in[c].validity) F.vword = 1;
  int w = 1 + F.vword;
  for (int c = 0; c < ninputs; c++) {
    F.off[c] = (uint8_t)w;
    w += P.in[c].type == KG_T_DECIMAL128 ? 2 : 1;
  }
  F.rw = w;
  AggProg Q = P;
  set_spill_layout(Q, F);
  // one 8-byte integer-like key and one-word 8-byte inputs: the loads need
  // no type dispatch
  bool fast = K.n == 1 && type_width(K.c[0].type) == 8 && !is_float(K.c[0].type) &&
              ninputs <= kJPre;
  for (int c = 0; c < ninputs; c++)
    if (type_width(P.in[c].type) != 8) fast = false;
  // window: C home slots + the probe overflow margin
  int optin = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, j->ctx->device) !=
          hipSuccess || optin < 64 * 1024) {
    (void)hipGetLastError();
    optin = 64 * 1024;
  }
  // one 1024-lane workgroup a CU with the whole LDS: twice the window of two
  // 512-lane ones, so half the partitions (longer runs for the scatter);
  // the tuner may pick the two 512-lane ones
  const TuneCfg tc = j->shape;
  const int mlanes = tc.merge_lanes == 512 ? 512 : 1024;
  const uint64_t lds_all = (uint64_t)std::min(optin, 160 * 1024);
  const uint64_t lds_cap = (mlanes == 1024 ? lds_all : lds_all / 2) - 1024;
  const uint64_t per_slot = 8 + 4 + (uint64_t)WL * 8;
  uint64_t C = 1;
  while (C * 2 <= j->cap && (C * 2 + kJMargin) * per_slot <= lds_cap) C *= 2;
  if ((C + kJMargin) * per_slot > lds_cap && C > 1) C /= 2;
  // at least two partitions a CU where the table allows: a small table cut
  // into a few large windows would leave most CUs without a partition
  const int64_t cus = j->ctx->num_cus > 0 ? j->ctx->num_cus : 256;
  while (C > 1 && j->cap / C < (uint64_t)cus * 2) C /= 2;
  const int log2p = (int)(j->log2cap - __builtin_ctzll(C));
  const uint32_t np = 1u << log2p;
  if (np > 16384) return false;  // the per-block counters would not fit in LDS
  const int Wn = (int)std::min<uint64_t>(C + kJMargin, j->cap);
  const size_t lds = (size_t)Wn * per_slot;
  if (lds > lds_cap) fail(KG_ERR_UNSUPPORTED, "aggregate state too wide for the join window");
  KG_HIP(set_dynamic_lds((const void *)k_jp_merge, (int)lds));
  // staged scatter: tiles of kJSBlock * R rows sorted in LDS (R tuned, or
  // by the row width; a tuned R whose tile does not fit falls back)
  // (a FAST tile keeps at most js_nv(R) values a row in registers)
  int R = F.rw <= 2 ? 12 : F.rw <= 3 ? 8 : 4;
  if ((tc.scatter_rows == 4 || tc.scatter_rows == 8 || tc.scatter_rows == 12) &&
      js_lds_bytes(tc.scatter_rows, F.rw, np) + 1024 <= (size_t)optin &&
      (!fast || F.ncols <= js_nv(tc.scatter_rows)))
    R = tc.scatter_rows;
  if (fast && F.ncols > js_nv(R)) fail(KG_ERR_INTERNAL, "staged scatter: too many values a row");
  const size_t slds = js_lds_bytes(R, F.rw, np);
  const bool staged = np <= kJSMaxParts && F.rw <= 6 && slds + 1024 <= (size_t)optin;
  int64_t per_block;
  if (staged) {
    const int64_t tt = (int64_t)kJSBlock * R;
    per_block = ceil_div(ceil_div(nrows, cus * 4), tt) * tt;
  } else {
    per_block = std::max<int64_t>(4096, std::min<int64_t>((int64_t)np * 32, ceil_div(nrows, cus * 2)));
  }
  const int64_t nb = ceil_div(nrows, per_block);
  if (!staged) per_block = ceil_div(nrows, nb);
  const uint64_t hn = (uint64_t)np * nb;
  DevBuf hist(j->ctx, hn * 4 + 4), base(j->ctx, hn * 8 + 8);
  DevBuf rows(j->ctx, (uint64_t)nrows * F.rw * 8 + 8);
  JTable T = jtable(j);
  ProfScope prof("joinagg_probe", j->s);
  {
    ProfScope ph("jp_hist", j->s);
    if (fast)
      k_jp_hist<true><<<(unsigned)nb, kJPBlock, np * 4, j->s>>>(K, nrows, T, log2p, per_block,
                                                                hist.get<uint32_t>());
    else
      k_jp_hist<false><<<(unsigned)nb, kJPBlock, np * 4, j->s>>>(K, nrows, T, log2p, per_block,
                                                                 hist.get<uint32_t>());
    KG_HIP(hipGetLastError());
  }
  DevBuf temp(j->ctx, (uint64_t)scan_partials((int64_t)hn) * 8);
  exclusive_scan(hist.get<uint32_t>(), base.get<uint64_t>(), (int64_t)hn, temp.get<uint64_t>(), j->s);
  {
    ProfScope ps("jp_scatter", j->s);
    const uint64_t *bp = base.get<uint64_t>();
    uint64_t *rp = rows.get<uint64_t>(), *rec = j->rec.get<uint64_t>();
    const unsigned g = (unsigned)nb;
    if (staged) {
      const void *fn;
#define KG_JS(RR, FF)                                                                          \
  fn = (const void *)k_jp_scatter_staged<RR, FF>;                                              \
  KG_HIP(set_dynamic_lds(fn, slds));                                                           \
  k_jp_scatter_staged<RR, FF><<<g, kJSBlock, slds, j->s>>>(K, nrows, T, P, F, log2p, per_block, \
                                                           bp, rp, rec)
      if (fast) {
        if (R == 12) { KG_JS(12, true); } else if (R == 8) { KG_JS(8, true); } else { KG_JS(4, true); }
      } else {
        if (R == 12) { KG_JS(12, false); } else if (R == 8) { KG_JS(8, false); } else { KG_JS(4, false); }
      }
#undef KG_JS
    } else if (fast) {
      k_jp_scatter<true><<<g, kJPBlock, np * 4, j->s>>>(K, nrows, T, P, F, log2p, per_block, bp, rp,
                                                        rec);
    } else {
      k_jp_scatter<false><<<g, kJPBlock, np * 4, j->s>>>(K, nrows, T, P, F, log2p, per_block, bp, rp,
                                                         rec);
    }
    KG_HIP(hipGetLastError());
  }
  const unsigned grid = (unsigned)std::min<int64_t>(np, cus * (mlanes == 1024 ? 2 : 4));
  {
    ProfScope pm("jp_merge", j->s);
    k_jp_merge<<<grid, mlanes, lds, j->s>>>(T, Q, F, rows.get<uint64_t>(), hist.get<uint32_t>(),
                                               hn, base.get<uint64_t>(), (uint32_t)nb, log2p,
                                               (int)C, Wn, j->rec.get<uint64_t>());
    KG_HIP(hipGetLastError());
  }
  // the temporaries are freed in this stream's order: no synchronize
  return true;
}

}  // namespace kg

// One probe batch through the fused aggregate.
static void joinagg_probe_batch(kg_hash_join *j, int32_t ncols, const kg_column *cols,
                                int64_t nrows, const int32_t *input_cols) {
  Staged st(j->ctx, j->s);
  std::vector<DCol> d;
  stage_all(st, ncols, cols, nrows, d);
  KeySpec K = key_spec(j, d.data(), j->spec.probe_keys);
  AggProg P = j->prog;
  for (int c = 0; c < (int)j->agg_types.size(); c++) {
    int ic = input_cols[c];
    if (ic < 0 || ic >= ncols) fail(KG_ERR_INTERNAL, "aggregate input column index");
    if (d[ic].type != j->agg_types[c].type) fail(KG_ERR_INTERNAL, "aggregate input type");
    P.in[c] = d[ic];
  }
  if (nrows >= (1ll << 32)) fail(KG_ERR_UNSUPPORTED, "a probe batch of 2^32 rows or more");
  // checked sums: proven for this batch before anything is applied (a
  // batch that cannot be proven is refused, the state unchanged: the CPU
  // path replays such sums in row order)
  std::vector<double> add(j->checked.size(), 0.0);
  for (size_t k = 0; k < j->checked.size(); k++)
    if (!checked_batch_proven(j->ctx, j->s, j->abs_scratch, j->checked[k],
                              P.in[j->checked[k].col], nrows, j->agg_rows, j->max_dup, add[k]))
      fail(KG_ERR_UNSUPPORTED,
           "a decimal or ANSI sum of the fused join aggregate that this batch could overflow");
  const int WL = P.W | 1;
  const size_t dlds = (size_t)j->dspan * (WL * 8 + 4);
  int optin = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, j->ctx->device) !=
          hipSuccess || optin < 64 * 1024) {
    (void)hipGetLastError();
    optin = 64 * 1024;
  }
  if (j->dspan && dlds + 1024 <= (size_t)optin && nrows >= (1 << 16)) {
    // direct index, records privatized in LDS
    ProfScope prof("joinagg_direct", j->s);
    const bool dfast = K.n == 1 && !K.null_safe && type_width(K.c[0].type) == 8 &&
                       !is_float(K.c[0].type);
    const void *dfn = dfast ? (const void *)k_joinagg_direct<true>
                            : (const void *)k_joinagg_direct<false>;
    KG_HIP(set_dynamic_lds(dfn, (int)dlds));
    const int64_t cus = j->ctx->num_cus > 0 ? j->ctx->num_cus : 256;
    const int64_t per_cu = dlds <= 40 * 1024 ? 4 : dlds <= 80 * 1024 ? 2 : 1;
    const int64_t grid = std::min<int64_t>(cus * per_cu,
                                           ceil_div(nrows, (int64_t)kJDBlock * kJDTile * 8));
    if (dfast)
      k_joinagg_direct<true><<<(unsigned)std::max<int64_t>(grid, 1), kJDBlock, dlds, j->s>>>(
          K, nrows, jtable(j), P, j->rec.get<uint64_t>());
    else
      k_joinagg_direct<false><<<(unsigned)std::max<int64_t>(grid, 1), kJDBlock, dlds, j->s>>>(
          K, nrows, jtable(j), P, j->rec.get<uint64_t>());
    KG_HIP(hipGetLastError());
  } else if (nrows < (1 << 20) || !joinagg_partitioned(j, K, P, nrows)) {
    ProfScope prof(j->dspan ? "joinagg_probe_direct" : "joinagg_probe", j->s);
    k_joinagg_probe<<<stream_grid(nrows, 256), 256, 0, j->s>>>(K, nrows, jtable(j), P,
                                                               j->rec.get<uint64_t>());
    KG_HIP(hipGetLastError());
  }
  if (st.staged_host()) KG_HIP(hipStreamSynchronize(j->s));
  j->agg_rows += (uint64_t)nrows;
  for (size_t k = 0; k < j->checked.size(); k++) j->checked[k].cum_abs += add[k];
}

// The gathered probe batches through the fused aggregate (kept gathered when
// that fails: a refusal leaves the aggregate as it was).
static void flush_probe_gather(kg_hash_join *j) {
  if (!j->pgather || j->pgather->rows() == 0) return;
  HostGather &g = *j->pgather;
  kg_column cols[kGatherCols];
  g.submit(j->s, cols);
  try {
    joinagg_probe_batch(j, g.ncols(), cols, g.rows(), j->pgather_inputs.data());
  } catch (...) {
    g.keep();
    throw;
  }
  g.clear();
}

int kg_joinagg_probe(kg_hash_join *j, int32_t ncols, const kg_column *cols, int64_t nrows,
                     const int32_t *input_cols) {
  return guarded([&] {
    StreamScope kg_scope(j->s);
    if (!j->has_agg) fail(KG_ERR_INTERNAL, "no fused aggregate");
    if (nrows <= 0) return;
    check_keys(j->spec, ncols, cols, j->spec.probe_keys);
    const int na = (int)j->agg_types.size();
    for (int c = 0; c < na; c++) {
      const int ic = input_cols[c];
      if (ic < 0 || ic >= ncols) fail(KG_ERR_INTERNAL, "aggregate input column index");
      if (cols[ic].type != j->agg_types[c].type) fail(KG_ERR_INTERNAL, "aggregate input type");
    }
    if (gathers_host(ncols, cols, nrows)) {
      for (int c = 0; c < ncols; c++) check_column(cols[c], nrows, "join column");
      const bool same_inputs = j->pgather_inputs.size() == (size_t)na &&
                               std::equal(input_cols, input_cols + na, j->pgather_inputs.begin());
      if (!same_inputs) {
        flush_probe_gather(j);
        j->pgather_inputs.assign(input_cols, input_cols + na);
      }
      gather_into(j, j->pgather, ncols, cols, nrows, [&] { flush_probe_gather(j); });
      return;
    }
    flush_probe_gather(j);
    joinagg_probe_batch(j, ncols, cols, nrows, input_cols);
  });
}

int kg_joinagg_result(kg_hash_join *j, kg_batch **out) {
  return guarded([&] {
    StreamScope kg_scope(j->s);
    *out = nullptr;
    if (!j->has_agg) fail(KG_ERR_INTERNAL, "no fused aggregate");
    flush_probe_gather(j);
    DevBuf pos;
    uint64_t ns = j->cap + 2;
    int64_t n = group_positions(j->ctx, j->rec.get<uint64_t>(), ns, j->prog.W, false, pos, j->s);
    auto b = std::make_unique<kg_batch>();
    b->ctx = j->ctx;
    b->nrows = n;
    b->cols.resize(1 + j->out.n);
    const auto &kb = j->bcols[j->spec.build_keys[0]];
    auto &kc = b->cols[0];
    kc.type = kb.type;
    kc.precision = kb.precision;
    kc.scale = kb.scale;
    int kw = type_width(kb.type);
    if (kw <= 0) fail(KG_ERR_UNSUPPORTED, "boolean join key in the fused aggregate");
    uint64_t vb = (uint64_t)ceil_div(n, 32) * 4 + 4;
    // validity words are not cleared: the output kernel writes every word
    // that holds a bit below n
    kc.values.reset(j->ctx, (uint64_t)n * kw + 8);
    kc.validity.reset(j->ctx, vb);
    OutProg O = j->out;
    O.cap = j->cap;
    O.key_type = kb.type;
    O.direct = 0;
    O.global = 0;
    for (int a = 0; a < O.n; a++) {
      auto &c = b->cols[1 + a];
      c.type = O.a[a].out_type;
      c.precision = O.a[a].out_p;
      c.scale = O.a[a].out_s;
      int w = type_width(c.type);
      c.values.reset(j->ctx, (uint64_t)n * (w > 0 ? w : 1) + 16);
      c.validity.reset(j->ctx, vb);
      O.a[a].values = c.values.get();
      O.a[a].validity = c.validity.get<uint8_t>();
    }
    DevBuf err(j->ctx, 4);
    KG_HIP(hipMemsetAsync(err.get(), 0, 4, j->s));
    launch_agg_output(O, &j->slots.get<Slot>()->key, (int)(sizeof(Slot) / 8),
                      j->rec.get<uint64_t>(), pos.get<uint64_t>(), n, kc.values.get(),
                      kc.validity.get<uint8_t>(), kw, err.get<unsigned int>(), j->s);
    PinnedReads pin;
    unsigned int *e = pin.take<unsigned int>(1);
    KG_HIP(hipMemcpyAsync(e, err.get(), 4, hipMemcpyDeviceToHost, j->s));
    KG_HIP(hipStreamSynchronize(j->s));
    if (*e & 1u) fail(KG_ERR_OVERFLOW, "Overflow in sum");
    if (*e & 2u) fail(KG_ERR_OUT_OF_RANGE, "average out of the result decimal's range");
    *out = b.release();
  });
}

void kg_join_destroy(kg_hash_join *j) {
  if (!j) return;
  if (j->s) (void)hipStreamSynchronize(j->s);
  hipStream_t s = j->s;
  {
    StreamScope kg_scope(s);
    delete j;
  }
  destroy_stream(s);
}

}  // extern "C"
