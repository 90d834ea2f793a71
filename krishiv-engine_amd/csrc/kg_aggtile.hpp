// kg_aggtile.hpp -- op-major application of a tile of rows to LDS records,
// and the row format of the aggregation spill.
//
// A lane holds R rows of a tile.  The op program is walked once per tile
// (each op's uniform dispatch is paid once for R rows), and a column's
// values are loaded once per tile with one type dispatch.  The rows come
// either from the input columns (ColSrc) or from spill rows (SpillSrc).
//
// Spill row (64-bit words): [0] key, [1] validity of input column c in bit
// vshift + c, then each input column's value at SpillFmt::off[c]: one word
// (integers sign-extended, floats as float64 bits) or two (decimal128 lo,
// hi).  One 32-bit integer column may be packed into the low half of word 1
// (packed = its index + 1; the validity bits then start at bit 32), which
// makes the benchmark's row 32 bytes instead of 40.  Bit 63 of word 1 marks
// a dead row (validity bits stop below it).
// Spilled record: [0] key, then the W state words.
#pragma once

#include "kg_aggops.hpp"

namespace kg {

struct SpillFmt {
  int rw;      // words per spill row
  int ncols;   // input columns carried
  int vword;   // 1: word 1 holds the validity bits; 0: every value is valid
  uint8_t off[kAggInputs];
  uint8_t packed = 0;  // 1 + the column held in word 1's low half (0: none)
  uint8_t vshift = 0;  // the first validity bit of word 1
};

// Fills n LDS records (stride WL words) with the program's initial state.
// The word index is uniform across the workgroup: a lane-dependent index
// into the kernel argument can make the compiler copy the whole program to
// scratch.
__device__ __forceinline__ void fill_lds_records(const AggProg &P, uint64_t *s_rec, int n, int WL) {
  for (int w = 0; w < WL; w++) {
    const uint64_t v = P.init[w];
    for (int s = threadIdx.x; s < n; s += blockDim.x) s_rec[(size_t)s * WL + w] = v;
  }
}

// Copies a spill row layout into a program (host side, before a launch whose
// kernels read spill rows).
inline void set_spill_layout(AggProg &P, const SpillFmt &F) {
  P.spill_rw = F.rw;
  P.spill_vword = F.vword;
  P.spill_packed = F.packed;
  P.spill_vshift = F.vshift;
  for (int c = 0; c < kAggInputs; c++) P.spill_off[c] = F.off[c];
}

// Loads of input column c for the R rows of a lane (rows with act[r] false
// are skipped; their ok[r] is false).
template <int R>
struct ColSrc {
  int64_t idx[R];
  // a row every lane may read (>= 0): rows not active then load it instead
  // of branching, so the loads of a lane's rows are in flight together; -1
  // when there is none (the loads are then taken per active row)
  int64_t safe = -1;
  __device__ __forceinline__ void load(const AggProg &P, int c, const bool (&act)[R],
                                       uint64_t (&lo)[R], uint64_t (&hi)[R], double (&f)[R],
                                       bool (&ok)[R]) const {
    const DCol &d = P.in[c];
    const int t = d.type;
    const int64_t off = d.offset;
    if (safe >= 0 && (t == KG_T_FLOAT64 || t == KG_T_INT64 || t == KG_T_DECIMAL64 ||
                      t == KG_T_TIMESTAMP || t == KG_T_INT32 || t == KG_T_DATE32)) {
      // one word a row, no per-lane branch (a row not active reads `safe`);
      // the validity bytes stay raw until every load is issued
      uint32_t vb[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t j = (act[r] ? idx[r] : safe) + off;
        lo[r] = (t == KG_T_INT32 || t == KG_T_DATE32)
                    ? (uint64_t)(int64_t) static_cast<const int32_t *>(d.values)[j]
                    : static_cast<const uint64_t *>(d.values)[j];
        vb[r] = d.validity ? d.validity[j >> 3] : 0xFFu;
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t j = (act[r] ? idx[r] : safe) + off;
        f[r] = __longlong_as_double((long long)lo[r]);  // read for FLOAT64 only
        hi[r] = (uint64_t)((int64_t)lo[r] >> 63);
        ok[r] = act[r] && ((vb[r] >> (j & 7)) & 1);
      }
      return;
    }
    if (t == KG_T_FLOAT64) {
      const double *v = static_cast<const double *>(d.values);
#pragma unroll
      for (int r = 0; r < R; r++)
        if (act[r]) f[r] = v[idx[r] + off];
    } else if (t == KG_T_FLOAT32) {
      const float *v = static_cast<const float *>(d.values);
#pragma unroll
      for (int r = 0; r < R; r++)
        if (act[r]) f[r] = (double)v[idx[r] + off];
    } else if (t == KG_T_INT64 || t == KG_T_DECIMAL64 || t == KG_T_TIMESTAMP) {
      const int64_t *v = static_cast<const int64_t *>(d.values);
#pragma unroll
      for (int r = 0; r < R; r++)
        if (act[r]) {
          int64_t x = v[idx[r] + off];
          lo[r] = (uint64_t)x;
          hi[r] = (uint64_t)(x >> 63);
        }
    } else if (t == KG_T_INT32 || t == KG_T_DATE32) {
      const int32_t *v = static_cast<const int32_t *>(d.values);
#pragma unroll
      for (int r = 0; r < R; r++)
        if (act[r]) {
          int64_t x = v[idx[r] + off];
          lo[r] = (uint64_t)x;
          hi[r] = (uint64_t)(x >> 63);
        }
    } else {
#pragma unroll
      for (int r = 0; r < R; r++)
        if (act[r]) {
          int64_t h;
          dev::load_i128(d, idx[r], lo[r], h);
          hi[r] = (uint64_t)h;
          if (dev::is_f(t)) f[r] = dev::load_f64(d, idx[r]);
        }
    }
    if (d.validity) {
#pragma unroll
      for (int r = 0; r < R; r++) ok[r] = act[r] && dev::valid(d, idx[r]);
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) ok[r] = act[r];
    }
  }
};

// Loads from spill rows: rows[idx[r] * P.spill_rw ...] (the layout is in the
// program, a kernel argument: no pointer to a kernel argument is kept, which
// would make the compiler copy it to scratch).
template <int R>
struct SpillSrc {
  const uint64_t *rows;
  int64_t idx[R];
  // a row index every lane may read (>= 0): rows not active then load it
  // instead of branching, so a lane's loads are all in flight together;
  // -1 when there is none (the loads are then taken per active row)
  int64_t safe = -1;
  __device__ __forceinline__ void load(const AggProg &P, int c, const bool (&act)[R],
                                       uint64_t (&lo)[R], uint64_t (&hi)[R], double (&f)[R],
                                       bool (&ok)[R]) const {
    const int t = P.in[c].type;
    const int rw = P.spill_rw, o = P.spill_off[c];
    const bool wide = t == KG_T_DECIMAL128, uns = t == KG_T_UINT64;
    uint64_t v[R], vb[R], w2[R];
    // every array is indexed by the unrolled r only (a loop that is not
    // unrolled would put lo/hi/f in scratch)
    if (safe >= 0) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const uint64_t *row = rows + (act[r] ? idx[r] : safe) * (int64_t)rw;
        v[r] = row[o];
        vb[r] = P.spill_vword ? row[1] : ~0ull;
        w2[r] = wide ? row[o + 1] : 0;
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const uint64_t *row = rows + idx[r] * (int64_t)rw;
        v[r] = act[r] ? row[o] : 0;
        vb[r] = act[r] && P.spill_vword ? row[1] : ~0ull;
        w2[r] = act[r] && wide ? row[o + 1] : 0;
      }
    }
    const bool packed = P.spill_packed == c + 1;
    const int vs = P.spill_vshift + c;
#pragma unroll
    for (int r = 0; r < R; r++) {
      ok[r] = act[r] && ((vb[r] >> vs) & 1);
      if (packed) v[r] = (uint64_t)(int64_t)(int32_t)(uint32_t)v[r];
      f[r] = __longlong_as_double((long long)v[r]);
      lo[r] = v[r];
      hi[r] = wide ? w2[r] : uns ? 0 : (uint64_t)((int64_t)v[r] >> 63);
    }
  }
};

// Applies the op program to records srec + ls[r] * WL for the lane's rows
// with ls[r] >= 0 (LDS records, or with LDS false global records).
template <int R, bool LDS = true, class Src>
__device__ __forceinline__ void tile_apply(const AggProg &P, uint64_t *srec, int WL,
                                           const int (&ls)[R], const Src &src) {
  bool act[R];
#pragma unroll
  for (int r = 0; r < R; r++) act[r] = ls[r] >= 0;
  uint64_t lo[R], hi[R];
  double f[R];
  bool ok[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    lo[r] = hi[r] = 0;
    f[r] = 0.0;
    ok[r] = false;
  }
  int last = -1;
  for (int o = 0; o < P.nops; o++) {
    const AggOp op = op_at(P, o);
    if (op.kind == OP_COUNT_STAR) {
#pragma unroll
      for (int r = 0; r < R; r++)
        if (act[r]) add_u64_nr<LDS>(srec + ls[r] * WL, 1);
      continue;
    }
    if (op.col != last) {
      last = op.col;
      src.load(P, op.col, act, lo, hi, f, ok);
    }
    if (op.kind == OP_NULLS) {
#pragma unroll
      for (int r = 0; r < R; r++)
        if (act[r] && !ok[r]) add_u64_nr<LDS>(srec + ls[r] * WL + op.word, 1);
      continue;
    }
    if (op.kind == OP_STICKY) {
      uint64_t lo2[R], hi2[R];
      double f2[R];
      bool ok2[R];
      src.load(P, op.col2, act, lo2, hi2, f2, ok2);
#pragma unroll
      for (int r = 0; r < R; r++)
        if (act[r] && !ok[r] && ok2[r] && lo2[r] != 0)
          max_u64<LDS>(srec + ls[r] * WL + op.st_word, 2);
      continue;
    }
    if (op.checked && P.skip_checked) continue;
    switch (op.kind) {
      case OP_ADDV:
      case OP_ADD64:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r] && lo[r]) add_u64_nr<LDS>(srec + ls[r] * WL + op.word, lo[r]);
        break;
      case OP_COUNT_TRUE:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r] && lo[r]) add_u64_nr<LDS>(srec + ls[r] * WL + op.word, 1);
        break;
      case OP_ADD128: {
        // the rows' low-word adds issued together, one wait, then the carries
        uint64_t old[R];
#pragma unroll
        for (int r = 0; r < R; r++)
          old[r] = ok[r] ? add_u64<LDS>(srec + ls[r] * WL + op.word, lo[r]) : 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const uint64_t h = hi[r] + ((old[r] + lo[r]) < old[r] ? 1 : 0);
          if (ok[r] && h) add_u64_nr<LDS>(srec + ls[r] * WL + op.word + 1, h);
        }
        break;
      }
      case OP_ADD192: {
        uint64_t old[R], mid[R], c1[R];
#pragma unroll
        for (int r = 0; r < R; r++)
          old[r] = ok[r] ? add_u64<LDS>(srec + ls[r] * WL + op.word, lo[r]) : 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const uint64_t c0 = (old[r] + lo[r]) < old[r] ? 1 : 0;
          mid[r] = hi[r] + c0;
          c1[r] = mid[r] < hi[r] ? 1 : 0;  // hi = ~0 with a carry in
        }
        uint64_t old1[R];
#pragma unroll
        for (int r = 0; r < R; r++)
          old1[r] = (ok[r] && mid[r]) ? add_u64<LDS>(srec + ls[r] * WL + op.word + 1, mid[r]) : 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const uint64_t c2 = (ok[r] && mid[r] && (old1[r] + mid[r]) < old1[r]) ? 1 : 0;
          const uint64_t top = (uint64_t)((int64_t)hi[r] >> 63) + c1[r] + c2;
          if (ok[r] && top) add_u64_nr<LDS>(srec + ls[r] * WL + op.word + 2, top);
        }
        break;
      }
      case OP_ADDF64:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) add_f64<LDS>(srec + ls[r] * WL + op.word, f[r]);
        break;
      case OP_MIN_I64:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) min_i64<LDS>(srec + ls[r] * WL + op.word, (int64_t)lo[r]);
        break;
      case OP_MAX_I64:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) max_i64<LDS>(srec + ls[r] * WL + op.word, (int64_t)lo[r]);
        break;
      case OP_MIN_F:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) min_u64<LDS>(srec + ls[r] * WL + op.word, f64_order(f[r]));
        break;
      case OP_MAX_F:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) max_u64<LDS>(srec + ls[r] * WL + op.word, f64_order(f[r]));
        break;
      default: break;
    }
  }
}

// The join fused with GROUP BY the build key: row r counts m[r] times (the
// join pairs it makes).  Counts and sums scale by m; min/max do not.
// Programs with checked or 192-bit sums are not built for this path.
template <int R, bool LDS = true, class Src>
__device__ __forceinline__ void tile_apply_mult(const AggProg &P, uint64_t *srec, int WL,
                                                const int (&ls)[R], const uint64_t (&m)[R],
                                                const Src &src) {
  bool act[R];
#pragma unroll
  for (int r = 0; r < R; r++) act[r] = ls[r] >= 0;
  uint64_t lo[R], hi[R];
  double f[R];
  bool ok[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    lo[r] = hi[r] = 0;
    f[r] = 0.0;
    ok[r] = false;
  }
  int last = -1;
  for (int o = 0; o < P.nops; o++) {
    const AggOp op = op_at(P, o);
    if (op.kind == OP_COUNT_STAR) {
#pragma unroll
      for (int r = 0; r < R; r++)
        if (act[r]) add_u64_nr<LDS>(srec + ls[r] * WL, m[r]);
      continue;
    }
    if (op.col != last) {
      last = op.col;
      src.load(P, op.col, act, lo, hi, f, ok);
    }
    uint64_t *base = srec + op.word;
    switch (op.kind) {
      case OP_NULLS:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (act[r] && !ok[r]) add_u64_nr<LDS>(base + ls[r] * WL, m[r]);
        break;
      case OP_ADD64:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) add_u64_nr<LDS>(base + ls[r] * WL, lo[r] * m[r]);
        break;
      case OP_ADD128:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) {
            // (hi:lo) * m modulo 2^128
            const uint64_t plo = lo[r] * m[r];
            const uint64_t phi = __umul64hi(lo[r], m[r]) + hi[r] * m[r];
            add128<LDS>(base + ls[r] * WL, plo, phi);
          }
        break;
      case OP_ADDF64:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) add_f64<LDS>(base + ls[r] * WL, f[r] * (double)m[r]);
        break;
      case OP_MIN_I64:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) min_i64<LDS>(base + ls[r] * WL, (int64_t)lo[r]);
        break;
      case OP_MAX_I64:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) max_i64<LDS>(base + ls[r] * WL, (int64_t)lo[r]);
        break;
      case OP_MIN_F:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) min_u64<LDS>(base + ls[r] * WL, f64_order(f[r]));
        break;
      case OP_MAX_F:
#pragma unroll
        for (int r = 0; r < R; r++)
          if (ok[r]) max_u64<LDS>(base + ls[r] * WL, f64_order(f[r]));
        break;
      default: break;
    }
  }
}

// Merges a state record (W words) into an LDS record (LDS atomics).
__device__ __forceinline__ void merge_into_lds(const AggProg &P, uint64_t *d, const uint64_t *s) {
  for (int o = 0; o < P.nops; o++) {
    const AggOp op = op_at(P, o);
    const uint64_t *a = s + op.word;
    uint64_t *b = d + op.word;
    switch (op.kind) {
      case OP_COUNT_STAR:
      case OP_NULLS:
      case OP_ADDV:
      case OP_ADD64:
      case OP_COUNT_TRUE:
        if (a[0]) add_u64_nr<true>(b, a[0]);
        break;
      case OP_STICKY:
        if (s[op.st_word] == 2) max_u64<true>(d + op.st_word, 2);
        break;
      case OP_ADD128:
        if (!(op.checked && P.skip_checked)) add128<true>(b, a[0], a[1]);
        break;
      case OP_ADD192:
        if (!(op.checked && P.skip_checked)) add192<true>(b, a[0], a[1], a[2]);
        break;
      case OP_ADDF64: add_f64<true>(b, __longlong_as_double((long long)a[0])); break;
      case OP_MIN_I64:
        if (a[0] != P.init[op.word]) min_i64<true>(b, (int64_t)a[0]);
        break;
      case OP_MAX_I64:
        if (a[0] != P.init[op.word]) max_i64<true>(b, (int64_t)a[0]);
        break;
      case OP_MIN_F:
        if (a[0] != P.init[op.word]) min_u64<true>(b, a[0]);
        break;
      case OP_MAX_F:
        if (a[0] != P.init[op.word]) max_u64<true>(b, a[0]);
        break;
      default: break;
    }
  }
}

// Merges record s into record d with plain loads and stores: d belongs to
// the calling lane alone (its key's owner).
__device__ __forceinline__ void merge_plain(const AggProg &P, uint64_t *d, const uint64_t *s) {
  for (int o = 0; o < P.nops; o++) {
    const AggOp op = op_at(P, o);
    const uint64_t *a = s + op.word;
    uint64_t *b = d + op.word;
    switch (op.kind) {
      case OP_COUNT_STAR:
      case OP_NULLS:
      case OP_ADDV:
      case OP_ADD64:
      case OP_COUNT_TRUE: b[0] += a[0]; break;
      case OP_STICKY:
        if (s[op.st_word] == 2) d[op.st_word] = 2;
        break;
      case OP_ADD128:
        if (!(op.checked && P.skip_checked)) {
          uint64_t l = b[0] + a[0];
          b[1] += a[1] + (l < b[0] ? 1 : 0);
          b[0] = l;
        }
        break;
      case OP_ADD192:
        if (!(op.checked && P.skip_checked)) {
          unsigned __int128 x = (unsigned __int128)b[0] + a[0];
          unsigned __int128 y = (unsigned __int128)b[1] + a[1] + (uint64_t)(x >> 64);
          b[0] = (uint64_t)x;
          b[1] = (uint64_t)y;
          b[2] += a[2] + (uint64_t)(y >> 64);
        }
        break;
      case OP_ADDF64:
        b[0] = (uint64_t)__double_as_longlong(__longlong_as_double((long long)b[0]) +
                                              __longlong_as_double((long long)a[0]));
        break;
      case OP_MIN_I64:
        if ((int64_t)a[0] < (int64_t)b[0]) b[0] = a[0];
        break;
      case OP_MAX_I64:
        if ((int64_t)a[0] > (int64_t)b[0]) b[0] = a[0];
        break;
      case OP_MIN_F:
        if (a[0] < b[0]) b[0] = a[0];
        break;
      case OP_MAX_F:
        if (a[0] > b[0]) b[0] = a[0];
        break;
      default: break;
    }
  }
}

}  // namespace kg
