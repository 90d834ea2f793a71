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


#include <cmath>
#include <cstring>
#include <memory>

#include "kg_aggops.hpp"
#include "kg_scan.hpp"
#include "kg_aggtile.hpp"
#include "kg_gather.hpp"

namespace kg {
// ---- few groups: a one-byte key or a global aggregate ----------------------
// The used slots of a one-byte key (256 values and NULL at 256), as bits.
__global__ void k_small_slots(AggProg P, int64_t nrows, unsigned int *bits) {
  __shared__ unsigned int s_bits[9];
  if (threadIdx.x < 9) s_bits[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrows;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t key;
    const int cls = row_key(P, i, key);
    const int slot = cls == 2 ? 256 : (int)(key & 0xFF);
    if (!((s_bits[slot >> 5] >> (slot & 31)) & 1)) atomicOr(&s_bits[slot >> 5], 1u << (slot & 31));
  }
  __syncthreads();
  if (threadIdx.x < 9 && s_bits[threadIdx.x]) atomicOr(&bits[threadIdx.x], s_bits[threadIdx.x]);
}

// Each workgroup keeps K copies of the D used records in LDS; lane l of
// wave w writes copy (w * 4 + l % 4) % K, so LDS atomics of one instruction
// meet on a record only when their rows share a key and a copy.  Rows are
// applied op-major in tiles of R a lane (tile_apply); at the end the copies
// are summed into copy 0 in LDS and each used record is merged into its
// global record once.  remap: one-byte slot -> compact record (-1 unused);
// slot_of: compact record -> slot.  A global aggregate has D = 1.
constexpr int kSmallBlock = 256;
template <int R>
__global__ void __launch_bounds__(kSmallBlock) k_agg_small(AggProg P, Table T, int64_t nrows,
                                                           int64_t chunk, const int16_t *remap,
                                                           const int16_t *slot_of, int D,
                                                           int K) {
  extern __shared__ uint64_t smem[];
  __shared__ int16_t s_map[257];
  const int WL = P.W | 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  fill_lds_records(P, smem, K * D, WL);
  if (P.direct)
    for (int t = threadIdx.x; t < 257; t += blockDim.x) s_map[t] = remap[t];
  __syncthreads();
  uint64_t *mine = smem + (size_t)((wave * 4 + (lane & 3)) % K) * D * WL;
  const int64_t base = blockIdx.x * chunk;
  const int64_t end = base + chunk < nrows ? base + chunk : nrows;
  for (int64_t t0 = base; t0 < end; t0 += (int64_t)R * kSmallBlock) {
    int ls[R];
    ColSrc<R> src;
    src.safe = base;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t i = t0 + (int64_t)r * kSmallBlock + threadIdx.x;
      src.idx[r] = i;
      ls[r] = -1;
      if (i < end) {
        if (P.direct) {
          uint64_t key;
          const int cls = row_key(P, i, key);
          ls[r] = s_map[cls == 2 ? 256 : (int)(key & 0xFF)];
        } else {
          ls[r] = 0;
        }
      }
    }
    tile_apply<R>(P, mine, WL, ls, src);
  }
  __syncthreads();
  // copies 1..K-1 into copy 0 (one thread a record of a copy), then out
  for (int k = 1; k < K; k++) {
    for (int d = threadIdx.x; d < D; d += blockDim.x)
      merge_record<true>(P, smem + (size_t)d * WL, smem + ((size_t)k * D + d) * WL);
    __syncthreads();
  }
  for (int d = threadIdx.x; d < D; d += blockDim.x) {
    const uint64_t *l = smem + (size_t)d * WL;
    if (l[0] == 0) continue;  // no row
    const int slot = P.direct ? slot_of[d] : 0;
    merge_record(P, global_record(P, T, (P.direct && slot == 256) ? 2 : 0, (uint64_t)slot), l);
  }
}

// Few groups whose records are too wide for LDS copies (more than 64 KiB
// for every used slot): each row applied to its global record with atomics.
__global__ void __launch_bounds__(kAggBlock) k_agg_rows_atomic(AggProg P, Table T, int64_t nrows) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nrows;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t key;
    const int cls = row_key(P, i, key);
    apply_row<false>(P, global_record(P, T, cls, key), i);
  }
}

// ---------------------------------------------------------------- hashed keys
// Pass 1 (k_agg_pass1): each workgroup takes a contiguous chunk and applies
// it tile by tile (R rows a lane, op-major) to an LDS table of SLOTS records
// plus two records for the key INT64_MIN and the NULL key.  Rows whose key
// finds no LDS slot within kLdsProbe probes are appended to the row spill.
// A key takes an LDS slot only at its second sighting within an eviction
// window: a 16K-bit LDS doorkeeper records first sightings (one atomicOr),
// whose rows spill as 40-byte rows instead of becoming records that the next
// eviction would spill at 8 x (W + 1) bytes.  Every kEvictRows rows the
// records with fewer than kHotRows rows (the Zipf tail) move to the record
// spill and free their slots, and the doorkeeper is cleared, so the LDS
// table keeps the heavy keys; at the end every record moves there.  The
// global table is not touched by ordinary keys.
//
// Spill space: each wave takes kWaveChunk rows of the row spill at a time
// with one atomic; rows it leaves unused are marked dead (bit 63 of the
// validity word) and skipped by the partition pass.
//
// Drain: both spills are partitioned by home-slot range (k_spill_hist,
// k_spill_scatter); k_agg_merge gives each partition to one workgroup, which
// aggregates its inputs in LDS and, when the LDS table fills or the
// partition ends, merges each LDS record into the global record of its key
// with plain loads and stores: every key's global record is written by its
// partition's workgroup alone.  New keys take a slot with one CAS each,
// within a per-partition quota that keeps the table below 7/8 full; a
// workgroup that would pass it stops, leaving its unmerged records and
// pending inputs for the host, which grows the table, merges them with
// atomics (k_agg_retry) and resumes the partition.
constexpr int kP1Block = 1024;
constexpr int kMergeBlock = 1024;  // 16 waves a CU: latency of the spill reads hidden
constexpr int kMergeRows = 2;      // rows a lane per tile (96 VGPRs at 1024 lanes)
constexpr int kWMergeRows = 3;     // rows a lane per tile of the windowed merge
constexpr int kTileRows = 4;
constexpr int kWaveChunk = 512;
constexpr int64_t kEvictRows = 65536;
constexpr uint64_t kHotRows = 2;
constexpr uint64_t kDeadRow = 1ull << 63;

struct SpillDev {
  uint64_t *rows;
  unsigned long long *nrows;
  uint64_t cap_rows;
  uint64_t *recs;
  unsigned long long *nrecs;
  uint64_t cap_recs;
  unsigned int *overflow;
  SpillFmt F;
};

// A HyperLogLog sketch of the spilled keys (Flajolet et al., 2007), counted
// by the drain's partition histogram: register j = top kHllBits bits of an
// independent 64-bit hash of the key keeps the largest rank (1 + leading
// zeros) of the hash's remaining bits.  The drain sizes the table from its
// estimate before the merge (standard error 1.04 / sqrt(kHllRegs), about
// 3 %).
constexpr int kHllBits = 10;
constexpr int kHllRegs = 1 << kHllBits;
// A key's sketch update from its slot hash h (the hash that also places it
// in the table and the spill's partitions): the register from the low
// kHllBits bits, the rank from the leading zeros of the other 54.
__device__ __forceinline__ void hll_add(unsigned int *reg, uint64_t h) {
  const uint32_t j = (uint32_t)(h & (kHllRegs - 1));
  const uint64_t w = h >> kHllBits;
  const unsigned int rho = w ? (unsigned int)(__clzll(w) - kHllBits + 1) : 64u - kHllBits + 1u;
  if (rho > reg[j]) atomicMax(&reg[j], rho);
}

// Slots for the lanes in `mask` (every lane of the wave calls it with the
// same mask): one atomic per wave; returns this lane's index.
__device__ __forceinline__ int64_t spill_reserve(unsigned long long *counter, uint64_t mask) {
  const int lane = __lane_id();
  const int lead = __ffsll((unsigned long long)mask) - 1;
  unsigned long long base = 0;
  if (lane == lead) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
  base = __shfl(base, lead);
  return (int64_t)base + __popcll(mask & ((1ull << lane) - 1));
}

// A wave's current piece of the row spill: rows [base + used, base + kWaveChunk).
struct WaveChunk {
  int64_t base = -1;
  int used = kWaveChunk;
};

// Marks the unused rows of the wave's piece dead (every lane calls it).
__device__ __forceinline__ void chunk_retire(const SpillDev &S, WaveChunk &wc) {
  if (wc.base < 0) return;
  for (int j = wc.used + __lane_id(); j < kWaveChunk; j += 64) {
    const int64_t pos = wc.base + j;
    if ((uint64_t)pos < S.cap_rows) S.rows[pos * (int64_t)S.F.rw + 1] = kDeadRow;
  }
  wc.base = -1;
}

// Row-spill positions for the lanes in `mask` (every lane calls it).
__device__ __forceinline__ int64_t chunk_take(const SpillDev &S, WaveChunk &wc, uint64_t mask) {
  const int k = __popcll(mask);
  if (wc.used + k > kWaveChunk) {
    chunk_retire(S, wc);
    unsigned long long b = 0;
    if (__lane_id() == 0) b = atomicAdd(S.nrows, (unsigned long long)kWaveChunk);
    wc.base = (int64_t)__shfl(b, 0);
    wc.used = 0;
  }
  const int64_t pos = wc.base + wc.used + __popcll(mask & ((1ull << __lane_id()) - 1));
  wc.used += k;
  return pos;
}

// Writes the spill rows of the lane's rows with miss[r] (pos[r] their row):
// key, validity bits, then the column values (one type dispatch a column).
template <int R>
__device__ __forceinline__ void write_spill_rows(const AggProg &P, const SpillDev &S,
                                                 const ColSrc<R> &src, const bool (&miss)[R],
                                                 const uint64_t (&key)[R],
                                                 const int64_t (&pos)[R]) {
  const int rw = S.F.rw;
  uint64_t vbits[R];
#pragma unroll
  for (int r = 0; r < R; r++) vbits[r] = 0;
  uint64_t lo[R], hi[R];
  double f[R];
  bool ok[R];
  for (int c = 0; c < S.F.ncols; c++) {
    src.load(P, c, miss, lo, hi, f, ok);
    const bool fl = dev::is_f(P.in[c].type), wide = P.in[c].type == KG_T_DECIMAL128;
    const bool packed = S.F.packed == c + 1;
    const int o = S.F.off[c];
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (!miss[r]) continue;
      if (ok[r]) vbits[r] |= 1ull << (S.F.vshift + c);
      if (packed) {
        if (ok[r]) vbits[r] |= (uint32_t)lo[r];
        continue;
      }
      uint64_t *w = S.rows + pos[r] * rw + o;
      w[0] = !ok[r] ? 0 : fl ? (uint64_t)__double_as_longlong(f[r]) : lo[r];
      if (wide) w[1] = ok[r] ? hi[r] : 0;
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++)
    if (miss[r]) {
      S.rows[pos[r] * rw] = key[r];
      S.rows[pos[r] * rw + 1] = vbits[r];
    }
}

// LDS home slot of a key among `slots` (low 32 bits of the slot hash; the
// global home uses the top bits).
__device__ __forceinline__ uint32_t lds_home(const Table &T, uint64_t key, uint32_t slots) {
  return (uint32_t)(((slot_hash(T, key) & 0xFFFFFFFFull) * slots) >> 32);
}

// The doorkeeper: kDoorWords 32-bit words of LDS (16K bits), bit from the
// key's multiplicative hash.
constexpr int kDoorWords = 512;

__device__ __forceinline__ int lds_probe(uint64_t *s_keys, uint32_t slots, const Table &T,
                                         uint64_t key, uint32_t *s_door) {
  uint32_t s = lds_home(T, key, slots);
  for (int p = 0; p < kLdsProbe; p++) {
    uint64_t k = s_keys[s];
    if (k == key) return (int)s;
    if (k == kEmpty) {
      // a free slot: only a key already seen in this window takes it
      const uint32_t b = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 50);
      const uint32_t old = atomicOr(&s_door[b >> 5], 1u << (b & 31));
      if (!((old >> (b & 31)) & 1)) return -1;  // first sighting: the row spills
      unsigned long long prev = atomicCAS((unsigned long long *)&s_keys[s], kEmpty, key);
      if (prev == kEmpty || prev == key) return (int)s;
    }
    s = s + 1 == slots ? 0 : s + 1;
  }
  return -1;
}

// Moves LDS records to the record spill and clears their slots: all of them
// (all = true) or those with fewer than kHotRows rows.  Every thread calls it.
__device__ __forceinline__ void spill_lds_records(const AggProg &P, const SpillDev &S,
                                                  uint64_t *s_keys, uint64_t *s_rec, int SLOTS,
                                                  bool all) {
  const int W = P.W, WL = W | 1;
  __syncthreads();
  for (int t0 = 0; t0 < SLOTS; t0 += blockDim.x) {
    const int t = t0 + threadIdx.x;
    bool has = false;
    if (t < SLOTS && s_keys[t] != kEmpty) {
      const uint64_t rows = s_rec[t * WL];
      has = rows != 0 && (all || rows < kHotRows);
    }
    const uint64_t m = __ballot(has);
    if (!m) continue;
    const int64_t pos = spill_reserve(S.nrecs, m);
    if (has) {
      uint64_t *l = s_rec + t * WL;
      if ((uint64_t)pos < S.cap_recs) {
        uint64_t *d = S.recs + pos * (int64_t)(1 + W);
        d[0] = s_keys[t];
        for (int w = 0; w < W; w++) d[1 + w] = l[w];
      } else {
        atomicOr(S.overflow, 2u);
      }
      s_keys[t] = kEmpty;
      for (int w = 0; w < W; w++) l[w] = P.init[w];
    }
  }
  __syncthreads();
}

template <int R>
__global__ void __launch_bounds__(kP1Block, 4) k_agg_pass1(AggProg P, Table T, SpillDev S,
                                                        int64_t nrows, int64_t chunk, int SLOTS) {
  extern __shared__ uint64_t smem[];
  uint64_t *s_keys = smem;
  uint64_t *s_rec = smem + SLOTS;  // SLOTS + 2 records
  const int W = P.W;
  const int WL = W | 1;  // odd LDS record stride: records spread over the banks
  uint32_t *s_door = reinterpret_cast<uint32_t *>(s_rec + (size_t)(SLOTS + 2) * WL);
  for (int t = threadIdx.x; t < kDoorWords; t += blockDim.x) s_door[t] = 0;
  for (int t = threadIdx.x; t < SLOTS; t += blockDim.x) s_keys[t] = kEmpty;
  fill_lds_records(P, s_rec, SLOTS + 2, WL);
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < nrows ? lo + chunk : nrows;
  WaveChunk wc;
  int64_t next_evict = lo + kEvictRows;
  const int kt = P.key.type;
  const bool k64 = P.has_key && (kt == KG_T_INT64 || kt == KG_T_UINT64 || kt == KG_T_TIMESTAMP ||
                                 kt == KG_T_DECIMAL64);
  for (int64_t base = lo; base < hi; base += (int64_t)R * blockDim.x) {
    if (base >= next_evict) {
      spill_lds_records(P, S, s_keys, s_rec, SLOTS, false);
      for (int t = threadIdx.x; t < kDoorWords; t += blockDim.x) s_door[t] = 0;
      __syncthreads();
      next_evict = base + kEvictRows;
    }
    int ls[R];
    bool miss[R];
    uint64_t key[R];
    int64_t pos[R];
    // (no `safe` row: the branch-free column loads need more registers
    // than this kernel's 128 and would spill)
    ColSrc<R> src;
    // one-word integer keys: every row's key and validity loads are issued
    // before the first probe (clamped row index, the validity byte kept raw)
    uint32_t kvb[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      key[r] = 0;
      kvb[r] = 0;
    }
    if (k64) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t i = base + (int64_t)r * blockDim.x + threadIdx.x, ic = i < hi ? i : hi - 1;
        key[r] = dev::load_w64(P.key, ic);
        kvb[r] = P.key.validity ? P.key.validity[(ic + P.key.offset) >> 3] : 0xFFu;
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t i = base + (int64_t)r * blockDim.x + threadIdx.x;
      src.idx[r] = i;
      ls[r] = -1;
      miss[r] = false;
      pos[r] = 0;
      if (i < hi) {
        int cls;
        if (k64) {
          const bool ok = (kvb[r] >> ((i + P.key.offset) & 7)) & 1;
          cls = !ok ? 2 : key[r] == kEmpty ? 1 : 0;
          if (!ok) key[r] = 0;
        } else {
          cls = row_key(P, i, key[r]);
        }
        if (cls) {
          ls[r] = SLOTS + cls - 1;
        } else {
          ls[r] = lds_probe(s_keys, (uint32_t)SLOTS, T, key[r], s_door);
          miss[r] = ls[r] < 0;
        }
      }
    }
    tile_apply<R>(P, s_rec, WL, ls, src);
    bool any = false;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const uint64_t m = __ballot(miss[r]);
      if (!m) continue;
      pos[r] = chunk_take(S, wc, m);
      if (miss[r] && (uint64_t)pos[r] >= S.cap_rows) {
        atomicOr(S.overflow, 1u);
        miss[r] = false;
      }
      any = true;
    }
    if (any) write_spill_rows<R>(P, S, src, miss, key, pos);
  }
  chunk_retire(S, wc);
  spill_lds_records(P, S, s_keys, s_rec, SLOTS, true);
  // the key INT64_MIN and the NULL key: their fixed slots, with atomics
  if (threadIdx.x < 2) {
    const uint64_t *l = s_rec + (SLOTS + threadIdx.x) * WL;
    if (l[0]) merge_record(P, T.rec + (T.cap + threadIdx.x) * (uint64_t)W, l);
  }
}

// Partition of a spilled key: its home slot's range, top log2p bits of the
// slot hash (home slot = top log2cap bits).
__device__ __forceinline__ uint32_t spill_part(const Table &T, uint64_t key, int log2p) {
  return log2p ? (uint32_t)(slot_hash(T, key) >> (64 - log2p)) : 0u;
}

// Per-block partition histogram: hist[p * gridDim + b] (dead rows of the row
// spill, `dead` set, are left out).
// With hll, the entries' keys also go into the sketch of the spilled keys.
__global__ void k_spill_hist(const uint64_t *rows, int w, uint64_t n, Table T, int log2p,
                             int64_t per_block, int dead, uint32_t *hist, unsigned int *hll) {
  extern __shared__ uint32_t s_h[];
  const uint32_t np = 1u << log2p;
  unsigned int *s_reg = s_h + np;
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) s_h[p] = 0;
  if (hll)
    for (int j = threadIdx.x; j < kHllRegs; j += blockDim.x) s_reg[j] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = lo + per_block < (int64_t)n ? lo + per_block : (int64_t)n;
  // four rows a thread: their loads are in flight together
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += 4 * (int64_t)blockDim.x) {
    uint64_t key[4], fl[4];
    bool live[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {  // clamped row index: no branch between the loads
      const int64_t i = i0 + r * (int64_t)blockDim.x, ic = i < hi ? i : hi - 1;
      key[r] = rows[ic * w];
      fl[r] = dead ? rows[ic * w + 1] : 0;
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
      live[r] = i0 + r * (int64_t)blockDim.x < hi && !(fl[r] & kDeadRow);
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (live[r]) {
        const uint64_t h = slot_hash(T, key[r]);
        atomicAdd(&s_h[log2p ? (uint32_t)(h >> (64 - log2p)) : 0u], 1u);
        if (hll) hll_add(s_reg, h);
      }
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) hist[(uint64_t)p * gridDim.x + blockIdx.x] = s_h[p];
  if (hll)
    for (int j = threadIdx.x; j < kHllRegs; j += blockDim.x) {
      const unsigned int v = s_reg[j];
      if (v && v > hll[j]) atomicMax(&hll[j], v);
    }
}

// Moves each row to its partition (order within a partition is not kept).
// Rows of up to kRegWords words are read into registers first, so every load
// of a thread's four rows is in flight before the cursor atomics.
__global__ void k_spill_scatter(const uint64_t *rows, int w, uint64_t n, Table T, int log2p,
                                int64_t per_block, int dead, const uint64_t *base, uint64_t *out) {
  extern __shared__ uint32_t s_c[];
  constexpr int kRegWords = 6;
  const uint32_t np = 1u << log2p;
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x)
    s_c[p] = (uint32_t)base[(uint64_t)p * gridDim.x + blockIdx.x];
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = lo + per_block < (int64_t)n ? lo + per_block : (int64_t)n;
  const bool regs = w <= kRegWords;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += 4 * (int64_t)blockDim.x) {
    uint64_t v[4][kRegWords];
    bool live[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int64_t i = i0 + r * (int64_t)blockDim.x;
      const bool in = i < hi;
#pragma unroll
      for (int k = 0; k < kRegWords; k++)
        v[r][k] = (in && (k < w) && (regs || k < 2)) ? rows[i * w + k] : 0;
      live[r] = in && !(dead && (v[r][1] & kDeadRow));
    }
    uint32_t pos[4];
#pragma unroll
    for (int r = 0; r < 4; r++)
      pos[r] = live[r] ? atomicAdd(&s_c[spill_part(T, v[r][0], log2p)], 1u) : 0u;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (!live[r]) continue;
      uint64_t *dst = out + (uint64_t)pos[r] * w;
      if (regs) {
#pragma unroll
        for (int k = 0; k < kRegWords; k++)
          if (k < w) dst[k] = v[r][k];
      } else {
        const uint64_t *src = rows + (i0 + r * (int64_t)blockDim.x) * w;
        for (int k = 0; k < w; k++) dst[k] = src[k];
      }
    }
  }
}

struct MergeDev {
  const uint64_t *rows;   // partitioned row spill
  uint64_t nrows;
  const uint64_t *rbase;  // scanned histogram: partition p starts at rbase[p * rnb]
  uint32_t rnb;
  const uint64_t *recs;   // partitioned record spill
  uint64_t nrecs;
  const uint64_t *cbase;
  uint32_t cnb;
  int log2p;
  int64_t *resume;        // per partition: -1 merged, else the next input
  uint32_t *inserted;     // per partition: keys inserted by this launch
  uint32_t quota;         // keys a partition may insert in this launch
  uint64_t *retry;        // per workgroup: up to S records (key + W words)
  uint32_t *retry_n;
  int64_t *pend;          // per workgroup: inputs of the stopped tile not applied
  uint32_t *pend_n;
  uint32_t *pend_part;
  unsigned int *stops;
  SpillFmt F;
};

__device__ __forceinline__ void part_range(const uint64_t *base, uint32_t nb, uint64_t n,
                                           uint32_t p, uint32_t np, int64_t &a, int64_t &b) {
  if (!base || n == 0) {
    a = b = 0;
    return;
  }
  a = (int64_t)base[(uint64_t)p * nb];
  b = p + 1 < np ? (int64_t)base[(uint64_t)(p + 1) * nb] : (int64_t)n;
}

// Key slot in the global table, or -1 (relaxed device-scope loads: keys are
// inserted by other workgroups while this one probes).
__device__ __forceinline__ int64_t table_find(const Table &T, uint64_t key) {
  uint64_t s = slot_hash(T, key) >> (64 - T.log2cap);
  for (uint64_t p = 0; p <= T.wmask; p++) {
    uint64_t k = __hip_atomic_load(&T.keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == key) return (int64_t)s;
    if (k == kEmpty) return -1;
    s = next_in_window(s, T.wmask);
  }
  return -1;
}

// Find-or-insert without counting (the caller accounts for new keys).  A full
// window flags the table (see global_slot).
__device__ __forceinline__ uint64_t table_insert(const Table &T, uint64_t key) {
  uint64_t s = slot_hash(T, key) >> (64 - T.log2cap);
  for (uint64_t p = 0; p <= T.wmask; p++) {
    uint64_t k = __hip_atomic_load(&T.keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == key) return s;
    if (k == kEmpty) {
      unsigned long long prev = atomicCAS((unsigned long long *)&T.keys[s], kEmpty, key);
      if (prev == kEmpty || prev == key) return s;
    }
    s = next_in_window(s, T.wmask);
  }
  atomicOr(T.occupied + 1, 1ull);
  return T.cap + 1;
}

// Insert into the merge kernel's LDS table, unbounded probing while fewer
// than `limit` keys are held (a full table fails the insert).
__device__ __forceinline__ int lds_insert(uint64_t *s_keys, uint32_t S, int *s_occ, int limit,
                                          const Table &T, uint64_t key) {
  uint32_t s = lds_home(T, key, S);
  while (true) {
    uint64_t k = s_keys[s];
    if (k == key) return (int)s;
    if (k == kEmpty) {
      if (atomicAdd(s_occ, 1) >= limit) {
        atomicSub(s_occ, 1);
        // a key another lane is inserting may still land here: look once more
        return s_keys[s] == key ? (int)s : -1;
      }
      unsigned long long prev = atomicCAS((unsigned long long *)&s_keys[s], kEmpty, key);
      if (prev == kEmpty) return (int)s;
      atomicSub(s_occ, 1);
      if (prev == key) return (int)s;
    }
    s = s + 1 == S ? 0 : s + 1;
  }
}

// Merges every LDS record into the global table and clears the LDS table.
// Returns false when the partition's quota of new keys would be passed: the
// records of new keys then go to the workgroup's retry area.
__device__ bool merge_flush(const AggProg &P, const Table &T, const MergeDev &M, int S,
                            uint64_t *s_keys, uint64_t *s_rec, int *s_occ, uint32_t *s_new,
                            uint32_t *s_ins, int *s_fail) {
  const int W = P.W, WL = W | 1;
  if (threadIdx.x == 0) *s_new = 0;
  __syncthreads();
  for (int t = threadIdx.x; t < S; t += blockDim.x) {
    const uint64_t k = s_keys[t];
    if (k == kEmpty) continue;
    const int64_t g = table_find(T, k);
    if (g < 0) {
      atomicAdd(s_new, 1u);
      continue;
    }
    uint64_t *l = s_rec + t * WL;
    merge_plain(P, T.rec + (uint64_t)g * W, l);
    s_keys[t] = kEmpty;
    for (int w = 0; w < W; w++) l[w] = P.init[w];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    *s_fail = *s_ins + *s_new > M.quota;
    if (!*s_fail) *s_ins += *s_new;
  }
  __syncthreads();
  const bool fail = *s_fail;
  for (int t = threadIdx.x; t < S; t += blockDim.x) {
    const uint64_t k = s_keys[t];
    if (k == kEmpty) continue;
    uint64_t *l = s_rec + t * WL;
    if (!fail) {
      const uint64_t g = table_insert(T, k);
      merge_plain(P, T.rec + g * W, l);
    } else {
      const uint32_t j = atomicAdd(&M.retry_n[blockIdx.x], 1u);
      uint64_t *d = M.retry + ((uint64_t)blockIdx.x * S + j) * (1 + W);
      d[0] = k;
      for (int w = 0; w < W; w++) d[1 + w] = l[w];
    }
    s_keys[t] = kEmpty;
    for (int w = 0; w < W; w++) l[w] = P.init[w];
  }
  if (threadIdx.x == 0) *s_occ = 0;
  __syncthreads();
  return !fail;
}

template <int R>
__global__ void __launch_bounds__(kMergeBlock) k_agg_merge(AggProg P, Table T, MergeDev M, int S) {
  extern __shared__ uint64_t smem[];
  uint64_t *s_keys = smem;
  uint64_t *s_rec = smem + S;
  __shared__ int s_occ, s_fail;
  __shared__ uint32_t s_new, s_ins;
  const int W = P.W, WL = W | 1;
  const int limit = S - S / 8;
  const uint32_t np = 1u << M.log2p;
  for (int t = threadIdx.x; t < S; t += blockDim.x) s_keys[t] = kEmpty;
  fill_lds_records(P, s_rec, S, WL);
  if (threadIdx.x == 0) s_occ = 0;
  __syncthreads();
  for (uint32_t p = blockIdx.x; p < np; p += gridDim.x) {
    const int64_t start = M.resume[p];
    if (start < 0) continue;
    int64_t r0, r1, c0, c1;
    part_range(M.rbase, M.rnb, M.nrows, p, np, r0, r1);
    part_range(M.cbase, M.cnb, M.nrecs, p, np, c0, c1);
    const int64_t nrec = c1 - c0, total = nrec + (r1 - r0);
    if (threadIdx.x == 0) s_ins = 0;
    __syncthreads();
    bool stopped = false;
    for (int64_t t0 = start; t0 < total && !stopped; t0 += (int64_t)R * blockDim.x) {
      int64_t u[R];
      bool done[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        u[r] = t0 + (int64_t)r * blockDim.x + threadIdx.x;
        done[r] = u[r] >= total;
      }
      while (true) {
        int ls[R], lsw[R];
        SpillSrc<R> src;
        src.rows = M.rows;
        src.safe = r1 > r0 ? r0 : -1;
        // every row's key load is issued before the first insert (a row
        // index clamped into the partition, so no per-lane branch)
        uint64_t key[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int64_t uc = u[r] < total ? u[r] : total - 1;
          const uint64_t *kp = uc < nrec ? M.recs + (c0 + uc) * (1 + W)
                                         : M.rows + (r0 + uc - nrec) * M.F.rw;
          key[r] = *kp;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
          ls[r] = -1;
          src.idx[r] = r0 + u[r] - nrec;
          if (!done[r]) ls[r] = lds_insert(s_keys, (uint32_t)S, &s_occ, limit, T, key[r]);
          if (ls[r] >= 0 && u[r] < nrec)
            merge_into_lds(P, s_rec + ls[r] * WL, M.recs + (c0 + u[r]) * (1 + W) + 1);
          lsw[r] = (ls[r] >= 0 && u[r] >= nrec) ? ls[r] : -1;
        }
        tile_apply<R>(P, s_rec, WL, lsw, src);
        bool left = false;
#pragma unroll
        for (int r = 0; r < R; r++) {
          if (ls[r] >= 0) done[r] = true;
          left |= !done[r];
        }
        if (!__syncthreads_or(left)) break;
        // the LDS table is full: merge it into the global table, then retry
        if (!merge_flush(P, T, M, S, s_keys, s_rec, &s_occ, &s_new, &s_ins, &s_fail)) {
#pragma unroll
          for (int r = 0; r < R; r++)
            if (!done[r]) {
              const uint32_t j = atomicAdd(&M.pend_n[blockIdx.x], 1u);
              M.pend[(uint64_t)blockIdx.x * kMergeRows * kMergeBlock + j] = u[r];
            }
          if (threadIdx.x == 0) {
            M.pend_part[blockIdx.x] = p;
            M.resume[p] = t0 + (int64_t)R * blockDim.x;
            M.inserted[p] = s_ins;
            atomicAdd(M.stops, 1u);
          }
          stopped = true;
          break;
        }
      }
    }
    if (stopped) break;
    const bool ok = merge_flush(P, T, M, S, s_keys, s_rec, &s_occ, &s_new, &s_ins, &s_fail);
    if (threadIdx.x == 0) {
      M.resume[p] = -1;
      M.inserted[p] = s_ins;
      if (!ok) {
        M.pend_part[blockIdx.x] = p;
        atomicAdd(M.stops, 1u);
      }
    }
    if (!ok) break;  // the retry area is in use: this workgroup stops
  }
}

// The stopped workgroups' leftovers, merged with atomics after the table has
// grown: retry records, then the pending inputs of the stopped tile.
__global__ void k_agg_retry(AggProg P, Table T, MergeDev M, int S, uint32_t nwg) {
  const int W = P.W;
  for (uint32_t g = blockIdx.x; g < nwg; g += gridDim.x) {
    const uint32_t nr = M.retry_n[g], npd = M.pend_n[g];
    for (uint32_t j = threadIdx.x; j < nr; j += blockDim.x) {
      const uint64_t *d = M.retry + ((uint64_t)g * S + j) * (1 + W);
      merge_record(P, T.rec + global_slot(T, d[0]) * (uint64_t)W, d + 1);
    }
    if (npd == 0) continue;
    const uint32_t p = M.pend_part[g];
    const uint32_t np = 1u << M.log2p;
    int64_t r0, r1, c0, c1;
    part_range(M.rbase, M.rnb, M.nrows, p, np, r0, r1);
    part_range(M.cbase, M.cnb, M.nrecs, p, np, c0, c1);
    const int64_t nrec = c1 - c0;
    for (uint32_t j = threadIdx.x; j < npd; j += blockDim.x) {
      const int64_t u = M.pend[(uint64_t)g * kMergeRows * kMergeBlock + j];
      if (u < nrec) {
        const uint64_t *d = M.recs + (c0 + u) * (1 + W);
        merge_record(P, T.rec + global_slot(T, d[0]) * (uint64_t)W, d + 1);
      } else {
        SpillSrc<1> src;
        src.rows = M.rows;
        src.idx[0] = r0 + u - nrec;
        uint64_t *r = T.rec + global_slot(T, M.rows[src.idx[0] * M.F.rw]) * (uint64_t)W;
        int ls[1] = {0};
        tile_apply<1, false>(P, r, 0, ls, src);
      }
    }
  }
}

// ---- windowed merge ---------------------------------------------------------
// One workgroup owns a window of the table (2^log2span slots: keys probe only
// within their window, so the window holds exactly its own keys): the window
// comes into LDS (its occupied slots loaded, or all empty while the table
// holds no key), every spilled input of the window is applied there (an LDS
// insert within the window, LDS atomics), and the window goes back with
// coalesced stores -- its keys, and the records of its occupied slots.  A
// partition of the spill covers 2^g windows, and its inputs are read once
// per window (those homed in another window are skipped).  A window that
// fills is not written back (the table keeps that window as it was) and is
// counted in *failed; the host grows the table and runs the merge again over
// the windows not yet done (done[w] = 1).
struct WMergeDev {
  const uint64_t *rows;
  uint64_t nrows;
  const uint64_t *rbase;
  uint32_t rnb;
  const uint64_t *recs;
  uint64_t nrecs;
  const uint64_t *cbase;
  uint32_t cnb;
  int log2p;     // spill partitions
  int g;         // log2 windows a partition
  int log2span;  // log2 slots a window
  int load;      // 1: windows are read from the table; 0: the table holds no key
  uint8_t *done;
  unsigned int *failed;
  SpillFmt F;
};

// Find-or-insert of key in an LDS window of `span` slots from slot s; -1
// when the window is full.
__device__ __forceinline__ int win_insert(uint64_t *s_keys, uint32_t span, uint32_t s,
                                          uint64_t key) {
  for (uint32_t p = 0; p < span; p++) {
    const uint64_t k = s_keys[s];
    if (k == key) return (int)s;
    if (k == kEmpty) {
      const unsigned long long prev = atomicCAS((unsigned long long *)&s_keys[s], kEmpty, key);
      if (prev == kEmpty || prev == key) return (int)s;
    }
    s = (s + 1) & (span - 1);
  }
  return -1;
}

// Spill rows of up to kStageWords words are staged in LDS a tile at a time.
constexpr int kStageWords = 8;

template <int R, bool STAGE>
__global__ void __launch_bounds__(kMergeBlock) k_agg_wmerge(AggProg P, Table T, WMergeDev M) {
  extern __shared__ uint64_t smem[];
  const uint32_t span = 1u << M.log2span;
  uint64_t *s_keys = smem;
  uint64_t *s_rec = smem + span;
  uint64_t *s_stage = s_rec + (size_t)span * (P.W | 1);  // STAGE: blockDim rows
  __shared__ int s_fail, s_keys_before, s_keys_after;
  const int W = P.W, WL = W | 1;
  const uint32_t np = 1u << M.log2p;
  const uint32_t nw = 1u << M.g;
  for (uint32_t p = blockIdx.x; p < np; p += gridDim.x) {
    int64_t r0, r1, c0, c1;
    part_range(M.rbase, M.rnb, M.nrows, p, np, r0, r1);
    part_range(M.cbase, M.cnb, M.nrecs, p, np, c0, c1);
    const int64_t nrec = c1 - c0, total = nrec + (r1 - r0);
    for (uint32_t wi = 0; wi < nw; wi++) {
      const uint64_t win = ((uint64_t)p << M.g) + wi;
      if (M.done[win]) continue;  // the same for every lane
      if (total == 0) {           // nothing changes in this window
        if (threadIdx.x == 0) M.done[win] = 1;
        continue;
      }
      const uint64_t s0 = win << M.log2span;
      // the window into LDS: keys, then each slot's record (its own when
      // occupied, the initial state otherwise)
      if (threadIdx.x == 0) {
        s_fail = 0;
        s_keys_before = 0;
        s_keys_after = 0;
      }
      for (uint32_t t = threadIdx.x; t < span; t += blockDim.x) {
        const uint64_t k = M.load ? T.keys[s0 + t] : kEmpty;
        s_keys[t] = k;
        const uint64_t m = __ballot(k != kEmpty);
        if (__lane_id() == 0 && m) atomicAdd(&s_keys_before, __popcll(m));
      }
      fill_lds_records(P, s_rec, (int)span, WL);
      __syncthreads();
      if (M.load) {  // the occupied slots' own records
        for (uint32_t i = threadIdx.x; i < span * (uint32_t)W; i += blockDim.x) {
          const uint32_t sl = i / (uint32_t)W, w = i - sl * (uint32_t)W;
          if (s_keys[sl] != kEmpty) s_rec[sl * WL + w] = T.rec[s0 * W + i];
        }
        __syncthreads();
      }
      // inputs [0, nrec) are pass 1's spilled records, [nrec, total) the
      // spilled rows; with STAGE the rows are taken a tile at a time from LDS
      const int64_t end = STAGE ? nrec : total;
      for (int64_t t0 = 0; t0 < end; t0 += (int64_t)R * blockDim.x) {
        if (*(volatile int *)&s_fail) break;  // the window runs again in a larger table
        int ls[R], lsw[R];
        uint64_t key[R];
        SpillSrc<R> src;
        src.rows = M.rows;
        src.safe = r1 > r0 ? r0 : -1;
        // every row's key load is issued before the first insert (a row
        // index clamped into the partition, so no per-lane branch)
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int64_t u = t0 + (int64_t)r * blockDim.x + threadIdx.x;
          const int64_t uc = u < end ? u : end - 1;
          const uint64_t *kp = uc < nrec ? M.recs + (c0 + uc) * (1 + W)
                                         : M.rows + (r0 + uc - nrec) * M.F.rw;
          key[r] = *kp;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int64_t u = t0 + (int64_t)r * blockDim.x + threadIdx.x;
          ls[r] = -1;
          src.idx[r] = r0 + u - nrec;
          if (u < end) {
            const uint64_t home = slot_hash(T, key[r]) >> (64 - T.log2cap);
            if ((home >> M.log2span) == win) {
              ls[r] = win_insert(s_keys, span, (uint32_t)(home & (span - 1)), key[r]);
              if (ls[r] < 0) s_fail = 1;
            }
          }
          if (ls[r] >= 0 && u < nrec)
            merge_into_lds(P, s_rec + ls[r] * WL, M.recs + (c0 + u) * (1 + W) + 1);
          lsw[r] = (ls[r] >= 0 && u >= nrec) ? ls[r] : -1;
        }
        tile_apply<R>(P, s_rec, WL, lsw, src);
      }
      if (STAGE) {
        // a tile of blockDim rows into LDS with coalesced loads (a lane's
        // words all in flight together), then applied from LDS
        const int rw = M.F.rw;
        for (int64_t t0 = r0; t0 < r1; t0 += blockDim.x) {
          // (also the barrier after the previous tile's reads of the stage)
          if (__syncthreads_or(*(volatile int *)&s_fail)) break;
          const int64_t nt = r1 - t0 < (int64_t)blockDim.x ? r1 - t0 : (int64_t)blockDim.x;
          const int64_t lim = nt * rw;
          const uint64_t *g = M.rows + t0 * rw;
          uint64_t v[kStageWords];
#pragma unroll
          for (int k = 0; k < kStageWords; k++) {
            const int64_t i = (int64_t)k * blockDim.x + threadIdx.x;
            v[k] = g[i < lim ? i : lim - 1];
          }
#pragma unroll
          for (int k = 0; k < kStageWords; k++) {
            const int64_t i = (int64_t)k * blockDim.x + threadIdx.x;
            if (i < lim) s_stage[i] = v[k];
          }
          __syncthreads();
          int ls[1] = {-1};
          if ((int64_t)threadIdx.x < nt) {
            const uint64_t key = s_stage[threadIdx.x * rw];
            const uint64_t home = slot_hash(T, key) >> (64 - T.log2cap);
            if ((home >> M.log2span) == win) {
              ls[0] = win_insert(s_keys, span, (uint32_t)(home & (span - 1)), key);
              if (ls[0] < 0) s_fail = 1;
            }
          }
          SpillSrc<1> src;
          src.rows = s_stage;
          src.idx[0] = threadIdx.x;
          src.safe = 0;
          tile_apply<1>(P, s_rec, WL, ls, src);
        }
      }
      __syncthreads();
      if (!s_fail) {
        for (uint32_t t = threadIdx.x; t < span; t += blockDim.x) {
          const uint64_t k = s_keys[t];
          T.keys[s0 + t] = k;
          const uint64_t m = __ballot(k != kEmpty);
          if (__lane_id() == 0 && m) atomicAdd(&s_keys_after, __popcll(m));
        }
        for (uint32_t i = threadIdx.x; i < span * (uint32_t)W; i += blockDim.x) {
          const uint32_t sl = i / (uint32_t)W, w = i - sl * (uint32_t)W;
          if (s_keys[sl] != kEmpty) T.rec[s0 * W + i] = s_rec[sl * WL + w];
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        if (s_fail) {
          atomicAdd(M.failed, 1u);
        } else {
          M.done[win] = 1;
          if (s_keys_after != s_keys_before)
            atomicAdd(T.occupied, (unsigned long long)(s_keys_after - s_keys_before));
        }
      }
      __syncthreads();  // the LDS window is refilled next
    }
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

__device__ __forceinline__ void store_dec(int type, void *values, int64_t o, uint64_t lo,
                                          uint64_t hi) {
  if (type == KG_T_DECIMAL128) {
    uint64_t *v = static_cast<uint64_t *>(values) + 2 * o;
    v[0] = lo;
    v[1] = hi;
  } else {
    static_cast<uint64_t *>(values)[o] = lo;
  }
}

__device__ __forceinline__ void store_int(int type, void *values, int64_t o, int64_t v) {
  switch (type) {
    case KG_T_INT8:
    case KG_T_UINT8:
    case KG_T_BOOL: static_cast<int8_t *>(values)[o] = (int8_t)v; break;
    case KG_T_INT16:
    case KG_T_UINT16: static_cast<int16_t *>(values)[o] = (int16_t)v; break;
    case KG_T_INT32:
    case KG_T_UINT32:
    case KG_T_DATE32: static_cast<int32_t *>(values)[o] = (int32_t)v; break;
    default: static_cast<int64_t *>(values)[o] = v;
  }
}

// The exact 128-bit sum rounded once to double.
__device__ __forceinline__ double i128_to_double(uint64_t lo, uint64_t hi) {
  bool n = neg128(lo, hi);
  if (n) negate128(lo, hi);
  double s = (double)(((unsigned __int128)hi << 64) | lo);
  return n ? -s : s;
}

// One lane per output row o (the compacted slot list[o]).  Validity words
// are whole 32-bit words from wave ballots (no atomics): lanes past the end
// redo the last row, whose stores repeat the same values; a wave wholly past
// the end stores nothing.
__device__ __forceinline__ void wave_bits(uint8_t *bm, int64_t obase, int64_t n, bool bit) {
  uint64_t m = __ballot(bit);
  const int lane = __lane_id();
  if (lane == 0 && obase < n) reinterpret_cast<uint32_t *>(bm)[obase >> 5] = (uint32_t)m;
  if (lane == 32 && obase + 32 < n) reinterpret_cast<uint32_t *>(bm)[(obase >> 5) + 1] = (uint32_t)(m >> 32);
}

__global__ void k_agg_output(OutProg O, const uint64_t *keys, int key_stride, const uint64_t *rec,
                             const uint64_t *list, int64_t n, void *out_key,
                             uint8_t *out_key_valid, int key_width, unsigned int *err) {
  for (int64_t o0 = blockIdx.x * (int64_t)blockDim.x; o0 < n; o0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t obase = o0 + (threadIdx.x & ~63);
    const bool active = o0 + threadIdx.x < n;
    const int64_t o = active ? o0 + threadIdx.x : n - 1;
    const uint64_t s = list[o];
    // the record into LDS with back-to-back loads (its lines are read once;
    // odd stride: lanes' records fall in different banks)
    extern __shared__ uint64_t s_out_rec[];
    uint64_t *r = s_out_rec + threadIdx.x * (uint64_t)(O.W | 1);
    {
      const uint64_t *g = rec + s * (uint64_t)O.W;
      for (int w = 0; w < O.W; w++) r[w] = g[w];
    }
    uint64_t rows = r[0];
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
    }
    if (out_key_valid) wave_bits(out_key_valid, obase, n, active && !null_key);
    for (int j = 0; j < O.n; j++) {
      const OutAgg &a = O.a[j];
      const uint64_t *w = r + a.word;
      uint64_t valid_rows = a.valid_word >= 0   ? r[a.valid_word]
                            : a.nulls_word >= 0 ? rows - r[a.nulls_word]
                                                : rows;
      uint64_t divisor = a.cnt_word >= 0 ? r[a.cnt_word] : valid_rows;
      bool overflowed = a.st_word >= 0 && r[a.st_word] == 2;
      bool ok = true;
      if (O.partial) {
        // the partial state columns (accumulator/output.rs:8-60)
        switch (a.kind) {
          case R_COUNT_STAR:
          case R_COUNT:
            static_cast<int64_t *>(a.values)[o] = (int64_t)(a.kind == R_COUNT_STAR ? (a.cnt_word >= 0 ? divisor : rows) : divisor);
            break;
          case R_SUM_INT:
            ok = valid_rows > 0;
            if (ok && overflowed) atomicOr(err, 1u);
            static_cast<int64_t *>(a.values)[o] = (int64_t)w[0];
            break;
          case R_SUM_F:
            ok = valid_rows > 0;
            static_cast<double *>(a.values)[o] = __longlong_as_double((long long)w[0]);
            break;
          case R_SUM_DEC:
            ok = valid_rows > 0 && !overflowed;
            if (ok) store_dec(a.st_type, a.values, o, w[0], w[1]);
            static_cast<uint8_t *>(a.values2)[o] = valid_rows > 0 ? 1 : 0;
            break;
          case R_AVG_INT:
            static_cast<double *>(a.values)[o] = i128_to_double(w[0], w[1]);
            static_cast<int64_t *>(a.values2)[o] = (int64_t)divisor;
            break;
          case R_AVG_F:
            static_cast<double *>(a.values)[o] = __longlong_as_double((long long)w[0]);
            static_cast<int64_t *>(a.values2)[o] = (int64_t)divisor;
            break;
          case R_AVG_DEC:
            ok = !overflowed;
            if (ok) store_dec(a.st_type, a.values, o, w[0], w[1]);
            static_cast<int64_t *>(a.values2)[o] = (int64_t)divisor;
            break;
          case R_MINMAX_I:
            ok = valid_rows > 0;
            store_int(a.out_type, a.values, o, (int64_t)w[0]);
            break;
          case R_MINMAX_F: {
            ok = valid_rows > 0;
            double d = f64_unorder(w[0]);
            if (a.out_type == KG_T_FLOAT32) static_cast<float *>(a.values)[o] = (float)d;
            else static_cast<double *>(a.values)[o] = d;
            break;
          }
        }
        wave_bits(a.validity, obase, n, active && ok);
        if (a.validity2) wave_bits(a.validity2, obase, n, active);
        continue;
      }
      switch (a.kind) {
        case R_COUNT_STAR:
          static_cast<int64_t *>(a.values)[o] = (int64_t)(a.cnt_word >= 0 ? divisor : rows);
          break;
        case R_COUNT: static_cast<int64_t *>(a.values)[o] = (int64_t)divisor; break;
        case R_SUM_INT: {
          ok = valid_rows > 0;
          // ANSI: the exact 128-bit sum must fit; otherwise the one wrapping
          // word is the result
          bool fits = !O.ansi || (int64_t)w[1] == ((int64_t)w[0] >> 63);
          if (ok && (overflowed || (O.ansi && !fits))) atomicOr(err, 1u);
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
          if (overflowed) {
            ok = false;
            if (O.ansi) atomicOr(err, 1u);
            break;
          }
          uint64_t lo = w[0], hi = w[1];
          bool wide_ok = !a.wide || (int64_t)w[2] == ((int64_t)w[1] >> 63);
          uint64_t mlo = lo, mhi = hi;
          if (neg128(lo, hi)) negate128(mlo, mhi);
          if (!(wide_ok && lt_mag(mlo, mhi, a.pow_lo, a.pow_hi))) {
            ok = false;
            if (O.ansi) atomicOr(err, 1u);
            break;
          }
          store_dec(a.out_type, a.values, o, lo, hi);
          break;
        }
        case R_MINMAX_I:
          ok = valid_rows > 0;
          store_int(a.out_type, a.values, o, (int64_t)w[0]);
          break;
        case R_MINMAX_F: {
          ok = valid_rows > 0;
          double d = f64_unorder(w[0]);
          if (a.out_type == KG_T_FLOAT32) static_cast<float *>(a.values)[o] = (float)d;
          else static_cast<double *>(a.values)[o] = d;
          break;
        }
        case R_AVG_INT:
          ok = divisor > 0;
          if (ok) static_cast<double *>(a.values)[o] = i128_to_double(w[0], w[1]) / (double)divisor;
          break;
        case R_AVG_F:
          ok = divisor > 0;
          if (ok)
            static_cast<double *>(a.values)[o] =
                __longlong_as_double((long long)w[0]) / (double)divisor;
          break;
        case R_AVG_DEC: {
          ok = divisor > 0;
          if (!ok) break;
          if (overflowed) {
            ok = false;
            if (O.ansi) atomicOr(err, 1u);
            break;
          }
          // avg_fast (accumulator/helpers.rs:595-609): |sum| * 10^(s - ss)
          // / count, HALF_UP; the digit-loop value where that overflows.
          uint64_t lo = w[0], hi = w[1];
          bool negative = neg128(lo, hi);
          if (negative) negate128(lo, hi);
          int k = a.out_s - a.in_s;
          bool ovf = false;
          for (int t = 0; t < k && !ovf; t++) ovf = mul128(lo, hi, 10);
          uint64_t qlo, qhi, rem;
          if (!ovf) {
            divmod128(lo, hi, divisor, qlo, qhi, rem);
            if (rem >= divisor - rem) {
              qlo += 1;
              if (qlo == 0) qhi += 1;
            }
          } else {
            uint64_t slo = w[0], shi = w[1];
            if (negative) negate128(slo, shi);
            divmod128(slo, shi, divisor, qlo, qhi, rem);
            ovf = false;
            for (int t = 0; t < k; t++) {
              if (mul128(qlo, qhi, 10)) {
                ovf = true;
                break;
              }
              unsigned __int128 r10 = (unsigned __int128)rem * 10;
              uint64_t dgt = (uint64_t)(r10 / divisor);
              rem = (uint64_t)(r10 % divisor);
              uint64_t old = qlo;
              qlo += dgt;
              if (qlo < old) qhi += 1;
            }
            if (!ovf) {
              unsigned __int128 r10 = (unsigned __int128)rem * 10;
              if ((uint64_t)(r10 / divisor) >= 5) {
                qlo += 1;
                if (qlo == 0) qhi += 1;
              }
            }
          }
          if (ovf || !lt_mag(qlo, qhi, a.pow_lo, a.pow_hi)) {
            ok = false;
            if (O.ansi) atomicOr(err, 2u);
            break;
          }
          if (negative) negate128(qlo, qhi);
          store_dec(a.out_type, a.values, o, qlo, qhi);
          break;
        }
      }
      wave_bits(a.validity, obase, n, active && ok);
    }
  }
}

}  // namespace kg

// ======================================================================
// host side
// ======================================================================
namespace kg {

// A record's initial words, passed by value (no upload, no wait).
struct InitWords {
  uint64_t w[kAggWords];
};

// The words into LDS first (uniform indices into the argument), then each
// lane's words from LDS.
__global__ void k_fill_words(uint64_t *rec, uint64_t nrec, int W, InitWords iw) {
  __shared__ uint64_t s_w[kAggWords];
  for (int w = 0; w < W; w++)
    if (threadIdx.x == 0) s_w[w] = iw.w[w];
  __syncthreads();
  uint64_t total = nrec * (uint64_t)W;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x)
    rec[t] = s_w[t % W];
}

static InitWords init_words(const AggProg &P) {
  InitWords iw{};
  for (int w = 0; w < P.W && w < kAggWords; w++) iw.w[w] = P.init[w];
  return iw;
}

// One byte per row (0/1) to an LSB bitmap: each thread owns a 32-bit word.
__global__ void k_bytes_to_bits(const uint8_t *in, int64_t n, uint32_t *out) {
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w * 32 < n;
       w += (int64_t)gridDim.x * blockDim.x) {
    uint32_t bits = 0;
    for (int k = 0; k < 32 && w * 32 + k < n; k++)
      if (in[w * 32 + k]) bits |= 1u << k;
    out[w] = bits;
  }
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
      d = table_insert(T, k);  // the host sets the occupancy count
    }
    for (int w = 0; w < P.W; w++) T.rec[d * P.W + w] = orec[s * P.W + w];
  }
}

// The initial state into every record whose slot holds no key.
__global__ void k_init_free_records(Table T, int W, InitWords iw) {
  __shared__ uint64_t s_w[kAggWords];
  for (int w = 0; w < W; w++)
    if (threadIdx.x == 0) s_w[w] = iw.w[w];
  __syncthreads();
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < T.cap * (uint64_t)W;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t sl = t / (uint64_t)W;
    if (T.keys[sl] == kEmpty) T.rec[t] = s_w[t - sl * (uint64_t)W];
  }
}

// Group flags of a hashed table: slots [0, cap) hold a group when they hold
// a key, the two fixed slots when they have rows.
__global__ void k_group_flags_keys(const uint64_t *keys, const uint64_t *rec, uint64_t cap, int W,
                                   uint32_t *flags) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < cap + 2;
       s += (uint64_t)gridDim.x * blockDim.x)
    flags[s] = (s < cap ? keys[s] != kEmpty : rec[s * (uint64_t)W] > 0) ? 1u : 0u;
}

__global__ void k_group_flags(const uint64_t *rec, uint64_t nslots, int W, int global,
                              uint32_t *flags) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < nslots;
       s += (uint64_t)gridDim.x * blockDim.x)
    flags[s] = (rec[s * (uint64_t)W] > 0 || (global && s == 0)) ? 1u : 0u;
}

__global__ void k_compact_slots(const uint32_t *flags, const uint64_t *pos, uint64_t ns,
                                uint64_t *list) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < ns;
       s += (uint64_t)gridDim.x * blockDim.x)
    if (flags[s]) list[pos[s]] = s;
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
  bool has_key = false, direct = false, ansi = false, merge = false, partial = false;
  bool hashed = false;  // ordinary keys: pass 1 + spill + drain
  kg_column key_type{};
  std::vector<kg_column> in_types;
  int nsink = 0;  // input columns per sink call (state columns when merging)
  AggProg prog{};
  OutProg out{};
  int p1_slots = 0, m_slots = 0;  // LDS tables of pass 1 and of the merge
  size_t p1_lds = 0, m_lds = 0;
  int log2span_max = 0;  // the largest window (log2 slots) the windowed merge holds in LDS
  size_t w_lds_max = 0;  // the windowed merge's dynamic LDS limit
  bool stage_rows = false;  // the windowed merge stages spill rows in LDS
  bool rec_init = false;  // every window record holds the initial state (else only occupied ones)
  int merge_form = KG_AGG_MERGE_AUTO;
  bool size_from_sketch = true;
  DevBuf hll;             // kHllRegs registers of the spill's sketch
  // table
  DevBuf keys, rec, occ;
  uint64_t cap = 0;
  int log2cap = 0;
  uint64_t seed = 0x9E3779B97F4A7C15ULL;
  uint64_t occ_bound = 0;
  uint64_t merged_keys = 0;  // keys the merge inserted (not in the device count)
  // spill (hashed keys)
  SpillFmt fmt{};
  DevBuf sp_rows, sp_recs, sp_cnt;  // sp_cnt: rows, records, overflow flags
  uint64_t cap_rows = 0, cap_recs = 0, bound_rows = 0, bound_recs = 0;
  uint64_t spill_budget = 0;
  // ordered sums
  std::vector<CheckedSum> checked;
  uint64_t rows_seen = 0;
  uint64_t pending_rows = 0;  // rows of the sink call in progress
  DevBuf small_bits;          // the few-groups path: used-slot bits and its maps
  // PartialAggregate's BypassPolicy (aggregate/partial.rs:20-50): partial
  // state batches emitted by flushes and by pass-through, taken with
  // kg_agg_partial_next
  bool has_policy = false, bypass = false;
  uint64_t pol_min_rows = 0, pol_pct = 0, pol_max_bytes = 0;
  uint64_t part_rows = 0, flushed_groups = 0, flushes = 0, bypassed_rows = 0;
  std::vector<kg_batch *> emitted;
  size_t emitted_next = 0;
  DevBuf scratch_slots, scratch_abs;
  bool finalized = false;
  int64_t ngroups = 0;
  DevBuf pos;
  std::unique_ptr<kg::HostGather> gather;  // small host batches waiting for pass 1
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
  t.wmask = (a->hashed ? std::min<uint64_t>(a->cap, 1ull << a->log2span_max) : a->cap) - 1;
  return t;
}

static int log2_span(const kg_hash_agg *a) {
  return a->hashed ? std::min(a->log2cap, a->log2span_max) : a->log2cap;
}

// A hashed table's window records are initialized lazily: every path that
// reads a slot's record before a key is in it calls ensure_rec_init first
// (the windowed merge never does: it starts empty slots from the initial
// state in LDS).  The two fixed records are initialized at once.
static void alloc_table(kg_hash_agg *a, uint64_t cap, DevBuf &keys, DevBuf &rec) {
  uint64_t ns = a->direct ? 257 : cap + 2;
  keys.reset(a->ctx, (a->direct ? 257 : cap) * 8);
  rec.reset(a->ctx, ns * (uint64_t)a->prog.W * 8);
  k_fill_u64<<<stream_grid(keys.bytes() / 8, 256), 256, 0, a->s>>>(keys.get<uint64_t>(),
                                                                   keys.bytes() / 8, kEmpty);
  KG_HIP(hipGetLastError());
  if (a->hashed) {
    init_records(a->ctx, rec.get<uint64_t>() + cap * (uint64_t)a->prog.W, 2, a->prog, a->s);
    a->rec_init = false;
  } else {
    init_records(a->ctx, rec.get<uint64_t>(), ns, a->prog, a->s);
  }
}

static void ensure_rec_init(kg_hash_agg *a) {
  if (!a->hashed || a->rec_init) return;
  // only slots without a key: occupied records hold state
  const Table t = table_of(a);
  k_init_free_records<<<stream_grid(a->cap, 256), 256, 0, a->s>>>(t, a->prog.W,
                                                                  init_words(a->prog));
  KG_HIP(hipGetLastError());
  a->rec_init = true;
}

static int log2_ceil(uint64_t v) {
  int b = 0;
  while ((1ull << b) < v) b++;
  return b;
}

// Keys in the table: the device count (inserts made with global_slot) plus
// the merge's inserts.
static uint64_t read_occupied(kg_hash_agg *a) {
  PinnedReads pin;
  uint64_t *v = pin.take<uint64_t>(2);
  KG_HIP(hipMemcpyAsync(v, a->occ.get(), 16, hipMemcpyDeviceToHost, a->s));
  KG_HIP(hipStreamSynchronize(a->s));
  if (v[1]) fail(KG_ERR_INTERNAL, "a window of the aggregation table filled");
  return v[0] + a->merged_keys;
}

// ---- NeedMemory: a call checks the HBM its worst case may take against the
// context's budget before it changes any state, so that a refusal leaves the
// aggregate as it was and the same batch can be retried (operator.rs:144-161,
// 314-329: NeedMemory{bytes} with the state unchanged).
static void need_room(kg_ctx *ctx, uint64_t need, const char *what) {
  const uint64_t held = ctx->in_use.load();
  if (ctx->budget && held + need > ctx->budget)
    fail(KG_ERR_MEMORY, std::string(what) + " needs " + std::to_string(need) +
                            " more bytes of HBM (" + std::to_string(held) + " of " +
                            std::to_string(ctx->budget) + " held)");
}

static uint64_t table_bytes(const kg_hash_agg *a, uint64_t cap) {
  const uint64_t ns = a->direct ? 257 : cap + 2;
  return (a->direct ? 257 : cap) * 8 + ns * (uint64_t)a->prog.W * 8;
}

// At most how many distinct keys rows can hold: a one-to-four-byte key has
// few values.
static uint64_t key_domain(const kg_hash_agg *a) {
  const int w = type_width(a->key_type.type);
  return w >= 1 && w <= 4 ? (1ull << (8 * w)) + 1 : ~0ull;
}

// The peak HBM a drain of nr spilled rows and nc spilled records may take
// beyond what is held: their partitioned copies, the merge's scratch, and
// the table grown for the worst case of every entry being a new key (at
// most `new_keys`; the old table is held while the new one fills).
// The table's size for a drain: at most half full once the keys held and
// the new keys (the sketch's estimate with a tenth more, at most one a
// spilled entry) are in.
static uint64_t drain_cap(kg_hash_agg *a, uint64_t nr, uint64_t nc, uint64_t occupied,
                          double est) {
  uint64_t want = occupied + std::min<uint64_t>((uint64_t)(est * 1.1) + 64, nr + nc);
  want = std::min<uint64_t>(want, key_domain(a));
  uint64_t cap = a->cap;
  while (cap / 2 < want) cap *= 2;
  return cap;
}

// The peak HBM a drain of nr spilled rows and nc spilled records may take
// beyond what is held: their partitioned copies, the merge's scratch, and
// the grown table (from the sketch's estimate est of new keys when known;
// otherwise for the worst case of every entry being a new key, at most
// `new_keys`; the old table is held while the new one fills).
static uint64_t drain_need(kg_hash_agg *a, uint64_t nr, uint64_t nc, uint64_t occupied,
                           uint64_t new_keys, uint64_t spill_bytes = 0, double est = -1.0) {
  if (!nr && !nc) return 0;
  const uint64_t W = (uint64_t)a->prog.W;
  const uint64_t grid = (uint64_t)(a->ctx->num_cus > 0 ? a->ctx->num_cus : 256);
  // the partitioned copies of the spill (its bytes, when given)
  uint64_t need = (spill_bytes ? spill_bytes : (nr * a->fmt.rw + nc * (1 + W)) * 8) + 16;
  need += grid * a->m_slots * (1 + W) * 8 + grid * kMergeRows * kMergeBlock * 8;
  // histograms and their scans (4 + 8 bytes a partition and block), counters,
  // window flags
  int log2p = 0;
  while (log2p < 15 && (1ull << log2p) * 4096 < nr + nc) log2p++;
  const uint64_t fill = grid * 2;
  for (uint64_t n : {nr, nc}) {
    if (!n) continue;
    const uint64_t per_block = std::max<uint64_t>(
        4096, std::min<uint64_t>((1ull << log2p) * 16, (n + fill - 1) / fill));
    need += ((n + per_block - 1) / per_block) * (12ull << log2p) * 2;
  }
  need += (16ull << 20);
  uint64_t cap = a->cap;
  if (est >= 0.0) {
    cap = drain_cap(a, nr, nc, occupied, est);
  } else {
    uint64_t want = std::min<uint64_t>(occupied + new_keys, occupied + nr + nc);
    want = std::min<uint64_t>(want, key_domain(a));
    while (cap / 2 < want) cap *= 2;
  }
  if (cap > a->cap) need += table_bytes(a, cap);
  return need;
}

// Doubles the table (rehash of every key and record).  keys_now: the
// occupied count when the caller has just read it.
static void grow(kg_hash_agg *a, uint64_t min_cap, uint64_t keys_now = ~0ull) {
  if (keys_now == ~0ull) keys_now = read_occupied(a);
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
  Table t = table_of(a);
  k_rehash<<<stream_grid(ocap + 2, 256), 256, 0, a->s>>>(a->prog, okeys.get<uint64_t>(),
                                                          orec.get<uint64_t>(), ocap, t);
  KG_HIP(hipGetLastError());
  k_fill_u64<<<1, 64, 0, a->s>>>(a->occ.get<uint64_t>(), 1, keys_now);
  KG_HIP(hipGetLastError());
  a->merged_keys = 0;
  (void)read_occupied(a);  // a window that filled during the rehash
}

// Rows the next launch may take so that the table stays below 3/4 full
// whatever keys they hold (launches that insert into the table directly).
static int64_t room(kg_hash_agg *a, int64_t want) {
  if (a->direct || !a->has_key) return want;  // fixed slots: nothing is inserted
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
  if (a->direct || !a->has_key) return;
  uint64_t occ = read_occupied(a);
  a->occ_bound = occ;
  while (occ + (uint64_t)n > a->cap / 4 * 3) grow(a, a->cap * 2);
  ensure_rec_init(a);  // the rows' new keys start from their slots' records
}

// The few-groups path (a one-byte key, or a global aggregate): the used
// slots first (one-byte keys), then k_agg_small over every CU.
static void launch_small(kg_hash_agg *a, const AggProg &Q, int64_t n) {
  int D = 1;
  int16_t remap[257], slot_of[257];
  if (Q.direct) {
    a->small_bits.ensure(a->ctx, 9 * 4 + 2 * 257 * 2 + 64);
    KG_HIP(hipMemsetAsync(a->small_bits.get(), 0, 9 * 4, a->s));
    k_small_slots<<<stream_grid(n, 256, 16), 256, 0, a->s>>>(Q, n, a->small_bits.get<unsigned>());
    KG_HIP(hipGetLastError());
    unsigned bits[9];
    {
      PinnedReads pin;
      unsigned *p = pin.take<unsigned>(9);
      KG_HIP(hipMemcpyAsync(p, a->small_bits.get(), 36, hipMemcpyDeviceToHost, a->s));
      KG_HIP(hipStreamSynchronize(a->s));
      memcpy(bits, p, sizeof(bits));
    }
    D = 0;
    for (int sl = 0; sl < 257; sl++) {
      remap[sl] = -1;
      if ((bits[sl >> 5] >> (sl & 31)) & 1) {
        remap[sl] = (int16_t)D;
        slot_of[D++] = (int16_t)sl;
      }
    }
    if (D == 0) return;
    int16_t *dev_map = reinterpret_cast<int16_t *>(a->small_bits.get<uint8_t>() + 64);
    KG_HIP(hipMemcpyAsync(dev_map, remap, 257 * 2, hipMemcpyHostToDevice, a->s));
    KG_HIP(hipMemcpyAsync(dev_map + 257, slot_of, (size_t)D * 2, hipMemcpyHostToDevice, a->s));
  }
  const int WL = Q.W | 1;
  const size_t rec_bytes = (size_t)D * WL * 8;
  // record copies: up to 4 a wave, within 48 KB of LDS a workgroup
  int K = kSmallBlock / 64 * 4;
  while (K > 1 && (size_t)K * rec_bytes > 48 * 1024) K--;
  if ((size_t)K * rec_bytes > 64 * 1024) {
    ProfScope prof("agg_update", a->s);
    k_agg_rows_atomic<<<stream_grid(n, kAggBlock), kAggBlock, 0, a->s>>>(Q, table_of(a), n);
    KG_HIP(hipGetLastError());
    return;
  }
  const int64_t cus = a->ctx->num_cus > 0 ? a->ctx->num_cus : 256;
  constexpr int R = 4;
  int64_t blocks = std::min<int64_t>(cus * 4, ceil_div(n, (int64_t)R * kSmallBlock * 4));
  if (blocks < 1) blocks = 1;
  const int64_t chunk = ceil_div(n, blocks);
  const int16_t *dm = Q.direct ? reinterpret_cast<int16_t *>(a->small_bits.get<uint8_t>() + 64)
                               : nullptr;
  ProfScope prof("agg_update", a->s);
  k_agg_small<R><<<(unsigned)blocks, kSmallBlock, (size_t)K * rec_bytes, a->s>>>(
      Q, table_of(a), n, chunk, dm, dm ? dm + 257 : nullptr, D, K);
  KG_HIP(hipGetLastError());
}

static void launch_update(kg_hash_agg *a, const AggProg &P, int64_t lo, int64_t n) {
  if (n <= 0) return;
  AggProg Q = P;
  // offset the columns to this piece of the batch
  for (int c = 0; c < kAggInputs; c++) Q.in[c].offset += lo;
  Q.key.offset += lo;
  if (a->hashed) fail(KG_ERR_INTERNAL, "hashed keys are aggregated by pass 1");
  launch_small(a, Q, n);
}

// ---------------------------------------------------------------- spill
static SpillDev spill_of(kg_hash_agg *a) {
  SpillDev S{};
  S.rows = a->sp_rows.get<uint64_t>();
  S.recs = a->sp_recs.get<uint64_t>();
  S.nrows = a->sp_cnt.get<unsigned long long>();
  S.nrecs = S.nrows + 1;
  S.overflow = reinterpret_cast<unsigned int *>(S.nrows + 2);
  S.cap_rows = a->cap_rows;
  S.cap_recs = a->cap_recs;
  S.F = a->fmt;
  return S;
}

// The spill counters and the table's occupied count with one round trip.
static void read_counts(kg_hash_agg *a, uint64_t &nrows, uint64_t &nrecs, uint64_t &occupied) {
  PinnedReads pin;
  uint64_t *c = pin.take<uint64_t>(3), *v = pin.take<uint64_t>(2);
  KG_HIP(hipMemcpyAsync(c, a->sp_cnt.get(), 24, hipMemcpyDeviceToHost, a->s));
  KG_HIP(hipMemcpyAsync(v, a->occ.get(), 16, hipMemcpyDeviceToHost, a->s));
  KG_HIP(hipStreamSynchronize(a->s));
  if (c[2]) fail(KG_ERR_INTERNAL, "aggregation spill overflow");
  if (v[1]) fail(KG_ERR_INTERNAL, "a window of the aggregation table filled");
  nrows = c[0];
  nrecs = c[1];
  occupied = v[0] + a->merged_keys;
}

static void read_spill_counts(kg_hash_agg *a, uint64_t &nrows, uint64_t &nrecs) {
  PinnedReads pin;
  uint64_t *c = pin.take<uint64_t>(3);
  KG_HIP(hipMemcpyAsync(c, a->sp_cnt.get(), 24, hipMemcpyDeviceToHost, a->s));
  KG_HIP(hipStreamSynchronize(a->s));
  if (c[2]) fail(KG_ERR_INTERNAL, "aggregation spill overflow");
  nrows = c[0];
  nrecs = c[1];
}

// A buffer of `cap` spill entries of w words keeping the first `used`.
static void resize_spill(kg_hash_agg *a, DevBuf &buf, uint64_t &cap, uint64_t want, uint64_t used,
                         int w) {
  DevBuf nb(a->ctx, want * w * 8 + 8);
  if (used)
    KG_HIP(hipMemcpyAsync(nb.get(), buf.get(), used * w * 8, hipMemcpyDeviceToDevice, a->s));
  KG_HIP(hipStreamSynchronize(a->s));
  buf = std::move(nb);
  cap = want;
}

static void drain(kg_hash_agg *a);

// Room in both spills for a pass-1 launch of n rows from `blocks`
// workgroups; drains when the spill budget would be passed.
static void spill_room(kg_hash_agg *a, int64_t n, int64_t blocks) {
  const uint64_t rw = a->fmt.rw, cw = 1 + a->prog.W;
  // rows: every row may miss, and each wave may leave a piece partly unused;
  // records: every slot at each eviction and at the end
  const int64_t chunk = ceil_div(n, blocks);
  const uint64_t new_recs = (uint64_t)blocks * a->p1_slots * (uint64_t)(chunk / kEvictRows + 2);
  n += blocks * (kP1Block / 64) * kWaveChunk;
  if (a->bound_rows + n <= a->cap_rows && a->bound_recs + new_recs <= a->cap_recs) return;
  uint64_t nr, nc;
  read_spill_counts(a, nr, nc);
  a->bound_rows = nr;
  a->bound_recs = nc;
  uint64_t want_rows = nr + n, want_recs = nc + new_recs;
  if ((want_rows * rw + want_recs * cw) * 8 > a->spill_budget && (nr || nc)) {
    drain(a);
    nr = nc = 0;
    want_rows = n;
    want_recs = new_recs;
  }
  if (want_rows > a->cap_rows) {
    uint64_t c = std::max<uint64_t>(want_rows, a->cap_rows + a->cap_rows / 2);
    if (c * rw * 8 > a->spill_budget) c = want_rows;
    resize_spill(a, a->sp_rows, a->cap_rows, c, nr, (int)rw);
  }
  if (want_recs > a->cap_recs) {
    uint64_t c = std::max<uint64_t>(want_recs, a->cap_recs + a->cap_recs / 2);
    resize_spill(a, a->sp_recs, a->cap_recs, c, nc, (int)cw);
  }
}

// The sink's column types, kept for reading spill rows back; the spill row
// layout follows them (key, validity bits, one word per input, two for
// decimal128).
static void record_types(kg_hash_agg *a, const AggProg &P) {
  if (a->fmt.rw == 0) {
    int w = 2;
    a->fmt.vword = 1;
    a->fmt.ncols = a->nsink;
    // the first 32-bit integer column rides in word 1's low half when the
    // validity bits fit above it (bits 32-62)
    a->fmt.packed = 0;
    a->fmt.vshift = 0;
    if (a->nsink <= 31)
      for (int c = 0; c < a->nsink; c++)
        if (P.in[c].type == KG_T_INT32 || P.in[c].type == KG_T_DATE32) {
          a->fmt.packed = (uint8_t)(c + 1);
          a->fmt.vshift = 32;
          break;
        }
    for (int c = 0; c < a->nsink; c++) {
      a->prog.in[c].type = P.in[c].type;
      a->prog.in[c].precision = P.in[c].precision;
      a->prog.in[c].scale = P.in[c].scale;
      a->prog.in[c].width = P.in[c].width;
      if (a->fmt.packed == c + 1) {
        a->fmt.off[c] = 1;
        continue;
      }
      a->fmt.off[c] = (uint8_t)w;
      w += P.in[c].type == KG_T_DECIMAL128 ? 2 : 1;
    }
    a->fmt.rw = w;
    return;
  }
  for (int c = 0; c < a->nsink; c++)
    if (a->prog.in[c].type != P.in[c].type) fail(KG_ERR_INTERNAL, "input column type changed");
}

static void pass1(kg_hash_agg *a, const AggProg &P, int64_t nrows) {
  record_types(a, P);
  const int64_t per_wg = (int64_t)kTileRows * kP1Block * 8;
  const int64_t maxb = (int64_t)(a->ctx->num_cus > 0 ? a->ctx->num_cus : 256);
  // pieces whose row spill fits the budget
  const int64_t piece_max =
      std::max<int64_t>(1 << 20, (int64_t)(a->spill_budget / 2 / (8 * (uint64_t)a->fmt.rw)));
  for (int64_t done = 0; done < nrows;) {
    const int64_t n = std::min<int64_t>(nrows - done, piece_max);
    int64_t blocks = std::min<int64_t>(ceil_div(n, per_wg), maxb);
    if (blocks < 1) blocks = 1;
    const int64_t chunk = ceil_div(n, blocks);
    spill_room(a, n, blocks);
    AggProg Q = P;
    for (int c = 0; c < kAggInputs; c++) Q.in[c].offset += done;
    Q.key.offset += done;
    {
      ProfScope prof("agg_update", a->s);
      k_agg_pass1<kTileRows><<<(unsigned)blocks, kP1Block, a->p1_lds, a->s>>>(
          Q, table_of(a), spill_of(a), n, chunk, a->p1_slots);
      KG_HIP(hipGetLastError());
    }
    a->bound_rows += (uint64_t)n + (uint64_t)blocks * (kP1Block / 64) * kWaveChunk;
    a->bound_recs += (uint64_t)blocks * a->p1_slots * (uint64_t)(chunk / kEvictRows + 2);
    done += n;
  }
}

// One spill stream (n entries of w words) partitioned by home-slot range:
// the per-block histogram first (spill_hist; with the sketch, the keys also
// go into a->hll), then the scan and the scatter into `out` (spill_scatter;
// partition p starts at base[p * nb]).
// Workgroup of the partition kernels: their LDS holds a counter a partition,
// so past 2^13 partitions one workgroup fills a CU; it then takes 1,024
// lanes (16 waves) to keep enough loads and stores in flight.
static unsigned spill_block(uint32_t np) { return np > (1u << 13) ? 1024u : 256u; }

struct SpillPart {
  const uint64_t *in = nullptr;
  uint64_t n = 0;
  int w = 0;
  bool dead = false;
  int log2p = 0;
  int64_t per_block = 0;
  uint32_t nb = 0;
  DevBuf hist;
};

static void spill_hist(kg_hash_agg *a, SpillPart &sp, const uint64_t *in, uint64_t n, int w,
                       bool dead, int log2p, bool sketch) {
  const uint32_t np = 1u << log2p;
  // blocks enough to fill the device, each with about 16 entries a partition
  const int64_t fill = (int64_t)(a->ctx->num_cus > 0 ? a->ctx->num_cus : 256) * 2;
  int64_t per_block = std::max<int64_t>(4096, std::min<int64_t>((int64_t)np * 16,
                                                                ceil_div((int64_t)n, fill)));
  int64_t nb = ceil_div((int64_t)n, per_block);
  if (nb < 1) nb = 1;
  per_block = ceil_div((int64_t)n, nb);
  if (per_block < 1) per_block = 1;
  sp.in = in;
  sp.n = n;
  sp.w = w;
  sp.dead = dead;
  sp.log2p = log2p;
  sp.per_block = per_block;
  sp.nb = (uint32_t)nb;
  sp.hist.reset(a->ctx, (uint64_t)np * nb * 4 + 4);
  Table t = table_of(a);
  ProfScope prof("agg_partition", a->s);
  k_spill_hist<<<(unsigned)nb, spill_block(np), np * 4 + (sketch ? 4 * kHllRegs : 0), a->s>>>(
      in, w, n, t, log2p, per_block, dead ? 1 : 0, sp.hist.get<uint32_t>(),
      sketch ? a->hll.get<unsigned int>() : nullptr);
  KG_HIP(hipGetLastError());
}

// The scan of the histogram into base (partition p of block b starts at
// base[p * nb + b]); the live total is read back with the caller's next
// round trip: base[hn - 1] + hist[hn - 1] (enqueue_total).
static void spill_scan(kg_hash_agg *a, SpillPart &sp, DevBuf &base) {
  const uint64_t hn = ((uint64_t)1 << sp.log2p) * sp.nb;
  base.reset(a->ctx, hn * 8 + 8);
  ProfScope prof("agg_partition", a->s);
  DevBuf temp(a->ctx, (uint64_t)scan_partials((int64_t)hn) * 8);
  exclusive_scan(sp.hist.get<uint32_t>(), base.get<uint64_t>(), (int64_t)hn, temp.get<uint64_t>(),
                 a->s);
}

// The last partition's base and count (pinned host memory, read after the
// caller's sync).
struct SpillTotal {
  uint64_t *last_base = nullptr;
  uint32_t *last_count = nullptr;
  uint64_t live() const { return *last_base + *last_count; }
};

static void enqueue_total(kg_hash_agg *a, const SpillPart &sp, const DevBuf &base,
                          const SpillTotal &t) {
  const uint64_t hn = ((uint64_t)1 << sp.log2p) * sp.nb;
  KG_HIP(hipMemcpyAsync(t.last_base, base.get<uint64_t>() + hn - 1, 8, hipMemcpyDeviceToHost,
                        a->s));
  KG_HIP(hipMemcpyAsync(t.last_count, sp.hist.get<uint32_t>() + hn - 1, 4,
                        hipMemcpyDeviceToHost, a->s));
}

// The live entries to `out`, partitioned by the scan in base.
static void spill_scatter(kg_hash_agg *a, SpillPart &sp, DevBuf &out, const DevBuf &base) {
  const uint32_t np = 1u << sp.log2p;
  out.reset(a->ctx, sp.n * sp.w * 8 + 8);
  Table t = table_of(a);
  ProfScope prof("agg_partition", a->s);
  k_spill_scatter<<<sp.nb, spill_block(np), np * 4, a->s>>>(sp.in, sp.w, sp.n, t, sp.log2p,
                                                           sp.per_block, sp.dead ? 1 : 0,
                                                           base.get<uint64_t>(),
                                                           out.get<uint64_t>());
  KG_HIP(hipGetLastError());
}

// The sketch's estimate of the distinct keys spilled since the last drain
// (HyperLogLog with the small-range correction of linear counting).
// reg: the kHllRegs registers read back.
static double spilled_keys_estimate(const uint32_t *reg) {
  const double m = kHllRegs;
  double sum = 0.0;
  int zeros = 0;
  for (int r = 0; r < kHllRegs; r++) {
    const uint32_t v = reg[r];
    sum += std::ldexp(1.0, -(int)v);
    zeros += v == 0;
  }
  double e = 0.7213 / (1.0 + 1.079 / m) * m * m / sum;
  if (e <= 2.5 * m && zeros) e = m * std::log(m / zeros);
  return e;
}

// The old merge (k_agg_merge: an LDS hash table per partition, merged into
// the global table by key), for drains whose partitions cover many windows
// each: few spilled inputs a window.
static void merge_by_key(kg_hash_agg *a, const MergeDev &M0, int log2p, uint32_t np) {
  const int W = a->prog.W;
  const int64_t grid = std::min<int64_t>(np, (int64_t)(a->ctx->num_cus > 0 ? a->ctx->num_cus : 256));
  DevBuf resume(a->ctx, (uint64_t)np * 8), inserted(a->ctx, (uint64_t)np * 4);
  DevBuf retry(a->ctx, (uint64_t)grid * a->m_slots * (1 + W) * 8);
  DevBuf pend(a->ctx, (uint64_t)grid * kMergeRows * kMergeBlock * 8);
  DevBuf ctr(a->ctx, (uint64_t)grid * 12 + 16);  // retry_n, pend_n, pend_part, stops
  KG_HIP(hipMemsetAsync(resume.get(), 0, resume.bytes(), a->s));
  AggProg P = a->prog;
  P.skip_checked = 0;
  set_spill_layout(P, a->fmt);
  MergeDev M = M0;
  M.log2p = log2p;
  M.resume = resume.get<int64_t>();
  M.inserted = inserted.get<uint32_t>();
  M.retry = retry.get<uint64_t>();
  M.retry_n = ctr.get<uint32_t>();
  M.pend_n = M.retry_n + grid;
  M.pend_part = M.pend_n + grid;
  M.stops = M.pend_part + grid;
  M.pend = pend.get<int64_t>();
  ensure_rec_init(a);  // new keys' records are merged into, not written
  for (int round = 0;; round++) {
    if (round > 40) fail(KG_ERR_INTERNAL, "aggregation merge made no progress");
    const uint64_t keys_now = read_occupied(a);
    const uint64_t limit = a->cap - a->cap / 8;
    M.quota = keys_now >= limit ? 0 : (uint32_t)std::min<uint64_t>((limit - keys_now) / np, 0xFFFFFFFFu);
    KG_HIP(hipMemsetAsync(inserted.get(), 0, inserted.bytes(), a->s));
    KG_HIP(hipMemsetAsync(ctr.get(), 0, ctr.bytes(), a->s));
    {
      ProfScope prof("agg_merge", a->s);
      k_agg_merge<kMergeRows><<<(unsigned)grid, kMergeBlock, a->m_lds, a->s>>>(P, table_of(a), M,
                                                                            a->m_slots);
      KG_HIP(hipGetLastError());
    }
    // keys inserted, and workgroups that stopped
    std::vector<uint32_t> ins(np);
    uint32_t stops = 0;
    KG_HIP(hipMemcpyAsync(ins.data(), inserted.get(), (size_t)np * 4, hipMemcpyDeviceToHost, a->s));
    KG_HIP(hipMemcpyAsync(&stops, M.stops, 4, hipMemcpyDeviceToHost, a->s));
    KG_HIP(hipStreamSynchronize(a->s));
    for (uint32_t v : ins) a->merged_keys += v;
    if (stops == 0) break;
    // the table is too full for some partition: grow so that every leftover
    // (retry records, pending inputs) fits as a new key below 3/4, merge
    // them with atomics, resume
    std::vector<uint32_t> left((size_t)grid * 2);
    KG_HIP(hipMemcpyAsync(left.data(), M.retry_n, (size_t)grid * 8, hipMemcpyDeviceToHost, a->s));
    KG_HIP(hipStreamSynchronize(a->s));
    uint64_t nleft = 0;
    for (uint32_t v : left) nleft += v;
    const uint64_t need = read_occupied(a) + nleft;
    uint64_t want = a->cap * 2;
    while (want / 4 * 3 < need) want *= 2;
    grow(a, want);
    ensure_rec_init(a);
    ProfScope prof("agg_merge", a->s);
    k_agg_retry<<<(unsigned)grid, 256, 0, a->s>>>(P, table_of(a), M, a->m_slots, (uint32_t)grid);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(a->s));
  }
}

// The windowed merge (k_agg_wmerge) over partitions of 2^g windows, run
// again over the windows not done after each growth of the table (a window
// that filled).
static void merge_windows(kg_hash_agg *a, const MergeDev &M0, int log2p, uint32_t np, bool load) {
  AggProg P = a->prog;
  P.skip_checked = 0;
  set_spill_layout(P, a->fmt);
  WMergeDev M{};
  M.rows = M0.rows;
  M.nrows = M0.nrows;
  M.rbase = M0.rbase;
  M.rnb = M0.rnb;
  M.recs = M0.recs;
  M.nrecs = M0.nrecs;
  M.cbase = M0.cbase;
  M.cnb = M0.cnb;
  M.log2p = log2p;
  M.F = a->fmt;
  M.load = load ? 1 : 0;
  int lw = a->log2cap - log2_span(a);
  DevBuf done(a->ctx, (1ull << lw) + 8), failed(a->ctx, 8);
  KG_HIP(hipMemsetAsync(done.get(), 0, done.bytes(), a->s));
  const int64_t grid = std::min<int64_t>(np, (int64_t)(a->ctx->num_cus > 0 ? a->ctx->num_cus : 256));
  for (int round = 0;; round++) {
    M.log2span = log2_span(a);
    M.g = lw - log2p;
    M.done = done.get<uint8_t>();
    M.failed = failed.get<unsigned int>();
    const size_t lds = ((size_t)1 << M.log2span) * (1 + (size_t)(a->prog.W | 1)) * 8;
    // spill rows staged in LDS when a tile of them fits beside the window
    const size_t stage = (size_t)kMergeBlock * a->fmt.rw * 8;
    const bool staged = a->stage_rows && a->fmt.rw <= kStageWords && lds + stage <= a->w_lds_max;
    KG_HIP(hipMemsetAsync(failed.get(), 0, 8, a->s));
    {
      ProfScope prof("agg_merge", a->s);
      if (staged)
        k_agg_wmerge<kWMergeRows, true><<<(unsigned)grid, kMergeBlock, lds + stage, a->s>>>(
            P, table_of(a), M);
      else
        k_agg_wmerge<kWMergeRows, false><<<(unsigned)grid, kMergeBlock, lds, a->s>>>(
            P, table_of(a), M);
      KG_HIP(hipGetLastError());
    }
    unsigned int nf = 0;
    {
      PinnedReads pin;
      unsigned int *p = pin.take<unsigned int>(1);
      KG_HIP(hipMemcpyAsync(p, failed.get(), 4, hipMemcpyDeviceToHost, a->s));
      KG_HIP(hipStreamSynchronize(a->s));
      nf = *p;
    }
    if (nf == 0) return;
    if (round >= 40) fail(KG_ERR_INTERNAL, "aggregation table windows kept filling");
    // twice the slots: the windows done keep their keys (rehashed), the rest
    // run again with the table's keys loaded
    std::vector<uint8_t> d0((size_t)1 << lw);
    KG_HIP(hipMemcpyAsync(d0.data(), done.get(), d0.size(), hipMemcpyDeviceToHost, a->s));
    KG_HIP(hipStreamSynchronize(a->s));
    try {
      // part of the spill is merged by now: a refusal cannot be retried
      need_room(a->ctx, table_bytes(a, a->cap * 2), "the growth of the aggregation table");
      grow(a, a->cap * 2);
    } catch (const Error &e) {
      if (e.code != KG_ERR_MEMORY) throw;
      fail(KG_ERR_INTERNAL, "cannot finish the aggregation merge: " + e.msg);
    }
    const int lw2 = a->log2cap - log2_span(a);
    std::vector<uint8_t> d1((size_t)1 << lw2);
    for (size_t w = 0; w < d1.size(); w++) d1[w] = d0[w >> (lw2 - lw)];
    DevBuf done2(a->ctx, d1.size() + 8);
    KG_HIP(hipMemcpyAsync(done2.get(), d1.data(), d1.size(), hipMemcpyHostToDevice, a->s));
    KG_HIP(hipStreamSynchronize(a->s));
    done = std::move(done2);
    lw = lw2;
    M.load = 1;
  }
}

// Merges both spills into the table and empties them.  The spill is
// partitioned by home-slot range into partitions of about 4096 entries; the
// partition histogram also counts the spilled keys' sketch, from which the
// table is sized (at most half full after the merge) with whole windows a
// partition.  Then each window is merged in LDS by its own workgroup, or,
// when the partitions hold few entries a window, by key (k_agg_merge).
static void drain(kg_hash_agg *a) {
  if (!a->hashed) return;
  uint64_t nr, nc, occ0;
  read_counts(a, nr, nc, occ0);
  if (nr == 0 && nc == 0) return;
  if (nr + nc >= (1ull << 32)) fail(KG_ERR_UNSUPPORTED, "more than 2^32 spilled entries");
  const int W = a->prog.W;
  // partitions of 4K-8K entries (at most 2^15), and whole windows each:
  // counted before the sketch sizes the table, again if that leaves fewer
  // windows than partitions.  The histograms' scans, their totals and the
  // sketch come back with one round trip.
  int log2p = 0;
  while (log2p < 15 && (1ull << log2p) * 8192 < nr + nc) log2p++;
  SpillPart pr, pc;
  DevBuf pbase, cbase;
  PinnedReads pin;
  uint32_t *reg = pin.take<uint32_t>(kHllRegs);
  SpillTotal tr{pin.take<uint64_t>(1), pin.take<uint32_t>(1)};
  SpillTotal tc{pin.take<uint64_t>(1), pin.take<uint32_t>(1)};
  auto hists = [&](bool sketch) {
    if (sketch) KG_HIP(hipMemsetAsync(a->hll.get(), 0, (size_t)kHllRegs * 4, a->s));
    if (nr) spill_hist(a, pr, a->sp_rows.get<uint64_t>(), nr, a->fmt.rw, true, log2p, sketch);
    if (nc) spill_hist(a, pc, a->sp_recs.get<uint64_t>(), nc, 1 + W, false, log2p, sketch);
    if (nr) spill_scan(a, pr, pbase);
    if (nc) spill_scan(a, pc, cbase);
    if (sketch)
      KG_HIP(hipMemcpyAsync(reg, a->hll.get(), (size_t)kHllRegs * 4, hipMemcpyDeviceToHost, a->s));
    if (nr) enqueue_total(a, pr, pbase, tr);
    if (nc) enqueue_total(a, pc, cbase, tc);
    KG_HIP(hipStreamSynchronize(a->s));
  };
  hists(a->size_from_sketch);
  const double est = a->size_from_sketch ? spilled_keys_estimate(reg) : -1.0;
  // everything the merge may take, checked before the spill or the table changes
  need_room(a->ctx, drain_need(a, nr, nc, occ0, a->rows_seen + a->pending_rows, 0, est),
            "the merge of the aggregation spill");
  if (est >= 0.0) {
    const uint64_t cap = drain_cap(a, nr, nc, occ0, est);
    if (cap > a->cap) grow(a, cap, occ0);
  }
  const int lw = a->log2cap - log2_span(a);  // log2 windows
  if (log2p > lw) {
    log2p = lw;
    hists(false);
  }
  const uint32_t np = 1u << log2p;
  DevBuf prow, prec;
  if (nr) spill_scatter(a, pr, prow, pbase);
  if (nc) spill_scatter(a, pc, prec, cbase);
  const uint64_t nr_live = nr ? tr.live() : 0, nc_live = nc ? tc.live() : 0;
  MergeDev M{};
  M.rows = prow.get<uint64_t>();
  M.nrows = nr_live;
  M.rbase = nr_live ? pbase.get<uint64_t>() : nullptr;
  M.rnb = pr.nb;
  M.recs = prec.get<uint64_t>();
  M.nrecs = nc_live;
  M.cbase = nc_live ? cbase.get<uint64_t>() : nullptr;
  M.cnb = pc.nb;
  M.F = a->fmt;
  const bool windows = a->merge_form == KG_AGG_MERGE_WINDOWS ||
                       (a->merge_form == KG_AGG_MERGE_AUTO && lw - log2p <= 2);
  if (windows)
    merge_windows(a, M, log2p, np, occ0 > 0);
  else
    merge_by_key(a, M, log2p, np);
  KG_HIP(hipMemsetAsync(a->sp_cnt.get(), 0, 24, a->s));
  KG_HIP(hipStreamSynchronize(a->s));
  a->bound_rows = a->bound_recs = 0;
}

bool checked_batch_proven(kg_ctx *ctx, hipStream_t s, DevBuf &scratch, const CheckedSum &c,
                          const DCol &col, int64_t n, uint64_t rows_before, uint64_t mult,
                          double &batch_abs) {
  // 1: every row so far at the type's largest magnitude, counted mult times
  const unsigned __int128 rows = (unsigned __int128)(rows_before + (uint64_t)n) * mult;
  const unsigned __int128 tb = c.type_bound;
  if (tb == 0 || rows <= (((unsigned __int128)0) - 1) / tb) {
    if (rows * tb < c.bound) {
      batch_abs = (double)n * (double)mult * (double)tb;
      return true;
    }
  }
  // 2: the stored sums' bound plus this batch's exact sum of magnitudes
  scratch.ensure(ctx, 8);
  KG_HIP(hipMemsetAsync(scratch.get(), 0, 8, s));
  k_abs_sum<<<stream_grid(n, 256), 256, 0, s>>>(col, n, scratch.get<double>());
  KG_HIP(hipGetLastError());
  double batch = 0.0;
  {
    PinnedReads pin;
    double *p = pin.take<double>(1);
    KG_HIP(hipMemcpyAsync(p, scratch.get(), 8, hipMemcpyDeviceToHost, s));
    KG_HIP(hipStreamSynchronize(s));
    batch = *p * (double)mult;
  }
  const double total = (c.cum_abs + batch) * (1.0 + 1e-9) + 1.0;
  batch_abs = batch;
  return total < (double)c.bound * (1.0 - 1e-9);
}

static bool proven(kg_hash_agg *a, const CheckedSum &c, const DCol &col, int64_t n,
                   double &batch_abs) {
  return checked_batch_proven(a->ctx, a->s, a->scratch_abs, c, col, n, a->rows_seen, 1,
                              batch_abs);
}

// The worst-case HBM a sink call of nrows takes beyond what is held (host
// columns are already staged by then): the spill buffers it may grow to,
// and (when the spill may pass its budget during the call, or the batch is
// replayed) the merge of a full spill.
static uint64_t sink_need(kg_hash_agg *a, const kg_column *inputs, int64_t nrows, bool replay) {
  uint64_t need = 0;
  const uint64_t new_keys = std::min<uint64_t>(a->rows_seen + (uint64_t)nrows, key_domain(a));
  if (!a->hashed) {
    if (replay) need += (uint64_t)nrows * 8;
    return need;
  }
  uint64_t nr = 0, nc = 0, occupied = 0;
  if (a->rows_seen) read_counts(a, nr, nc, occupied);  // else nothing is spilled or held
  if (replay) {  // drain, then every row into the table directly
    need += (uint64_t)nrows * 8 + drain_need(a, nr, nc, occupied, new_keys);
    uint64_t cap = a->cap;
    while (cap / 4 * 3 < std::min<uint64_t>(occupied + nr + nc + (uint64_t)nrows, key_domain(a)))
      cap *= 2;
    if (cap > a->cap) need += table_bytes(a, cap);
    return need;
  }
  uint64_t rw = a->fmt.rw;
  if (rw == 0) {  // spill row layout before the first sink: key, validity, inputs
    rw = 2;
    for (int c = 0; c < a->nsink; c++) rw += inputs[c].type == KG_T_DECIMAL128 ? 2 : 1;
  }
  const uint64_t cw = 1 + (uint64_t)a->prog.W;
  const int64_t maxb = (int64_t)(a->ctx->num_cus > 0 ? a->ctx->num_cus : 256);
  const uint64_t slack_rows = (uint64_t)maxb * (kP1Block / 64) * kWaveChunk;
  const uint64_t recs = (uint64_t)maxb * a->p1_slots *
                        ((uint64_t)ceil_div(nrows, maxb) / kEvictRows + 2);
  const uint64_t spill_after = ((nr + (uint64_t)nrows + slack_rows) * rw + (nc + recs) * cw) * 8;
  if (spill_after <= a->spill_budget) return need + spill_after;
  // pieces drain between them: fresh spill buffers up to the budget, and the
  // merge of a spill that size
  // (a full spill holds at most spill_budget bytes of rows and records together)
  const uint64_t full_rows = a->spill_budget / (rw * 8), full_recs = a->spill_budget / (cw * 8);
  return need + a->spill_budget +
         drain_need(a, full_rows, full_recs, occupied, new_keys, a->spill_budget);
}

static void emit_rows(kg_hash_agg *a, const AggProg &P, int64_t n);
static void after_partial_batch(kg_hash_agg *a, uint64_t nrows);
static uint64_t out_row_bytes(const kg_hash_agg *a);

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
  for (int c = 0; c < a->nsink; c++) {
    check_column(inputs[c], nrows, "aggregate input");
    if (!a->merge && inputs[c].type != a->in_types[c].type)
      fail(KG_ERR_INTERNAL, "input column type changed");
    P.in[c] = st.add(inputs[c], nrows);
  }
  if (a->has_policy && a->bypass) {
    // pass-through: the batch leaves as one partial state per row
    P.skip_checked = 0;
    const uint64_t rb = (uint64_t)a->prog.W * 8 + 8;
    need_room(a->ctx, (uint64_t)nrows * (rb + out_row_bytes(a)) + (1u << 20),
              "the pass-through batch");
    emit_rows(a, P, nrows);
    a->rows_seen += (uint64_t)nrows;
    a->part_rows += (uint64_t)nrows;
    return;
  }
  bool replay = false;
  std::vector<double> batch_abs(a->checked.size(), 0.0);
  for (size_t k = 0; k < a->checked.size(); k++)
    if (!proven(a, a->checked[k], P.in[a->checked[k].col], nrows, batch_abs[k])) replay = true;
  // NeedMemory before any state changes
  need_room(a->ctx, sink_need(a, inputs, nrows, replay), "the aggregation batch");
  a->pending_rows = (uint64_t)nrows;
  if (!replay) {
    P.skip_checked = 0;
    if (a->hashed) {
      pass1(a, P, nrows);
    } else {
      int64_t done = 0;
      while (done < nrows) {
        int64_t n = room(a, nrows - done);
        launch_update(a, P, done, n);
        done += n;
      }
    }
  } else {
    // rows straight to the global table (their record index kept), then the
    // checked sums in row order, after every spilled row is in the table
    drain(a);
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
  a->pending_rows = 0;
  for (size_t k = 0; k < a->checked.size(); k++) a->checked[k].cum_abs += batch_abs[k];
  // host inputs were staged into buffers freed on return: finish with them
  if (st.staged_host()) KG_HIP(hipStreamSynchronize(a->s));
  if (a->has_policy) after_partial_batch(a, (uint64_t)nrows);
}

// Aggregates the gathered rows.  The rows stay gathered when this fails, so
// a NeedMemory error leaves the operator as it was.
static void flush_gather(kg_hash_agg *a) {
  if (!a->gather || a->gather->rows() == 0) return;
  HostGather &g = *a->gather;
  kg_column cols[kGatherCols];
  g.submit(a->s, cols);
  const int k0 = a->has_key ? 1 : 0;
  try {
    sink(a, a->has_key ? &cols[0] : nullptr, cols + k0, g.rows());
  } catch (...) {
    g.keep();
    throw;
  }
  g.clear();
}

// The columns of a sink call: the grouping column (when there is one), then
// the inputs.
static int sink_columns(const kg_hash_agg *a, const kg_column *key, const kg_column *inputs,
                        const kg_column **cols) {
  int n = 0;
  if (a->has_key) cols[n++] = key;
  for (int c = 0; c < a->nsink; c++) cols[n++] = &inputs[c];
  return n;
}

static int64_t gather_limit(const kg_hash_agg *a, const kg_column *key, const kg_column *inputs) {
  const kg_column *cols[kGatherCols];
  const int n = sink_columns(a, key, inputs, cols);
  int width[kGatherCols];
  for (int c = 0; c < n; c++) width[c] = type_width(cols[c]->type);
  return gather_max_rows(width, n);
}

// Host batches of fewer rows than a slot of the largest gather are gathered
// (without a bypass policy, whose decisions follow each batch).
static bool gathers(const kg_hash_agg *a, const kg_column *key, const kg_column *inputs,
                    int64_t nrows) {
  if (a->has_policy || nrows <= 0 || (!a->has_key && a->nsink == 0)) return false;
  if (a->has_key && !key) return false;
  if (a->has_key && (key->flags & KG_COL_DEVICE)) return false;
  for (int c = 0; c < a->nsink; c++)
    if (inputs[c].flags & KG_COL_DEVICE) return false;
  return nrows < gather_limit(a, key, inputs) / 8;
}

// A host batch into the gather: checked as sink checks it, the gathered rows
// aggregated first when it does not fit, then copied to pinned staging.  A
// gather starts at 64 times the first batch's rows (64K to the largest), and
// is replaced by one four times larger when it fills, or by one sized for a
// batch that does not fit its slots, up to the largest for the row width.
static void gather_batch(kg_hash_agg *a, const kg_column *key, const kg_column *inputs,
                         int64_t nrows) {
  if (a->finalized) fail(KG_ERR_INTERNAL, "sink after finalize");
  if (a->has_key) {
    check_column(*key, nrows, "grouping column");
    if (key->type != a->key_type.type) fail(KG_ERR_INTERNAL, "grouping column type changed");
  }
  for (int c = 0; c < a->nsink; c++) {
    check_column(inputs[c], nrows, "aggregate input");
    if (!a->merge && inputs[c].type != a->in_types[c].type)
      fail(KG_ERR_INTERNAL, "input column type changed");
  }
  const kg_column *cols[kGatherCols];
  const int ncols = sink_columns(a, key, inputs, cols);
  int width[kGatherCols];
  for (int c = 0; c < ncols; c++) width[c] = type_width(cols[c]->type);
  const int64_t maxcap = gather_max_rows(width, ncols);
  const int64_t want = gather_rows_for(nrows, maxcap);
  int64_t cap = want;
  if (a->gather) {
    HostGather &g = *a->gather;
    bool widths = g.ncols() == ncols;
    for (int c = 0; c < ncols && widths; c++) widths = width[c] == g.width(c);
    const bool fits = g.rows() + nrows <= g.cap(), in_slot = nrows < g.slot_rows();
    if (g.rows() > 0 && (!widths || !g.same_types(cols) || !fits || !in_slot)) flush_gather(a);
    cap = g.cap();
    if (!fits && cap < maxcap) cap = std::min(cap * 4, maxcap);
    if (!in_slot || !widths) cap = std::max(cap, want);
    if (cap != g.cap() || !widths) a->gather.reset();  // empty by now
  }
  if (!a->gather) a->gather = std::make_unique<HostGather>(a->ctx, ncols, width, cap);
  a->gather->append(a->s, cols, nrows);
}

static void sink_or_gather(kg_hash_agg *a, const kg_column *key, const kg_column *inputs,
                           int64_t nrows) {
  if (gathers(a, key, inputs, nrows)) {
    gather_batch(a, key, inputs, nrows);
    return;
  }
  flush_gather(a);
  sink(a, key, inputs, nrows);
}

int compile_aggs(const std::vector<kg_column> &in_types, int32_t naggs, const kg_agg_spec *aggs,
                 bool ansi, bool merge, bool partial, AggProg &P, OutProg &O,
                 std::vector<CheckedSum> &checked) {
  int W = 1;  // word 0: rows (raw) or partial states (merge)
  std::vector<AggOp> ops;
  ops.push_back(AggOp{OP_COUNT_STAR, 0, 0, 0, 0, 0});
  std::vector<int> nulls_word(kAggInputs, -1);
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
  struct Words {
    int w, st;
  };
  std::vector<std::pair<Key, Words>> shared;
  // integral columns that an AVG sums exactly: a SUM of the same column
  // shares that 128-bit sum rather than keeping a word of its own
  std::vector<bool> avg_sums(in_types.size(), false);
  if (!merge)
    for (int i = 0; i < naggs; i++)
      if (aggs[i].func == KG_AGG_AVG && aggs[i].input >= 0 &&
          aggs[i].input < (int)in_types.size() && is_integral(in_types[aggs[i].input].type))
        avg_sums[aggs[i].input] = true;
  // one set of words per (op, column, checked): a decimal SUM and AVG of one
  // column share their checked sum and its state word (both results are
  // NULL, or an error under ANSI, once the running sum leaves the bound)
  auto op_words = [&](int kind, int col, bool chk, bool avg, int nwords, uint64_t iv,
                      int &st_word) {
    for (auto &kv : shared)
      if (kv.first.kind == kind && kv.first.col == col && kv.first.checked == (int)chk &&
          kv.first.avg == (int)avg) {
        st_word = kv.second.st;
        return kv.second.w;
      }
    int w = words(nwords, iv);
    st_word = chk ? words(1, 0) : 0;
    ops.push_back(AggOp{(uint8_t)kind, (uint8_t)col, (uint8_t)chk, 0, (uint16_t)w,
                        (uint16_t)st_word});
    shared.push_back({Key{kind, col, chk, avg}, Words{w, st_word}});
    return w;
  };
  // the ordered replay of a checked sum, once per set of words (SUM's
  // semantics when a SUM shares the words of an AVG)
  auto add_checked = [&](const CheckedSum &cs) {
    for (auto &x : checked)
      if (x.word == cs.word) {
        x.avg = x.avg && cs.avg;
        return;
      }
    checked.push_back(cs);
  };
  auto nulls_of = [&](int col) {
    if (col >= kAggInputs) fail(KG_ERR_UNSUPPORTED, "at most 32 input columns per sink");
    if (nulls_word[col] < 0) {
      nulls_word[col] = words(1, 0);
      ops.push_back(AggOp{OP_NULLS, (uint8_t)col, 0, 0, (uint16_t)nulls_word[col], 0});
    }
    return nulls_word[col];
  };
  auto plain = [&](int kind, int col, int word) {
    ops.push_back(AggOp{(uint8_t)kind, (uint8_t)col, 0, 0, (uint16_t)word, 0});
  };
  int next_state = 0;  // merge: the next state column
  O.n = naggs;
  O.partial = partial;
  for (int i = 0; i < naggs; i++) {
    OutAgg &o = O.a[i];
    o = OutAgg{};
    o.func = aggs[i].func;
    o.col = aggs[i].input;
    o.st_word = -1;
    o.nulls_word = -1;
    o.valid_word = -1;
    o.cnt_word = -1;
    int f = aggs[i].func;
    if (f == KG_AGG_COUNT_STAR || f == KG_AGG_COUNT) {
      o.kind = f == KG_AGG_COUNT_STAR ? R_COUNT_STAR : R_COUNT;
      o.out_type = o.st_type = KG_T_INT64;
      if (merge) {
        // state `#count` (aggregate/mod.rs:755): add the counts
        int c = next_state++;
        o.word = words(1, 0);
        plain(OP_ADDV, c, o.word);
        o.cnt_word = o.word;
      } else if (f == KG_AGG_COUNT) {
        int c = aggs[i].input;
        if (c < 0 || c >= (int)in_types.size()) fail(KG_ERR_INTERNAL, "aggregate input out of range");
        o.nulls_word = nulls_of(c);
      }
      continue;
    }
    int c0 = aggs[i].input;
    if (c0 < 0 || c0 >= (int)in_types.size()) fail(KG_ERR_INTERNAL, "aggregate input out of range");
    const kg_column &t = in_types[c0];
    int c = merge ? next_state : c0;  // the (first) column this aggregate reads
    o.in_type = t.type;
    o.in_s = t.scale;
    int st = 0;
    switch (f) {
      case KG_AGG_SUM:
      case KG_AGG_AVG: {
        bool avg = f == KG_AGG_AVG;
        if (is_float(t.type) || (avg && is_integral(t.type) && merge)) {
          // float sums; a merged integer avg state sums in double (avg_sum_type)
          o.kind = avg ? R_AVG_F : R_SUM_F;
          o.out_type = KG_T_FLOAT64;
          o.st_type = KG_T_FLOAT64;
          o.word = op_words(OP_ADDF64, c, false, false, 1, 0, st);
          o.nulls_word = nulls_of(c);
        } else if (is_integral(t.type)) {
          bool chk = !avg && ansi;
          o.kind = avg ? R_AVG_INT : R_SUM_INT;
          o.out_type = avg ? KG_T_FLOAT64 : KG_T_INT64;
          o.st_type = avg ? KG_T_FLOAT64 : KG_T_INT64;
          // a SUM that wraps (not ANSI) keeps one word; AVG and the ANSI SUM
          // keep the exact 128-bit sum
          if (!avg && !chk && (merge || !avg_sums[c0]))
            o.word = op_words(OP_ADD64, c, false, false, 1, 0, st);
          else o.word = op_words(OP_ADD128, c, chk, false, 2, 0, st);
          o.nulls_word = nulls_of(c);
          if (chk) {
            o.st_word = st;
            int tw = merge ? 8 : type_width(t.type);
            add_checked({c, o.word, st, false, false, (unsigned __int128)1 << 63,
                         (unsigned __int128)1 << (8 * tw - 1)});
          }
        } else if (is_decimal(t.type)) {
          if (t.scale < 0) fail(KG_ERR_UNSUPPORTED, "negative decimal scale");
          int sum_p = t.precision + 10 > 38 ? 38 : t.precision + 10;
          // the sum state's carrier (decimal_on, aggregate/mod.rs:1004-1011)
          int sum_type = (t.type == KG_T_DECIMAL64 && sum_p <= 18) ? KG_T_DECIMAL64 : KG_T_DECIMAL128;
          bool wide = merge ? sum_type == KG_T_DECIMAL128 : t.type == KG_T_DECIMAL128;
          // (merged partial states of a SUM and an AVG are separate columns)
          o.word = op_words(wide ? OP_ADD192 : OP_ADD128, c, true, false, wide ? 3 : 2, 0, st);
          o.wide = wide;
          o.st_word = st;
          o.nulls_word = nulls_of(c);
          add_checked({c, o.word, st, avg, wide, pow10_u128(sum_p),
                       pow10_u128(merge ? sum_p : t.precision)});
          o.st_type = sum_type;
          o.st_p = sum_p;
          o.st_s = t.scale;
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
        if (merge) {
          next_state++;
          bool two = avg || is_decimal(t.type);
          if (two) {
            int c2 = next_state++;
            if (c2 >= kAggInputs) fail(KG_ERR_UNSUPPORTED, "at most 32 state columns per sink");
            if (avg) {
              // state (#sum, #count): the rows averaged are the summed counts
              o.cnt_word = words(1, 0);
              plain(OP_ADDV, c2, o.cnt_word);
              o.valid_word = o.cnt_word;
            } else {
              // state (#sum, #nonempty): NULL only when no partial was nonempty
              o.valid_word = words(1, 0);
              plain(OP_COUNT_TRUE, c2, o.valid_word);
            }
            if (is_decimal(t.type))
              ops.push_back(AggOp{OP_STICKY, (uint8_t)c, 0, (uint8_t)c2, 0, (uint16_t)st});
          }
        }
        if (partial) o.st2 = (avg ? 2 : (is_decimal(t.type) ? 1 : 0));
        break;
      }
      case KG_AGG_MIN:
      case KG_AGG_MAX: {
        bool mx = f == KG_AGG_MAX;
        o.out_type = o.st_type = t.type;
        o.out_p = o.st_p = t.precision;
        o.out_s = o.st_s = t.scale;
        o.nulls_word = nulls_of(c);
        if (is_float(t.type)) {
          o.kind = R_MINMAX_F;
          o.word = op_words(mx ? OP_MAX_F : OP_MIN_F, c, false, false, 1, mx ? 0 : ~0ull, st);
        } else if (is_int_like(t.type) && t.type != KG_T_UINT64) {
          o.kind = R_MINMAX_I;
          o.word = op_words(mx ? OP_MAX_I64 : OP_MIN_I64, c, false, false, 1,
                            mx ? 0x8000000000000000ULL : 0x7FFFFFFFFFFFFFFFULL, st);
        } else {
          fail(KG_ERR_UNSUPPORTED, "min/max of this type on the device");
        }
        if (merge) next_state++;
        break;
      }
      default: fail(KG_ERR_UNSUPPORTED, "aggregate function");
    }
  }
  // columns' ops adjacent, so a row loads each value once
  std::stable_sort(ops.begin() + 1, ops.end(),
                   [](const AggOp &x, const AggOp &y) { return x.col < y.col; });
  // W < kAggWords: the LDS stride W|1 indexes init[] within bounds
  if ((int)ops.size() > kAggOps || W >= kAggWords) fail(KG_ERR_UNSUPPORTED, "too many aggregates");
  P.nops = (int)ops.size();
  for (int k = 0; k < P.nops; k++) P.ops[k] = ops[k];
  P.W = W;
  for (int k = 0; k < W; k++) P.init[k] = init[k];
  O.W = W;
  O.ansi = ansi;
  return merge ? next_state : (int)in_types.size();
}

void init_records(kg_ctx *ctx, uint64_t *rec, uint64_t nrec, const AggProg &P, hipStream_t s) {
  (void)ctx;
  k_fill_words<<<stream_grid(nrec * P.W, 256), 256, 0, s>>>(rec, nrec, P.W, init_words(P));
  KG_HIP(hipGetLastError());
}

// The slots holding a group, in slot order: list[0..n) (flags, an exclusive
// scan, then each flagged slot writes itself at its position).
int64_t group_positions(kg_ctx *ctx, const uint64_t *rec, uint64_t ns, int W, bool global,
                        DevBuf &list, hipStream_t s, const uint64_t *keys) {
  DevBuf flags(ctx, ns * 4), pos(ctx, (ns + 1) * 8);
  if (keys)  // a hashed table of ns - 2 slots and the two fixed ones
    k_group_flags_keys<<<stream_grid(ns, 256), 256, 0, s>>>(keys, rec, ns - 2, W,
                                                            flags.get<uint32_t>());
  else
    k_group_flags<<<stream_grid(ns, 256), 256, 0, s>>>(rec, ns, W, global, flags.get<uint32_t>());
  KG_HIP(hipGetLastError());
  DevBuf temp(ctx, (uint64_t)scan_partials((int64_t)ns) * 8);
  exclusive_scan(flags.get<uint32_t>(), pos.get<uint64_t>(), (int64_t)ns, temp.get<uint64_t>(), s);
  PinnedReads pin;
  uint64_t *last_pos = pin.take<uint64_t>(1);
  uint32_t *last_flag = pin.take<uint32_t>(1);
  KG_HIP(hipMemcpyAsync(last_pos, pos.get<uint64_t>() + ns - 1, 8, hipMemcpyDeviceToHost, s));
  KG_HIP(hipMemcpyAsync(last_flag, flags.get<uint32_t>() + ns - 1, 4, hipMemcpyDeviceToHost, s));
  KG_HIP(hipStreamSynchronize(s));
  int64_t n = (int64_t)(*last_pos + *last_flag);
  list.reset(ctx, (uint64_t)n * 8 + 8);
  if (n > 0) {
    k_compact_slots<<<stream_grid(ns, 256), 256, 0, s>>>(flags.get<uint32_t>(), pos.get<uint64_t>(),
                                                         ns, list.get<uint64_t>());
    KG_HIP(hipGetLastError());
  }
  return n;  // list is ready in stream order (the scratch is freed behind it)
}

// Bits [off, off + n) of src to dst from bit 0 (a column's validity, or a
// BOOL column's values, at its offset).
__global__ void k_copy_key_bits(const uint8_t *src, int64_t off, int64_t n, uint32_t *dst) {
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w * 32 < n;
       w += (int64_t)gridDim.x * blockDim.x) {
    uint32_t bits = 0;
    for (int k = 0; k < 32 && w * 32 + k < n; k++) {
      const int64_t i = off + w * 32 + k;
      bits |= (uint32_t)((src[i >> 3] >> (i & 7)) & 1) << k;
    }
    dst[w] = bits;
  }
}

__global__ void k_iota(uint64_t *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (uint64_t)i;
}

// Each row as its own record: the initial state, then the row applied (the
// records of PartialAggregate's pass-through, AggTable::states_of_rows).
__global__ void k_rows_as_records(AggProg P, int64_t n, uint64_t *rec) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t *r = rec + (uint64_t)i * P.W;
    for (int w = 0; w < P.W; w++) r[w] = P.init[w];
    apply_row<false>(P, r, i);
  }
}

void launch_agg_output(const OutProg &O, const uint64_t *keys, int key_stride,
                       const uint64_t *rec, const uint64_t *list, int64_t n, void *out_key,
                       uint8_t *out_key_valid, int key_width, unsigned int *err, hipStream_t s) {
  if (n <= 0) return;
  ProfScope prof("agg_output", s);
  const int block = O.W > 31 ? 128 : 256;
  const size_t lds = (size_t)block * (O.W | 1) * 8;
  k_agg_output<<<stream_grid(n, block), block, lds, s>>>(O, keys, key_stride, rec, list, n,
                                                         out_key, out_key_valid, key_width, err);
  KG_HIP(hipGetLastError());
}

// The aggregate's output batch [key, aggregates (or partial states)] of the
// n records at rec + list[o] * W.  key_src: copy the key column from these
// rows instead of the records' keys (pass-through rows).
static std::unique_ptr<kg_batch> result_batch(kg_hash_agg *a, const uint64_t *keys,
                                              const uint64_t *rec, const uint64_t *list,
                                              int64_t n, uint64_t cap, const DCol *key_src) {
  {
    auto b = std::make_unique<kg_batch>();
    b->ctx = a->ctx;
    b->nrows = n;
    // validity words are not cleared: the output kernel (and the key copy)
    // writes every word that holds a bit below n
    uint64_t vb = (uint64_t)ceil_div(n, 32) * 4 + 4;
    OutProg O = a->out;
    O.cap = cap;
    int ncols = (a->has_key ? 1 : 0);
    for (int j = 0; j < a->out.n; j++) ncols += (a->partial && a->out.a[j].st2) ? 2 : 1;
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
      key_vals = kc.values.get();
      key_valid = kc.validity.get<uint8_t>();
      base = 1;
    }
    int at = base;
    std::vector<int> bool_cols;  // partial `nonempty` columns, written a byte per group
    for (int j = 0; j < a->out.n; j++) {
      OutAgg &o = O.a[j];
      auto &c = b->cols[at++];
      c.type = a->partial ? o.st_type : o.out_type;
      c.precision = a->partial ? o.st_p : o.out_p;
      c.scale = a->partial ? o.st_s : o.out_s;
      int w = type_width(c.type);
      c.values.reset(a->ctx, (uint64_t)n * (w > 0 ? w : 1) + 16);
      c.validity.reset(a->ctx, vb);
      o.values = c.values.get();
      o.validity = c.validity.get<uint8_t>();
      o.values2 = nullptr;
      o.validity2 = nullptr;
      if (a->partial && o.st2) {
        auto &c2 = b->cols[at];
        c2.type = o.st2 == 1 ? KG_T_BOOL : KG_T_INT64;
        c2.values.reset(a->ctx, (uint64_t)n * 8 + 16);
        c2.validity.reset(a->ctx, vb);
        o.values2 = c2.values.get();
        o.validity2 = c2.validity.get<uint8_t>();
        if (o.st2 == 1) bool_cols.push_back(at);
        at++;
      }
    }
    DevBuf err(a->ctx, 4);
    KG_HIP(hipMemsetAsync(err.get(), 0, 4, a->s));
    launch_agg_output(O, keys, 1, rec, list, n, key_src ? nullptr : key_vals,
                      key_src ? nullptr : key_valid, kw, err.get<unsigned int>(), a->s);
    if (key_src && n > 0) {  // the rows' own keys
      const int w = type_width(key_src->type);
      if (w == 0)
        k_copy_key_bits<<<stream_grid(ceil_div(n, 32), 256, 1), 256, 0, a->s>>>(
            static_cast<const uint8_t *>(key_src->values), key_src->offset, n,
            static_cast<uint32_t *>(key_vals));
      else
        KG_HIP(hipMemcpyAsync(key_vals, static_cast<const uint8_t *>(key_src->values) +
                                            key_src->offset * w,
                              (size_t)n * w, hipMemcpyDeviceToDevice, a->s));
      if (key_src->validity)
        k_copy_key_bits<<<stream_grid(ceil_div(n, 32), 256, 1), 256, 0, a->s>>>(
            key_src->validity, key_src->offset, n, reinterpret_cast<uint32_t *>(key_valid));
      else
        KG_HIP(hipMemsetAsync(key_valid, 0xFF, (size_t)ceil_div(n, 32) * 4, a->s));
      KG_HIP(hipGetLastError());
    }
    unsigned int e = 0;
    {
      PinnedReads pin;
      unsigned int *p = pin.take<unsigned int>(1);
      KG_HIP(hipMemcpyAsync(p, err.get(), 4, hipMemcpyDeviceToHost, a->s));
      KG_HIP(hipStreamSynchronize(a->s));
      e = *p;
    }
    if (e & 1u) fail(KG_ERR_OVERFLOW, "Overflow in sum (use try_sum/try_avg to tolerate overflow)");
    if (e & 2u) fail(KG_ERR_OUT_OF_RANGE, "average out of the result decimal's range");
    // booleans (a boolean key, partial `nonempty` states) leave the kernel as
    // one byte per group; Arrow packs them as bits
    if (a->has_key && a->key_type.type == KG_T_BOOL && !key_src) bool_cols.push_back(0);
    for (int c : bool_cols) {
      DevBuf bits(a->ctx, vb);
      k_bytes_to_bits<<<stream_grid(ceil_div(n, 32), 256, 1), 256, 0, a->s>>>(
          b->cols[c].values.get<uint8_t>(), n, bits.get<uint32_t>());
      KG_HIP(hipGetLastError());
      b->cols[c].values = std::move(bits);
    }
    KG_HIP(hipStreamSynchronize(a->s));
    return b;
  }
}


// ---- PartialAggregate: flushes and pass-through (aggregate/partial.rs) -----
static void reset_table(kg_hash_agg *a) {
  alloc_table(a, a->cap, a->keys, a->rec);
  KG_HIP(hipMemsetAsync(a->occ.get(), 0, 16, a->s));
  if (a->hashed) {
    KG_HIP(hipMemsetAsync(a->sp_cnt.get(), 0, 24, a->s));
    KG_HIP(hipMemsetAsync(a->hll.get(), 0, (size_t)kHllRegs * 4, a->s));
  }
  a->occ_bound = a->merged_keys = 0;
  a->bound_rows = a->bound_recs = 0;
  a->finalized = false;
  a->ngroups = 0;
}

// Output bytes per group: the key and every state column with validity.
static uint64_t out_row_bytes(const kg_hash_agg *a) {
  uint64_t b = 8 + 1;
  for (int j = 0; j < a->out.n; j++) b += (a->out.a[j].st2 ? 2 : 1) * (16 + 1);
  return b;
}

// Flush (PartialAggregate::start_drain): the table's partial states are
// emitted and the table starts empty.  Emitting allocates its output first:
// refused for memory, the table stays as it was and false is returned.
static bool emit_table(kg_hash_agg *a) {
  drain(a);
  const int64_t groups = group_positions(a->ctx, a->rec.get<uint64_t>(), nslots(a), a->prog.W,
                                         !a->has_key, a->pos, a->s,
                                         a->hashed ? a->keys.get<uint64_t>() : nullptr);
  if (groups > 0) {
    try {
      a->emitted.push_back(result_batch(a, a->keys.get<uint64_t>(), a->rec.get<uint64_t>(),
                                        a->pos.get<uint64_t>(), groups, a->cap, nullptr)
                               .release());
    } catch (const Error &e) {
      if (e.code != KG_ERR_MEMORY) throw;
      return false;
    }
    a->flushed_groups += (uint64_t)groups;
  }
  a->flushes++;
  reset_table(a);
  for (auto &c : a->checked) c.cum_abs = 0;  // the sums start again from empty groups
  return true;
}

// BypassPolicy after a batch is folded (partial.rs:246-262): groups per 100
// rows, flushed and held, at or above max_group_percent once min_rows rows
// are seen switch the aggregate to pass-through (the table flushed first);
// otherwise a table whose groups' records pass max_table_bytes is flushed.
static void after_partial_batch(kg_hash_agg *a, uint64_t nrows) {
  a->part_rows += nrows;
  drain(a);
  const uint64_t held = a->direct ? (uint64_t)group_positions(a->ctx, a->rec.get<uint64_t>(),
                                                               nslots(a), a->prog.W, false,
                                                               a->pos, a->s)
                                  : read_occupied(a);
  const uint64_t groups = a->flushed_groups + held;
  if (a->part_rows >= a->pol_min_rows && groups * 100 >= a->part_rows * a->pol_pct) {
    if (emit_table(a)) a->bypass = true;
    return;
  }
  if (held * ((uint64_t)a->prog.W * 8 + 8) > a->pol_max_bytes) (void)emit_table(a);
}

// Pass-through: every row of the batch as its own partial state.
static void emit_rows(kg_hash_agg *a, const AggProg &P, int64_t n) {
  DevBuf rec(a->ctx, (uint64_t)n * a->prog.W * 8 + 8), list(a->ctx, (uint64_t)n * 8 + 8);
  k_rows_as_records<<<stream_grid(n, 256), 256, 0, a->s>>>(P, n, rec.get<uint64_t>());
  k_iota<<<stream_grid(n, 256), 256, 0, a->s>>>(list.get<uint64_t>(), n);
  KG_HIP(hipGetLastError());
  const DCol *ks = a->has_key ? &P.key : nullptr;
  // cap past every record index: no record is the NULL-key slot
  a->emitted.push_back(result_batch(a, nullptr, rec.get<uint64_t>(), list.get<uint64_t>(), n,
                                    (uint64_t)n + 16, ks)
                           .release());
  a->bypassed_rows += (uint64_t)n;
}

}  // namespace kg

extern "C" {

int kg_agg_set_bypass(kg_hash_agg *a, uint64_t min_rows, uint32_t max_group_percent,
                      uint64_t max_table_bytes) {
  return guarded([&] {
    if (!a->partial) fail(KG_ERR_INTERNAL, "a bypass policy needs a partial aggregate");
    if (a->rows_seen || (a->gather && a->gather->rows()))
      fail(KG_ERR_INTERNAL, "the bypass policy is set before the first sink");
    if (!a->has_key) fail(KG_ERR_UNSUPPORTED, "a bypass policy for a global aggregate");
    a->has_policy = true;
    a->pol_min_rows = min_rows;
    a->pol_pct = max_group_percent;
    a->pol_max_bytes = max_table_bytes;
  });
}

int kg_agg_partial_next(kg_hash_agg *a, kg_batch **out, int32_t *done) {
  return guarded([&] {
    *out = nullptr;
    *done = 1;
    if (a->emitted_next < a->emitted.size()) {
      *out = a->emitted[a->emitted_next];
      a->emitted[a->emitted_next++] = nullptr;
    }
    if (a->emitted_next >= a->emitted.size()) {
      a->emitted.clear();
      a->emitted_next = 0;
    }
    *done = a->emitted.empty() ? 1 : 0;
  });
}

int kg_agg_partial_stats(const kg_hash_agg *a, uint64_t *flushes, uint64_t *bypassed_rows) {
  return guarded([&] {
    if (flushes) *flushes = a->flushes;
    if (bypassed_rows) *bypassed_rows = a->bypassed_rows;
  });
}

int kg_agg_create(kg_ctx *ctx, const kg_column *key_type, int32_t ninputs,
                  const kg_column *input_types, int32_t naggs, const kg_agg_spec *aggs,
                  uint32_t flags, uint64_t expected_groups, kg_hash_agg **out) {
  return guarded([&] {
    *out = nullptr;
    if (ninputs < 0 || ninputs > kAggInputs) fail(KG_ERR_UNSUPPORTED, "at most 32 input columns");
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
    a->merge = (flags & KG_AGG_MERGE) != 0;
    a->partial = (flags & KG_AGG_PARTIAL) != 0;
    a->nsink = compile_aggs(a->in_types, naggs, aggs, a->ansi, a->merge, a->partial, a->prog,
                            a->out, a->checked);
    a->prog.has_key = a->has_key;
    a->prog.direct = a->direct;
    a->prog.key.type = a->has_key ? key_type->type : 0;
    a->out.key_type = a->prog.key.type;
    a->out.direct = a->direct;
    a->out.global = !a->has_key;
    a->s = make_stream(ctx->device);
    StreamScope kg_scope(a->s);
    int W = a->prog.W;
    a->hashed = a->has_key && !a->direct;
    if (a->hashed) {
      // one 1,024-lane workgroup a CU with the whole LDS: one table of the
      // heavy keys a CU (two 512-lane ones would each hold the same top keys)
      const int WL = W | 1;
      int p1_optin = 0;
      if (hipDeviceGetAttribute(&p1_optin, hipDeviceAttributeSharedMemPerBlockOptin,
                                ctx->device) != hipSuccess || p1_optin < 64 * 1024) {
        (void)hipGetLastError();
        p1_optin = 64 * 1024;
      }
      const size_t budget = (size_t)std::min(p1_optin, 160 * 1024) - 1024 - 4 * kDoorWords;
      a->p1_slots = (int)((budget / 8 - 2 * (size_t)WL) / (1 + (size_t)WL));
      // the merge: one workgroup a CU with as much LDS as a workgroup may take
      int optin = 0;
      if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, ctx->device) !=
              hipSuccess || optin < 64 * 1024) {
        (void)hipGetLastError();
        optin = 64 * 1024;
      }
      a->m_slots = (int)(((size_t)std::min(optin, 160 * 1024) - 1024) / 8 / (1 + (size_t)WL));
      if (a->p1_slots < 64 || a->m_slots < 64)
        fail(KG_ERR_UNSUPPORTED, "aggregate state too wide for the LDS table");
      a->p1_lds = ((size_t)a->p1_slots + (size_t)(a->p1_slots + 2) * WL) * 8 + 4 * kDoorWords;
      a->m_lds = (size_t)a->m_slots * (1 + WL) * 8;
      // windows of the windowed merge: the largest power of two of slots
      // (key + record) within the merge's LDS
      while ((2ull << a->log2span_max) <= (uint64_t)a->m_slots) a->log2span_max++;
      a->w_lds_max = (size_t)std::min(optin, 160 * 1024) - 1024;
      const std::pair<const void *, size_t> fns[4] = {
          {(const void *)k_agg_pass1<kTileRows>, a->p1_lds},
          {(const void *)k_agg_merge<kMergeRows>, a->m_lds},
          {(const void *)k_agg_wmerge<kWMergeRows, false>, a->w_lds_max},
          {(const void *)k_agg_wmerge<kWMergeRows, true>, a->w_lds_max}};
      // spill partition histograms (and the sketch's registers) and cursors:
      // 4 bytes a partition, up to 2^15
      for (const void *f : {(const void *)k_spill_hist, (const void *)k_spill_scatter})
        if (set_dynamic_lds(f, 128 * 1024 + 4 * kHllRegs) != hipSuccess) {
          (void)hipGetLastError();
          fail(KG_ERR_DEVICE, "dynamic LDS of 132 KiB refused");
        }
      for (auto &f : fns) {
        hipError_t e = set_dynamic_lds(f.first, (int)f.second);
        if (e != hipSuccess) {
          (void)hipGetLastError();
          fail(KG_ERR_DEVICE, "dynamic LDS of " + std::to_string(f.second) + " bytes refused");
        }
      }
      a->sp_cnt.reset(ctx, 32);
      KG_HIP(hipMemsetAsync(a->sp_cnt.get(), 0, 32, a->s));
      a->hll.reset(ctx, (uint64_t)kHllRegs * 4);
      KG_HIP(hipMemsetAsync(a->hll.get(), 0, (size_t)kHllRegs * 4, a->s));
      a->spill_budget = std::min<uint64_t>(16ull << 30, ctx->total_hbm / 16);
      if (ctx->budget) a->spill_budget = std::min<uint64_t>(a->spill_budget, ctx->budget / 4);
      if (a->spill_budget < (64u << 20)) a->spill_budget = 64u << 20;
    }
    uint64_t eg = expected_groups < 512 ? 512 : expected_groups;
    a->cap = a->direct ? 256 : (1ull << log2_ceil(eg + eg / 4));
    a->log2cap = log2_ceil(a->cap);
    a->occ.reset(ctx, 16);
    KG_HIP(hipMemsetAsync(a->occ.get(), 0, 16, a->s));
    alloc_table(a.get(), a->cap, a->keys, a->rec);
    *out = a.release();
  });
}

int kg_agg_set_merge(kg_hash_agg *a, int32_t form, int32_t size_from_sketch) {
  return guarded([&] {
    if (form < KG_AGG_MERGE_AUTO || form > KG_AGG_MERGE_BY_KEY) fail(KG_ERR_INTERNAL, "merge form");
    if (a->rows_seen || (a->gather && a->gather->rows()))
      fail(KG_ERR_INTERNAL, "the merge form is set before the first sink");
    a->merge_form = form;
    a->size_from_sketch = size_from_sketch != 0;
  });
}

int kg_agg_sink(kg_hash_agg *a, const kg_column *key, const kg_column *inputs, int64_t nrows) {
  return guarded([&] { StreamScope kg_scope(a->s); sink_or_gather(a, key, inputs, nrows); });
}

int kg_agg_finalize(kg_hash_agg *a) {
  return guarded([&] {
    StreamScope kg_scope(a->s);
    if (a->finalized) return;
    flush_gather(a);
    drain(a);
    a->ngroups = group_positions(a->ctx, a->rec.get<uint64_t>(), nslots(a), a->prog.W,
                                 !a->has_key, a->pos, a->s,
                                 a->hashed ? a->keys.get<uint64_t>() : nullptr);
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
    StreamScope kg_scope(a->s);
    *out = nullptr;
    if (!a->finalized) fail(KG_ERR_INTERNAL, "result before finalize");
    *out = result_batch(a, a->keys.get<uint64_t>(), a->rec.get<uint64_t>(), a->pos.get<uint64_t>(),
                        a->ngroups, a->cap, nullptr)
               .release();
  });
}

void kg_agg_destroy(kg_hash_agg *a) {
  if (!a) return;
  if (a->s) (void)hipStreamSynchronize(a->s);
  for (kg_batch *b : a->emitted) kg_batch_release(b);  // batches never taken
  hipStream_t s = a->s;
  delete a;
  destroy_stream(s);
}

}  // extern "C"
