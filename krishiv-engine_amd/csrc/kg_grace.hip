// kg_grace.hip -- the grace hash join (join/grace.rs:99-1031) fused with the
// per-partition aggregate of the plan in include/kgpu.h (kg_grace_*).
//
//  sink     H_kg -> partition_of -> LDS-staged radix scatter (k_gp_hist,
//           k_gp_scatter): a workgroup sorts each tile of rows by partition
//           in LDS and writes every partition's run with consecutive lanes;
//           rows become [key, payload] pairs (NULL keys, which join nothing,
//           are dropped; a NULL payload is stored as -1, never a payload
//           the finish accepts; the mod 2^31 reading keeps only each
//           payload's low 31 bits, on which it alone depends).  The histogram pass also takes the payload
//           range for the factorization check.  Each sunk batch stays
//           partitioned in HBM, or, once the context's HBM budget would be
//           passed, in pinned host memory (the spill).
//  finish   per partition, as one launch over all its pieces each: the
//           build rows aggregated per key straight into the partition's
//           lookup table (CAS insert, then one 64-bit atomic for the row
//           and non-NULL payload counts and one for the payload sum), then
//           the probe rows streamed against it.  A probe row with key k adds
//           nb_k pairs and pay * Sb_k to its partition (wave-reduced, one
//           atomic per wave), the factorized form of the pair walk
//           (join/probe.rs:266-337) for a SUM of payload products.  The
//           partitions run back to back on the operator's stream, reusing
//           one table buffer; the host waits once, at the end.
#include <hip/hip_runtime.h>


#include <climits>

#include "kg_common.hpp"
#include "kg_scan.hpp"
#include "kg_device.hpp"

namespace kg {

struct GSlot {
  uint64_t key;
  uint64_t cnt;  // build rows with the key (low 32 bits), of which with a
                 // non-NULL payload (high 32 bits)
  uint64_t sb;   // their payload sum (wrapping)
  uint64_t pad;
};

// A partition's rows of one sunk piece (device columns).
struct GPiece {
  DCol key, pay;
  int64_t n;
};

constexpr uint64_t kGEmpty = 0x8000000000000000ULL;

// Validity of row i in two steps: the byte (0xFF without a bitmap), loaded
// with the row's other words, then its bit once every load is issued.
__device__ __forceinline__ uint32_t vbyte(const DCol &c, int64_t i) {
  return c.validity ? c.validity[(i + c.offset) >> 3] : 0xFFu;
}
__device__ __forceinline__ bool vbit(const DCol &c, int64_t i, uint32_t b) {
  return (b >> ((i + c.offset) & 7)) & 1;
}

__device__ __forceinline__ uint64_t gslot_start(uint64_t key, int log2cap) {
  return dev::table_hash(key, 0x51ED27A1D3C4B9E5ULL) >> (64 - log2cap);
}

// Adds (row and non-NULL payload counts, payload sum) to key k's slot,
// inserting the key first when it is absent.
__device__ __forceinline__ void grace_add(GSlot *slots, uint64_t mask, int log2cap, uint64_t cap,
                                          uint64_t k, unsigned long long c,
                                          unsigned long long sb) {
  uint64_t s;
  if (k == kGEmpty) {
    s = cap;
  } else {
    s = gslot_start(k, log2cap);
    while (true) {
      const uint64_t cur = __hip_atomic_load(&slots[s].key, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
      if (cur == k) break;
      if (cur == kGEmpty) {
        const unsigned long long prev = atomicCAS((unsigned long long *)&slots[s].key, kGEmpty, k);
        if (prev == kGEmpty || prev == k) break;
      }
      s = (s + 1) & mask;
    }
  }
  if (c) atomicAdd((unsigned long long *)&slots[s].cnt, c);
  if (sb) atomicAdd((unsigned long long *)&slots[s].sb, sb);
}

// The build pieces of a partition aggregated per key into its table (a
// partition holds fewer than 2^32 build rows, so both counts fit one word).
// A workgroup first combines each tile of kGBTile rows in an LDS table of
// kGBSlots keys, then adds each of its keys to the global table once: a
// Zipf-hot key (a quarter of some partition's build rows at s = 0.8) costs
// one global atomic per tile instead of one per row, which the atomics on
// one address would serialize.  Rows whose key finds no LDS slot within
// kGBProbe steps go to the global table directly.
constexpr int kGBBlock = 512;
constexpr int kGBTile = 2048;
constexpr int kGBSlots = 4096;
constexpr int kGBProbe = 16;

__global__ void __launch_bounds__(kGBBlock) k_grace_build(const GPiece *pcs, int npcs, GSlot *slots,
                                                          uint64_t mask, int log2cap,
                                                          uint64_t cap) {
  extern __shared__ unsigned long long g_lds[];
  unsigned long long *s_key = g_lds;               // kGBSlots (kGEmpty: free)
  unsigned long long *s_cnt = s_key + kGBSlots;    // kGBSlots
  unsigned long long *s_sb = s_cnt + kGBSlots;     // kGBSlots
  for (int l = threadIdx.x; l < kGBSlots; l += kGBBlock) {
    s_key[l] = kGEmpty;
    s_cnt[l] = 0;
    s_sb[l] = 0;
  }
  __syncthreads();
  for (int q = 0; q < npcs; q++) {
    const DCol key = pcs[q].key, pay = pcs[q].pay;
    const int64_t n = pcs[q].n;
    const int64_t ntiles = (n + kGBTile - 1) / kGBTile;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
      // the tile's loads are all issued before the first LDS insert (clamped
      // row index; pieces are int64 key and payload columns, see slice)
      constexpr int kR = kGBTile / kGBBlock;
      uint64_t kr[kR];
      int64_t pw[kR];
      uint32_t kb[kR], pb[kR];
#pragma unroll
      for (int r = 0; r < kR; r++) {
        const int64_t i = t * kGBTile + (int64_t)r * kGBBlock + threadIdx.x, ic = i < n ? i : n - 1;
        kr[r] = dev::load_w64(key, ic);
        kb[r] = vbyte(key, ic);
        pw[r] = (int64_t)dev::load_w64(pay, ic);
        pb[r] = vbyte(pay, ic);
      }
#pragma unroll
      for (int r = 0; r < kR; r++) {
        const int64_t i = t * kGBTile + (int64_t)r * kGBBlock + threadIdx.x, ic = i < n ? i : n - 1;
        // a NULL key joins nothing
        if (i < n && vbit(key, ic, kb[r])) {
          const uint64_t k = kr[r];
          const int64_t pval = pw[r];
          const bool pv = vbit(pay, ic, pb[r]) && pval >= 0;  // -1: a NULL payload
          const unsigned long long c = 1ull | ((unsigned long long)pv << 32);
          const unsigned long long sb = pv ? (unsigned long long)pval : 0ull;
          int l = -1;
          if (k != kGEmpty) {
            int h = (int)((dev::table_hash(k, 0x9E3779B97F4A7C15ULL) >> 32) & (kGBSlots - 1));
            for (int pr = 0; pr < kGBProbe; pr++) {
              const unsigned long long cur = s_key[h];
              if (cur == k) {
                l = h;
                break;
              }
              if (cur == kGEmpty) {
                const unsigned long long prev = atomicCAS(&s_key[h], kGEmpty, k);
                if (prev == kGEmpty || prev == k) {
                  l = h;
                  break;
                }
              }
              h = (h + 1) & (kGBSlots - 1);
            }
          }
          if (l >= 0) {
            atomicAdd(&s_cnt[l], c);
            if (sb) atomicAdd(&s_sb[l], sb);
          } else {
            grace_add(slots, mask, log2cap, cap, k, c, sb);
          }
        }
      }
      __syncthreads();
      // flush: the home-slot loads of a lane's kGBSlots / kGBBlock keys are
      // all issued before the first insert
      constexpr int kF = kGBSlots / kGBBlock;
      uint64_t fk[kF], fs[kF], fcur[kF];
#pragma unroll
      for (int r = 0; r < kF; r++) {
        fk[r] = s_key[threadIdx.x + r * kGBBlock];
        fs[r] = gslot_start(fk[r], log2cap);
        fcur[r] = fk[r] != kGEmpty ? __hip_atomic_load(&slots[fs[r]].key, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT)
                                   : kGEmpty;
      }
#pragma unroll
      for (int r = 0; r < kF; r++) {
        const int l = threadIdx.x + r * kGBBlock;
        if (fk[r] != kGEmpty) {
          uint64_t at = fs[r], cur = fcur[r];
          while (cur != fk[r]) {
            if (cur == kGEmpty) {
              const unsigned long long prev =
                  atomicCAS((unsigned long long *)&slots[at].key, kGEmpty, fk[r]);
              if (prev == kGEmpty || prev == fk[r]) break;
            }
            at = (at + 1) & mask;
            cur = __hip_atomic_load(&slots[at].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          if (s_cnt[l]) atomicAdd((unsigned long long *)&slots[at].cnt, s_cnt[l]);
          if (s_sb[l]) atomicAdd((unsigned long long *)&slots[at].sb, s_sb[l]);
        }
        s_key[l] = kGEmpty;
        s_cnt[l] = 0;
        s_sb[l] = 0;
      }
      __syncthreads();
    }
  }
}

__global__ void k_grace_init(GSlot *slots, uint64_t n) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < n;
       s += (uint64_t)gridDim.x * blockDim.x)
    slots[s] = GSlot{kGEmpty, 0, 0, 0};
}

// The probe pieces of a partition against its table; out[0..2] of the
// partition: pairs, pairs with both payloads non-NULL, payload-product sum
// (all wrapping u64).  kGPU rows a lane: their key and payload loads, then
// their home slots (32 bytes: key, counts, sum), are issued before any is
// used.  The NULL-sentinel key lives in slot `cap`, whose key field stays
// EMPTY; a slot with no rows adds nothing, so no count test is needed.
constexpr int kGPU = 4;

__global__ void __launch_bounds__(256) k_grace_probe(const GPiece *pcs, int npcs,
                                                     const GSlot *slots, uint64_t mask,
                                                     int log2cap, uint64_t cap,
                                                     unsigned long long *out) {
  uint64_t cnt = 0, vcnt = 0, sum = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * kGPU;
  for (int q = 0; q < npcs; q++) {
    const DCol key = pcs[q].key, pay = pcs[q].pay;
    const int64_t n = pcs[q].n;
    for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x * kGPU + threadIdx.x; i0 < n;
         i0 += stride) {
      uint64_t k[kGPU], at[kGPU], sk[kGPU], sc[kGPU], ss[kGPU];
      int64_t pv[kGPU];
      bool act[kGPU];
      uint32_t kb[kGPU], pb[kGPU];
      // int64 piece columns (see slice): loads on a clamped row index, the
      // validity bits tested once all are issued
#pragma unroll
      for (int u = 0; u < kGPU; u++) {
        const int64_t i = i0 + (int64_t)u * blockDim.x, ic = i < n ? i : n - 1;
        k[u] = dev::load_w64(key, ic);
        kb[u] = vbyte(key, ic);
        pv[u] = (int64_t)dev::load_w64(pay, ic);
        pb[u] = vbyte(pay, ic);
      }
#pragma unroll
      for (int u = 0; u < kGPU; u++) {
        const int64_t i = i0 + (int64_t)u * blockDim.x, ic = i < n ? i : n - 1;
        act[u] = i < n && vbit(key, ic, kb[u]);
        if (!act[u]) k[u] = 0;
        if (!(act[u] && vbit(pay, ic, pb[u]))) pv[u] = -1;  // -1: a NULL payload
      }
#pragma unroll
      for (int u = 0; u < kGPU; u++) {
        at[u] = k[u] == kGEmpty ? cap : gslot_start(k[u], log2cap);
        const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(slots + at[u]);
        const ulonglong2 a = act[u] ? p[0] : make_ulonglong2(kGEmpty, 0);
        const ulonglong2 b = act[u] ? p[1] : make_ulonglong2(0, 0);
        sk[u] = a.x;
        sc[u] = a.y;
        ss[u] = b.x;
      }
#pragma unroll
      for (int u = 0; u < kGPU; u++) {
        while (act[u] && sk[u] != k[u] && sk[u] != kGEmpty) {
          at[u] = (at[u] + 1) & mask;
          const GSlot t = slots[at[u]];
          sk[u] = t.key;
          sc[u] = t.cnt;
          ss[u] = t.sb;
        }
        if (act[u] && sk[u] == k[u]) {
          cnt += sc[u] & 0xFFFFFFFFull;
          if (pv[u] >= 0) {
            vcnt += sc[u] >> 32;
            sum += (uint64_t)pv[u] * ss[u];
          }
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_xor(cnt, off);
    vcnt += __shfl_xor(vcnt, off);
    sum += __shfl_xor(sum, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (cnt) atomicAdd(out, (unsigned long long)cnt);
    if (vcnt) atomicAdd(out + 1, (unsigned long long)vcnt);
    if (sum) atomicAdd(out + 2, (unsigned long long)sum);
  }
}

// ---------------------------------------------------------------- sink
// Partition pass 1: per block (contiguous rows) the rows of each partition
// (NULL keys dropped), in wave-private LDS counters; and the range of the
// non-NULL payloads of rows with a key, one atomic per block.
constexpr int kGPBlock = 256;

// W64: the key is one raw 64-bit word hashed as such (int64, uint64,
// temporal) and the payload one raw 64-bit integer word: eight rows a lane
// with every load issued before any is tested (clamped row index, validity
// bytes kept raw until then).

template <bool W64>
__global__ void __launch_bounds__(kGPBlock) k_gp_hist(DCol key, DCol pay, int64_t n, uint32_t parts,
                                                      int64_t per_block, uint32_t *hist,
                                                      long long *mm) {
  extern __shared__ uint32_t s_h[];  // [4][parts]
  __shared__ long long s_lo[kGPBlock / 64], s_hi[kGPBlock / 64];
  const int wave = threadIdx.x >> 6;
  for (uint32_t p = threadIdx.x; p < 4 * parts; p += kGPBlock) s_h[p] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = lo + per_block < n ? lo + per_block : n;
  long long vlo = LLONG_MAX, vhi = LLONG_MIN;
  uint32_t *h = s_h + wave * parts;
  if (W64) {
    constexpr int U = 8;
    for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (int64_t)kGPBlock * U) {
      uint64_t kv[U], pv[U];
      uint32_t kb[U], pb[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t i = i0 + (int64_t)u * kGPBlock, ic = i < hi ? i : hi - 1;
        kv[u] = dev::load_w64(key, ic);
        kb[u] = vbyte(key, ic);
        pv[u] = dev::load_w64(pay, ic);
        pb[u] = vbyte(pay, ic);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t i = i0 + (int64_t)u * kGPBlock, ic = i < hi ? i : hi - 1;
        if (i < hi && vbit(key, ic, kb[u])) {
          atomicAdd(&h[dev::partition_of(dev::xxh64_fixed(kv[u], 0, 8), parts)], 1u);
          if (vbit(pay, ic, pb[u])) {
            const long long v = (long long)pv[u];
            vlo = v < vlo ? v : vlo;
            vhi = v > vhi ? v : vhi;
          }
        }
      }
    }
  } else {
    for (int64_t i = lo + threadIdx.x; i < hi; i += kGPBlock) {
      if (!dev::valid(key, i)) continue;
      atomicAdd(&h[dev::partition_of(dev::value_hash(key, i), parts)], 1u);
      if (dev::valid(pay, i)) {
        const long long v = (long long)dev::load_i64(pay, i);
        vlo = v < vlo ? v : vlo;
        vhi = v > vhi ? v : vhi;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const long long a = __shfl_xor(vlo, off), b = __shfl_xor(vhi, off);
    vlo = a < vlo ? a : vlo;
    vhi = b > vhi ? b : vhi;
  }
  if ((threadIdx.x & 63) == 0) {
    s_lo[wave] = vlo;
    s_hi[wave] = vhi;
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < parts; p += kGPBlock)
    hist[(uint64_t)p * gridDim.x + blockIdx.x] = s_h[p] + s_h[parts + p] + s_h[2 * parts + p] +
                                                 s_h[3 * parts + p];
  if (threadIdx.x == 0) {
    for (int w = 1; w < kGPBlock / 64; w++) {
      vlo = s_lo[w] < vlo ? s_lo[w] : vlo;
      vhi = s_hi[w] > vhi ? s_hi[w] : vhi;
    }
    if (vlo <= vhi) {
      atomicMin(mm, vlo);
      atomicMax(mm + 1, vhi);
    }
  }
}

// Exclusive scan of c[0..np) into s in LDS (np <= 8 * kGSBlock = 4096, the
// most partitions); returns the total.
constexpr int kGSBlock = 512;

__device__ __forceinline__ uint32_t gs_block_scan(const uint32_t *c, uint32_t *s, uint32_t np,
                                                  uint32_t *wsum) {
  constexpr int PER = 8;
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
  for (int w = 0; w < kGSBlock / 64; w++) {
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

__host__ __device__ constexpr size_t gs_lds_bytes(int R, uint32_t parts) {
  return (size_t)kGSBlock * R * 16 + (size_t)parts * 16 + (size_t)kGSBlock * R * 2 +
         kGSBlock / 64 * 4;
}

// Partition pass 2: tiles of kGSBlock * R rows sorted by partition in LDS,
// then each partition's run written by consecutive lanes: out_key/out_pay
// rows [key, payload or -1].
template <int R, bool W64>
__global__ void __launch_bounds__(kGSBlock) k_gp_scatter(DCol key, DCol pay, int64_t n,
                                                         uint32_t parts, int64_t per_block,
                                                         const uint64_t *base, int64_t *out_key,
                                                         int64_t *out_pay, uint64_t pmask) {
  constexpr int TT = kGSBlock * R;
  extern __shared__ uint64_t smem[];
  uint64_t *st_k = smem;                                       // TT
  uint64_t *st_v = st_k + TT;                                  // TT
  uint64_t *cur = st_v + TT;                                   // parts
  uint32_t *cnt = reinterpret_cast<uint32_t *>(cur + parts);   // parts
  uint32_t *start = cnt + parts;                               // parts
  uint16_t *st_p = reinterpret_cast<uint16_t *>(start + parts);  // TT
  uint32_t *wsum = reinterpret_cast<uint32_t *>(st_p + TT);      // kGSBlock / 64
  for (uint32_t p = threadIdx.x; p < parts; p += kGSBlock)
    cur[p] = base[(uint64_t)p * gridDim.x + blockIdx.x];
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = lo + per_block < n ? lo + per_block : n;
  for (int64_t t0 = lo; t0 < hi; t0 += TT) {
    for (uint32_t p = threadIdx.x; p < parts; p += kGSBlock) cnt[p] = 0;
    __syncthreads();
    uint64_t kk[R], vv[R];
    uint32_t pp[R], lp[R];
    if (W64) {
      uint32_t kb[R], pb[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t i = t0 + (int64_t)r * kGSBlock + threadIdx.x, ic = i < hi ? i : hi - 1;
        kk[r] = dev::load_w64(key, ic);
        kb[r] = vbyte(key, ic);
        vv[r] = dev::load_w64(pay, ic);
        pb[r] = vbyte(pay, ic);
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t i = t0 + (int64_t)r * kGSBlock + threadIdx.x, ic = i < hi ? i : hi - 1;
        const bool live = i < hi && vbit(key, ic, kb[r]);
        pp[r] = live ? dev::partition_of(dev::xxh64_fixed(kk[r], 0, 8), parts) : 0xffffffffu;
        vv[r] = live ? (vbit(pay, ic, pb[r]) ? vv[r] & pmask : ~0ull) : 0;
        kk[r] = live ? kk[r] : 0;
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t i = t0 + (int64_t)r * kGSBlock + threadIdx.x;
        const bool live = i < hi && dev::valid(key, i);
        kk[r] = live ? (uint64_t)dev::load_i64(key, i) : 0;
        vv[r] = live ? (dev::valid(pay, i) ? (uint64_t)dev::load_i64(pay, i) & pmask : ~0ull) : 0;
        pp[r] = live ? dev::partition_of(dev::value_hash(key, i), parts) : 0xffffffffu;
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) lp[r] = pp[r] != 0xffffffffu ? atomicAdd(&cnt[pp[r]], 1u) : 0u;
    __syncthreads();
    const uint32_t tn = gs_block_scan(cnt, start, parts, wsum);
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (pp[r] == 0xffffffffu) continue;
      const uint32_t sl = start[pp[r]] + lp[r];
      st_k[sl] = kk[r];
      st_v[sl] = vv[r];
      st_p[sl] = (uint16_t)pp[r];
    }
    __syncthreads();
    for (uint32_t sl = threadIdx.x; sl < tn; sl += kGSBlock) {
      const uint32_t p = st_p[sl];
      const uint64_t d = cur[p] + (sl - start[p]);
      out_key[d] = (int64_t)st_k[sl];
      out_pay[d] = (int64_t)st_v[sl];
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < parts; p += kGSBlock) cur[p] += cnt[p];
    __syncthreads();
  }
}

// Partition p's first row: base[p * nb] (the scan of the partition-major
// histogram), and the row count at offs[parts].
__global__ void k_gp_offsets(const uint64_t *base, const uint32_t *hist, uint64_t hn, uint32_t parts,
                             uint32_t nb, int64_t *offs) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= parts; p += gridDim.x * blockDim.x)
    offs[p] = p < parts ? (int64_t)base[(uint64_t)p * nb] : (int64_t)(base[hn - 1] + hist[hn - 1]);
}

static void ck(int rc) {
  if (rc != KG_OK) throw Error{rc, std::string(kg_last_error())};
}

static int glog2(uint64_t v) {
  int b = 0;
  while ((1ull << b) < v) b++;
  return b;
}

}  // namespace kg

using namespace kg;

struct kg_grace {
  kg_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  uint32_t parts = 64;
  uint64_t table_budget = 0;
  bool mod31 = false;  // SUM = ((sum of products) mod 2^64) mod 2^31
  // one partitioned piece per sunk batch: [key], [payload or -1 for NULL]
  // in HBM, or their copies in pinned host memory (hipHostMalloc)
  struct Piece {
    DevBuf key, pay;
    int64_t *hk = nullptr, *hp = nullptr;
    bool on_device = false;
    int64_t rows = 0;
    std::vector<int64_t> off;
  };
  std::vector<Piece> build, probe;
  DevBuf mm;  // [build min, build max, probe min, probe max]
  uint64_t spilled = 0;
};

namespace kg {

static void free_piece(kg_grace::Piece &p) {
  p.key.release();
  p.pay.release();
  p.on_device = false;
  for (int64_t *q : {p.hk, p.hp})
    if (q) (void)hipHostFree(q);
  p.hk = p.hp = nullptr;
}

// Moves a partitioned piece from HBM to pinned host memory.
static void spill_piece(kg_grace *g, kg_grace::Piece &piece) {
  if (!piece.on_device) return;
  const int64_t n = piece.rows;
  KG_HIP(hipHostMalloc((void **)&piece.hk, (size_t)n * 8 + 8, hipHostMallocDefault));
  KG_HIP(hipHostMalloc((void **)&piece.hp, (size_t)n * 8 + 8, hipHostMallocDefault));
  if (n) {
    KG_HIP(hipMemcpyAsync(piece.hk, piece.key.get(), (size_t)n * 8, hipMemcpyDeviceToHost, g->s));
    KG_HIP(hipMemcpyAsync(piece.hp, piece.pay.get(), (size_t)n * 8, hipMemcpyDeviceToHost, g->s));
  }
  KG_HIP(hipStreamSynchronize(g->s));
  piece.key.release();
  piece.pay.release();
  piece.on_device = false;
  g->spilled += (uint64_t)n * 16;
}

// HBM a sub-batch of `rows` needs while it is partitioned: the staged key and
// payload, their partitioned copy, partition ids and the block histograms.
static uint64_t sink_bytes(int64_t rows) { return (uint64_t)rows * 40 + (4u << 20); }

static bool fits(kg_grace *g, uint64_t need) {
  return !g->ctx->budget || g->ctx->in_use.load() + need + g->table_budget <= g->ctx->budget;
}

// Past the HBM budget (with one partition's tables kept free), partitioned
// pieces move to pinned host memory, oldest first; a batch is partitioned in
// sub-batches that fit beside what stays in HBM.
static void sink_rows(kg_grace *g, std::vector<kg_grace::Piece> &store, const kg_column *key,
                      const kg_column *pay, int64_t n, int side) {
  if (n <= 0) return;
  if (!is_int_like(key->type) || type_width(key->type) != 8 || !is_int_like(pay->type) ||
      type_width(pay->type) != 8)
    fail(KG_ERR_UNSUPPORTED, "grace join of 64-bit integer keys and payloads");
  check_column(*key, n, "grace key");
  check_column(*pay, n, "grace payload");
  for (int64_t done = 0; done < n;) {
    int64_t m = n - done;
    if (g->ctx->budget) {
      for (auto *v : {&g->build, &g->probe})
        for (auto &pc : *v)
          if (!fits(g, sink_bytes(m < (1 << 16) ? m : (1 << 16)))) spill_piece(g, pc);
      uint64_t used = g->ctx->in_use.load() + g->table_budget + (4u << 20);
      int64_t room = used < g->ctx->budget ? (int64_t)((g->ctx->budget - used) / 40) : 0;
      if (room < 1024 && room < m)
        fail(KG_ERR_MEMORY, "HBM budget too small to partition a grace batch");
      if (m > room) m = room;
    }
    kg_column k = *key, v = *pay;
    k.offset += done;
    v.offset += done;
    k.length -= done;
    v.length -= done;
    kg_grace::Piece piece;
    piece.off.assign(g->parts + 1, 0);
    {
      Staged st(g->ctx, g->s);
      const DCol dk = st.add(k, m), dv = st.add(v, m);
      // raw 64-bit key words hashed as such, raw 64-bit integer payload words
      const bool w64 = (dk.type == KG_T_INT64 || dk.type == KG_T_UINT64 ||
                        dk.type == KG_T_TIMESTAMP) &&
                       (dv.type == KG_T_INT64 || dv.type == KG_T_UINT64 ||
                        dv.type == KG_T_TIMESTAMP || dv.type == KG_T_DECIMAL64);
      const uint32_t parts = g->parts;
      const int64_t cus = g->ctx->num_cus > 0 ? g->ctx->num_cus : 256;
      const int R = parts <= 1024 ? 12 : 8;
      const int64_t tt = (int64_t)kGSBlock * R;
      int64_t per_block = ceil_div(ceil_div(m, cus * 4), tt) * tt;
      const int64_t nb = ceil_div(m, per_block);
      const uint64_t hn = (uint64_t)parts * nb;
      DevBuf hist(g->ctx, hn * 4 + 4), base(g->ctx, hn * 8 + 8), offs(g->ctx, (parts + 1) * 8);
      {
        ProfScope prof("hash_partition", g->s);
        const void *hf = w64 ? (const void *)k_gp_hist<true> : (const void *)k_gp_hist<false>;
        KG_HIP(set_dynamic_lds(hf, (int)(parts * 16)));
        if (w64)
          k_gp_hist<true><<<(unsigned)nb, kGPBlock, parts * 16, g->s>>>(
              dk, dv, m, parts, per_block, hist.get<uint32_t>(), g->mm.get<long long>() + 2 * side);
        else
          k_gp_hist<false><<<(unsigned)nb, kGPBlock, parts * 16, g->s>>>(
              dk, dv, m, parts, per_block, hist.get<uint32_t>(), g->mm.get<long long>() + 2 * side);
        KG_HIP(hipGetLastError());
      }
      DevBuf temp(g->ctx, (uint64_t)scan_partials((int64_t)hn) * 8);
      exclusive_scan(hist.get<uint32_t>(), base.get<uint64_t>(), (int64_t)hn, temp.get<uint64_t>(),
                     g->s);
      piece.key.reset(g->ctx, (uint64_t)m * 8 + 8);
      piece.pay.reset(g->ctx, (uint64_t)m * 8 + 8);
      {
        ProfScope prof("scatter", g->s);
        const size_t lds = gs_lds_bytes(R, parts);
      // the mod 2^31 reading depends on the payloads' low 31 bits only (2^31
      // divides 2^64): kept non-negative, so -1 still marks a NULL payload
      const uint64_t pmask = g->mod31 ? 0x7FFFFFFFull : ~0ull;
#define KG_GPS(RR, WW)                                                                        \
  do {                                                                                       \
    KG_HIP(set_dynamic_lds((const void *)k_gp_scatter<RR, WW>, lds));                        \
    k_gp_scatter<RR, WW><<<(unsigned)nb, kGSBlock, lds, g->s>>>(                             \
        dk, dv, m, parts, per_block, base.get<uint64_t>(), piece.key.get<int64_t>(),         \
        piece.pay.get<int64_t>(), pmask);                                                    \
  } while (0)
        if (R == 12 && w64) KG_GPS(12, true);
        else if (R == 12) KG_GPS(12, false);
        else if (w64) KG_GPS(8, true);
        else KG_GPS(8, false);
#undef KG_GPS
        KG_HIP(hipGetLastError());
      }
      k_gp_offsets<<<(parts + 256) / 256, 256, 0, g->s>>>(base.get<uint64_t>(), hist.get<uint32_t>(),
                                                          hn, parts, (uint32_t)nb,
                                                          offs.get<int64_t>());
      KG_HIP(hipGetLastError());
      KG_HIP(hipMemcpyAsync(piece.off.data(), offs.get(), (parts + 1) * 8, hipMemcpyDeviceToHost,
                            g->s));
      KG_HIP(hipStreamSynchronize(g->s));
    }
    piece.rows = piece.off[g->parts];
    piece.on_device = true;
    if (!fits(g, 0)) spill_piece(g, piece);
    store.push_back(std::move(piece));
    done += m;
  }
}

// Partition p's rows of a piece as key and payload columns (device or host).
static void slice(const kg_grace::Piece &pc, uint32_t p, kg_column &k, kg_column &v, int64_t &n) {
  const int64_t a = pc.off[p], b = pc.off[p + 1];
  n = b - a;
  const uint32_t fl = pc.on_device ? KG_COL_DEVICE : 0u;
  const void *kv = pc.on_device ? (const void *)pc.key.get() : (const void *)pc.hk;
  const void *pv = pc.on_device ? (const void *)pc.pay.get() : (const void *)pc.hp;
  k = kg_column{KG_T_INT64, 0, 0, fl, b, a, kv, nullptr};
  v = kg_column{KG_T_INT64, 0, 0, fl, b, a, pv, nullptr};
}

}  // namespace kg

extern "C" {

int kg_grace_create(kg_ctx *ctx, uint32_t partitions, uint64_t table_budget_bytes,
                    uint32_t flags, kg_grace **out) {
  return guarded([&] {
    *out = nullptr;
    if (partitions < 1 || partitions > 4096) fail(KG_ERR_UNSUPPORTED, "1 to 4096 partitions");
    if (flags & ~KG_GRACE_SUM_MOD31) fail(KG_ERR_INTERNAL, "unknown grace flags");
    auto g = std::make_unique<kg_grace>();
    g->ctx = ctx;
    g->parts = partitions;
    g->mod31 = (flags & KG_GRACE_SUM_MOD31) != 0;
    g->table_budget = table_budget_bytes ? table_budget_bytes : (8ull << 30);
    g->s = make_stream(ctx->device);
    StreamScope kg_scope(g->s);
    g->mm.reset(ctx, 32);
    long long init[4] = {INT64_MAX, INT64_MIN, INT64_MAX, INT64_MIN};
    KG_HIP(hipMemcpy(g->mm.get(), init, 32, hipMemcpyHostToDevice));
    *out = g.release();
  });
}

int kg_grace_build_sink(kg_grace *g, const kg_column *key, const kg_column *pay, int64_t nrows) {
  return guarded([&] { StreamScope kg_scope(g->s); sink_rows(g, g->build, key, pay, nrows, 0); });
}

int kg_grace_probe_sink(kg_grace *g, const kg_column *key, const kg_column *pay, int64_t nrows) {
  return guarded([&] { StreamScope kg_scope(g->s); sink_rows(g, g->probe, key, pay, nrows, 1); });
}

int kg_grace_finish(kg_grace *g, int64_t *counts, int64_t *sums, uint8_t *sum_valid,
                    uint64_t *spilled_bytes) {
  return guarded([&] {
    StreamScope kg_scope(g->s);
    long long mm[4];
    KG_HIP(hipMemcpyAsync(mm, g->mm.get(), 32, hipMemcpyDeviceToHost, g->s));
    KG_HIP(hipStreamSynchronize(g->s));
    bool any_b = mm[0] <= mm[1], any_p = mm[2] <= mm[3];
    if (any_b && any_p) {
      // every pair's product in [0, 2^31): the per-pair % 2^31 is the identity
      // (the mod 2^31 reading takes any payloads)
      bool nonneg = g->mod31 || (mm[0] >= 0 && mm[2] >= 0);
      unsigned __int128 prod = (unsigned __int128)(uint64_t)(mm[1] > 0 ? mm[1] : 0) *
                               (unsigned __int128)(uint64_t)(mm[3] > 0 ? mm[3] : 0);
      if (!g->mod31 && (!nonneg || prod >= ((unsigned __int128)1 << 31)))
        fail(KG_ERR_UNSUPPORTED,
             "payload products reach 2^31 or are negative: the per-pair modulus does not factor");
    }
    DevBuf out(g->ctx, (uint64_t)g->parts * 3 * 8);
    KG_HIP(hipMemsetAsync(out.get(), 0, out.bytes(), g->s));
    // one table buffer, sized for the partition with the most build rows
    // (an upper bound of its distinct keys), reused by every partition
    std::vector<int64_t> nbp(g->parts, 0), npp(g->parts, 0);
    int64_t maxb = 0;
    for (uint32_t p = 0; p < g->parts; p++) {
      for (auto &pc : g->build) nbp[p] += pc.off[p + 1] - pc.off[p];
      for (auto &pc : g->probe) npp[p] += pc.off[p + 1] - pc.off[p];
      if (nbp[p] >= (1ll << 32)) fail(KG_ERR_UNSUPPORTED, "a grace partition of 2^32 build rows");
      if (npp[p] > 0 && nbp[p] > maxb) maxb = nbp[p];
    }
    const uint64_t maxcap = 1ull << glog2((uint64_t)(maxb < 256 ? 256 : maxb) * 2);
    if ((maxcap + 1) * sizeof(GSlot) > g->table_budget)
      fail(KG_ERR_MEMORY, "a grace partition's table exceeds the table budget");
    DevBuf slots(g->ctx, (maxcap + 1) * sizeof(GSlot));
    // piece descriptors of every partition, in pinned host memory until
    // their uploads (enqueued in order on the stream) have run
    const size_t per_part = g->build.size() + g->probe.size();
    GPiece *hdesc = nullptr;
    KG_HIP(hipHostMalloc((void **)&hdesc, sizeof(GPiece) * per_part * g->parts + 64,
                         hipHostMallocDefault));
    std::unique_ptr<GPiece, void (*)(GPiece *)> hguard(hdesc, [](GPiece *q) {
      (void)hipHostFree(q);
    });
    DevBuf ddesc(g->ctx, sizeof(GPiece) * per_part * g->parts + 64);
    for (uint32_t p = 0; p < g->parts; p++) {
      if (nbp[p] == 0 || npp[p] == 0) continue;
      // host (spilled) pieces of this partition are staged into HBM; the
      // staging buffers are released in the stream's order after the launches
      Staged st(g->ctx, g->s);
      GPiece *hb = hdesc + per_part * p;
      int nb = 0, np = 0;
      for (auto &pc : g->build) {
        kg_column k, v;
        int64_t n;
        slice(pc, p, k, v, n);
        if (n) hb[nb++] = GPiece{st.add(k, n), st.add(v, n), n};
      }
      GPiece *hp = hb + nb;
      for (auto &pc : g->probe) {
        kg_column k, v;
        int64_t n;
        slice(pc, p, k, v, n);
        if (n) hp[np++] = GPiece{st.add(k, n), st.add(v, n), n};
      }
      GPiece *db = ddesc.get<GPiece>() + per_part * p;
      KG_HIP(hipMemcpyAsync(db, hb, sizeof(GPiece) * (nb + np), hipMemcpyHostToDevice, g->s));
      const uint64_t cap = 1ull << glog2((uint64_t)(nbp[p] < 256 ? 256 : nbp[p]) * 2);
      const int lc = glog2(cap);
      k_grace_init<<<stream_grid(cap + 1, 256), 256, 0, g->s>>>(slots.get<GSlot>(), cap + 1);
      {
        ProfScope prof("grace_build", g->s);
        const size_t lds = (size_t)kGBSlots * 3 * 8;
        KG_HIP(set_dynamic_lds((const void *)k_grace_build, (int)lds));
        const int64_t tiles = ceil_div(nbp[p], kGBTile);
        const int64_t cus = g->ctx->num_cus > 0 ? g->ctx->num_cus : 256;
        k_grace_build<<<(unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, cus)), kGBBlock,
                        lds, g->s>>>(db, nb, slots.get<GSlot>(), cap - 1, lc, cap);
        KG_HIP(hipGetLastError());
      }
      {
        ProfScope prof("grace_probe", g->s);
        k_grace_probe<<<stream_grid(npp[p], 256, kGPU * 2), 256, 0, g->s>>>(
            db + nb, np, slots.get<GSlot>(), cap - 1, lc, cap,
            out.get<unsigned long long>() + 3 * p);
        KG_HIP(hipGetLastError());
      }
      // spilled pieces were staged into HBM: let them go before the next
      // partition's, so the staging stays within one partition's rows
      if (st.staged_host()) KG_HIP(hipStreamSynchronize(g->s));
    }
    KG_HIP(hipStreamSynchronize(g->s));
    std::vector<uint64_t> h((size_t)g->parts * 3);
    KG_HIP(hipMemcpy(h.data(), out.get(), h.size() * 8, hipMemcpyDeviceToHost));
    for (uint32_t p = 0; p < g->parts; p++) {
      counts[p] = (int64_t)h[3 * p];
      sums[p] = (int64_t)(g->mod31 ? (h[3 * p + 2] & 0x7FFFFFFFull) : h[3 * p + 2]);
      if (sum_valid) sum_valid[p] = h[3 * p + 1] != 0;
    }
    if (spilled_bytes) *spilled_bytes = g->spilled;
  });
}

void kg_grace_destroy(kg_grace *g) {
  if (!g) return;
  if (g->s) (void)hipStreamSynchronize(g->s);
  for (auto &p : g->build) free_piece(p);
  for (auto &p : g->probe) free_piece(p);
  hipStream_t s = g->s;
  delete g;
  destroy_stream(s);
}

}  // extern "C"
