// kg_bloom.hip -- the runtime Bloom filter: the Parquet split-block Bloom
// filter (RuntimeBloom, crates/krishiv-io/src/runtime_filter/bloom.rs:83-320)
// over XXH64 seed 0 of the normalized key bytes (runtime_filter/hash.rs:56-131),
// built as KeyFilterBuilder::finish_with/bloom() (runtime_filter/filter.rs:533-624)
// and probed as check_batch (bloom.rs:211-247) by the scan's row stage.
//
// Build: one thread per key sets one bit in each of the block's eight words
// with atomicOr (OR is order-free, so the bit array is the CPU's byte for
// byte).  Probe: pass 1 streams the key column, tests the 32-byte block
// (L2/MALL resident for filters of a few MB) and writes a 64-row ballot word
// per wave plus a per-tile count; pass 2 scans the counts and expands the
// bitmap into the surviving row ids in input order.
#include <hip/hip_runtime.h>


#include <cmath>

#include "kg_common.hpp"
#include "kg_device.hpp"
#include "kg_scan.hpp"

namespace kg {

constexpr int kProbeBlock = 256;
constexpr int kProbeTile = kProbeBlock * 32;  // rows per workgroup tile

__device__ __forceinline__ uint64_t sbbf_block(uint64_t h, uint64_t nblocks) {
  return ((h >> 32) * nblocks) >> 32;
}

__device__ __forceinline__ bool select_row(const DCol *sel, int64_t i) {
  if (!sel) return true;
  return dev::valid(*sel, i) && dev::load_i64(*sel, i) != 0;
}

__global__ void k_bloom_insert(DCol key, DCol sel, int has_sel, int64_t n, uint32_t *words,
                               uint64_t nblocks, unsigned long long *items) {
  uint64_t mine = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!dev::valid(key, i) || !select_row(has_sel ? &sel : nullptr, i)) continue;
    uint64_t h = dev::value_hash(key, i);
    uint32_t *b = words + sbbf_block(h, nblocks) * 8;
    uint32_t k = (uint32_t)h;
#pragma unroll
    for (int w = 0; w < 8; w++) atomicOr(b + w, 1u << ((k * dev::sbbf_salt(w)) >> 27));
    mine++;
  }
  // one atomic per wave for the item count
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(items, (unsigned long long)mine);
}

__device__ __forceinline__ bool sbbf_check(const uint32_t *words, uint64_t nblocks, uint64_t h) {
  const uint4 *b = reinterpret_cast<const uint4 *>(words + sbbf_block(h, nblocks) * 8);
  uint4 x = b[0], y = b[1];
  uint32_t k = (uint32_t)h;
  uint32_t miss = 0;
  miss |= (1u << ((k * dev::sbbf_salt(0)) >> 27)) & ~x.x;
  miss |= (1u << ((k * dev::sbbf_salt(1)) >> 27)) & ~x.y;
  miss |= (1u << ((k * dev::sbbf_salt(2)) >> 27)) & ~x.z;
  miss |= (1u << ((k * dev::sbbf_salt(3)) >> 27)) & ~x.w;
  miss |= (1u << ((k * dev::sbbf_salt(4)) >> 27)) & ~y.x;
  miss |= (1u << ((k * dev::sbbf_salt(5)) >> 27)) & ~y.y;
  miss |= (1u << ((k * dev::sbbf_salt(6)) >> 27)) & ~y.z;
  miss |= (1u << ((k * dev::sbbf_salt(7)) >> 27)) & ~y.w;
  return miss == 0;
}

// The first four words of the block of h: the bits they miss (0: the row
// still passes).
__device__ __forceinline__ uint32_t sbbf_miss_lo(uint4 x, uint32_t k) {
  uint32_t miss = 0;
  miss |= (1u << ((k * dev::sbbf_salt(0)) >> 27)) & ~x.x;
  miss |= (1u << ((k * dev::sbbf_salt(1)) >> 27)) & ~x.y;
  miss |= (1u << ((k * dev::sbbf_salt(2)) >> 27)) & ~x.z;
  miss |= (1u << ((k * dev::sbbf_salt(3)) >> 27)) & ~x.w;
  return miss;
}
__device__ __forceinline__ uint32_t sbbf_miss_hi(uint4 y, uint32_t k) {
  uint32_t miss = 0;
  miss |= (1u << ((k * dev::sbbf_salt(4)) >> 27)) & ~y.x;
  miss |= (1u << ((k * dev::sbbf_salt(5)) >> 27)) & ~y.y;
  miss |= (1u << ((k * dev::sbbf_salt(6)) >> 27)) & ~y.z;
  miss |= (1u << ((k * dev::sbbf_salt(7)) >> 27)) & ~y.w;
  return miss;
}

// Pass 1: tile t covers rows [t*kProbeTile, (t+1)*kProbeTile); bitmap word
// (row / 64) holds the answers of 64 consecutive rows.  A lane takes kPU
// rows at a time: their key loads, then their blocks' first halves (16
// bytes), are in flight together; the second half is read only by the rows
// the first half did not already reject (most rows of a selective filter).
// W64: keys that hash as one raw 64-bit word (int64, uint64, temporal).
constexpr int kPU = 4;

template <bool W64>
__global__ void __launch_bounds__(kProbeBlock) k_bloom_probe_bits(DCol key, int64_t n,
                                                                  const uint32_t *words,
                                                                  uint64_t nblocks,
                                                                  uint64_t *bitmap,
                                                                  uint32_t *tile_count) {
  __shared__ uint32_t s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kProbeTile;
  uint32_t cnt = 0;
  const int lane = threadIdx.x & 63;
  for (int k0 = 0; k0 < kProbeTile / kProbeBlock; k0 += kPU) {
    uint64_t h[kPU];
    bool live[kPU];
    if (W64) {
      uint64_t kv[kPU];
      uint32_t kb[kPU];
#pragma unroll
      for (int u = 0; u < kPU; u++) {
        const int64_t i = base + (int64_t)(k0 + u) * kProbeBlock + threadIdx.x;
        const int64_t ic = i < n ? i : n - 1;
        kv[u] = dev::load_w64(key, ic);
        kb[u] = key.validity ? key.validity[(ic + key.offset) >> 3] : 0xFFu;
      }
#pragma unroll
      for (int u = 0; u < kPU; u++) {
        const int64_t i = base + (int64_t)(k0 + u) * kProbeBlock + threadIdx.x;
        const int64_t ic = i < n ? i : n - 1;
        live[u] = i < n && ((kb[u] >> ((ic + key.offset) & 7)) & 1);
        h[u] = dev::xxh64_fixed(kv[u], 0, 8);
      }
    } else {
#pragma unroll
      for (int u = 0; u < kPU; u++) {
        const int64_t i = base + (int64_t)(k0 + u) * kProbeBlock + threadIdx.x;
        live[u] = i < n && dev::valid(key, i);
        h[u] = live[u] ? dev::value_hash(key, i) : 0;
      }
    }
    // every lane reads its first half (a block in range for any h), so the
    // reads are not behind a per-lane branch; rows not live are masked after
    uint4 x[kPU];
#pragma unroll
    for (int u = 0; u < kPU; u++)
      x[u] = reinterpret_cast<const uint4 *>(words + sbbf_block(h[u], nblocks) * 8)[0];
    uint32_t miss[kPU];
#pragma unroll
    for (int u = 0; u < kPU; u++) miss[u] = live[u] ? sbbf_miss_lo(x[u], (uint32_t)h[u]) : 1u;
#pragma unroll
    for (int u = 0; u < kPU; u++) {
      if (!miss[u]) {
        const uint4 *b = reinterpret_cast<const uint4 *>(words + sbbf_block(h[u], nblocks) * 8);
        miss[u] = sbbf_miss_hi(b[1], (uint32_t)h[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kPU; u++) {
      const uint64_t m = __ballot(miss[u] == 0);
      const int64_t w0 = base + (int64_t)(k0 + u) * kProbeBlock + (threadIdx.x & ~63);
      if (lane == 0 && w0 < n) {
        bitmap[w0 >> 6] = m;
        cnt += (uint32_t)__popcll(m);
      }
    }
  }
  if (lane == 0 && cnt) atomicAdd(&s_cnt, cnt);
  __syncthreads();
  if (threadIdx.x == 0) tile_count[blockIdx.x] = s_cnt;
}

// Pass 2: thread t of a tile owns 32 consecutive rows (half a bitmap word);
// the tile's survivors go to out_rows from tile_off[tile] on, as row_base +
// row (the row ids of a chunk that starts at row_base).
__global__ void __launch_bounds__(kProbeBlock) k_bloom_probe_rows(const uint64_t *bitmap, int64_t n,
                                                                  const uint64_t *tile_off,
                                                                  int64_t row_base,
                                                                  int32_t *out_rows) {
  __shared__ uint32_t s_w[kProbeBlock / 64];
  int64_t base = (int64_t)blockIdx.x * kProbeTile;
  int64_t r0 = base + (int64_t)threadIdx.x * 32;
  uint32_t bits = 0;
  if (r0 < n) {
    uint64_t w = bitmap[r0 >> 6];
    bits = (uint32_t)(((r0 >> 5) & 1) ? (w >> 32) : w);
  }
  uint32_t tot;
  const uint32_t off = block_exclusive_scan<kProbeBlock>((uint32_t)__popc(bits), s_w, &tot);
  int64_t pos = (int64_t)tile_off[blockIdx.x] + off;
  while (bits) {
    int b = __ffs(bits) - 1;
    bits &= bits - 1;
    out_rows[pos++] = (int32_t)(row_base + r0 + b);
  }
}

// ---------------------------------------------------------------- Spark layout
// Spark's sketch BloomFilterImpl (version 1): a long[] bit array of
// `nwords` words; putLong(item) sets bit (h1 + i*h2, ~ when negative,
// 32-bit) mod (64 nwords) for i = 1..k, h1 = murmur3(item, 0), h2 =
// murmur3(item, h1).  A runtime filter's item is Spark's xxhash64(key), seed
// 42 (of the int value for keys of 4 bytes or less, of the long otherwise).
__device__ __forceinline__ int64_t spark_item(const DCol &c, int64_t i) {
  const int64_t v = dev::load_i64(c, i);
  return c.width <= 4 ? (int64_t)dev::xxh64_seeded((uint32_t)(int32_t)v, 4, 42)
                      : (int64_t)dev::xxh64_seeded((uint64_t)v, 8, 42);
}

__device__ __forceinline__ uint64_t spark_bit(int32_t h1, int32_t h2, int i, uint64_t bits) {
  int32_t c = (int32_t)((uint32_t)h1 + (uint32_t)i * (uint32_t)h2);
  if (c < 0) c = ~c;
  return (uint64_t)c % bits;
}

__global__ void k_spark_insert(DCol key, DCol sel, int has_sel, int64_t n,
                               unsigned long long *words, uint64_t nwords, int k,
                               unsigned long long *items) {
  uint64_t mine = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!dev::valid(key, i) || !select_row(has_sel ? &sel : nullptr, i)) continue;
    const int64_t item = spark_item(key, i);
    const int32_t h1 = dev::murmur3_long((uint64_t)item, 0);
    const int32_t h2 = dev::murmur3_long((uint64_t)item, h1);
    for (int j = 1; j <= k; j++) {
      const uint64_t b = spark_bit(h1, h2, j, nwords * 64);
      atomicOr(words + (b >> 6), 1ull << (b & 63));
    }
    mine++;
  }
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(items, (unsigned long long)mine);
}

__device__ __forceinline__ bool spark_check(const unsigned long long *words, uint64_t nwords, int k,
                                            int64_t item) {
  const int32_t h1 = dev::murmur3_long((uint64_t)item, 0);
  const int32_t h2 = dev::murmur3_long((uint64_t)item, h1);
  bool all = true;
  for (int j = 1; j <= k; j++) {
    const uint64_t b = spark_bit(h1, h2, j, nwords * 64);
    all = all && ((words[b >> 6] >> (b & 63)) & 1);
  }
  return all;
}

// Pass 1 of the probe for the Spark layout (the same bitmap and tile counts
// as k_bloom_probe_bits).
__global__ void __launch_bounds__(kProbeBlock) k_spark_probe_bits(DCol key, int64_t n,
                                                                  const unsigned long long *words,
                                                                  uint64_t nwords, int k,
                                                                  uint64_t *bitmap,
                                                                  uint32_t *tile_count) {
  __shared__ uint32_t s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kProbeTile;
  uint32_t cnt = 0;
  const int lane = threadIdx.x & 63;
  for (int t = 0; t < kProbeTile / kProbeBlock; t++) {
    int64_t i = base + (int64_t)t * kProbeBlock + threadIdx.x;
    bool pass = false;
    if (i < n && dev::valid(key, i)) pass = spark_check(words, nwords, k, spark_item(key, i));
    uint64_t m = __ballot(pass);
    int64_t w0 = base + (int64_t)t * kProbeBlock + (threadIdx.x & ~63);
    if (lane == 0 && w0 < n) {
      bitmap[w0 >> 6] = m;
      cnt += (uint32_t)__popcll(m);
    }
  }
  if (lane == 0 && cnt) atomicAdd(&s_cnt, cnt);
  __syncthreads();
  if (threadIdx.x == 0) tile_count[blockIdx.x] = s_cnt;
}

}  // namespace kg

using namespace kg;

struct kg_bloom {
  kg_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  bool spark = false;     // Spark's BloomFilterImpl layout instead of the SBBF
  uint64_t nblocks = 0;   // SBBF: 32-byte blocks
  uint64_t nwords = 0;    // Spark: 64-bit words
  int32_t k = 0;          // Spark: hash functions
  DevBuf words, items;
  DevBuf bitmap, tiles, tile_off, temp;  // probe scratch
};

// Pass 1 of a probe of rows [0, n) of k: the survivors' bitmap and a count
// per tile, for either layout.
static void probe_bits(kg_bloom *b, const DCol &k, int64_t n, uint64_t *bitmap, uint32_t *tiles) {
  const unsigned grid = (unsigned)ceil_div(n, kProbeTile);
  if (b->spark) {
    k_spark_probe_bits<<<grid, kProbeBlock, 0, b->s>>>(k, n, b->words.get<unsigned long long>(),
                                                      b->nwords, b->k, bitmap, tiles);
  } else if (k.type == KG_T_INT64 || k.type == KG_T_UINT64 || k.type == KG_T_TIMESTAMP) {
    k_bloom_probe_bits<true><<<grid, kProbeBlock, 0, b->s>>>(k, n, b->words.get<uint32_t>(),
                                                            b->nblocks, bitmap, tiles);
  } else {
    k_bloom_probe_bits<false><<<grid, kProbeBlock, 0, b->s>>>(k, n, b->words.get<uint32_t>(),
                                                             b->nblocks, bitmap, tiles);
  }
  KG_HIP(hipGetLastError());
}

__global__ void k_fill_u64(uint64_t *p, int64_t n, uint64_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- scratch of kg_bloom_probe_noalloc --------------------------------------
// The allocation-free probe works through its input in chunks of kNaChunk
// rows; everything beyond the caller's input and output lives in `scratch`,
// one 2.03 MB array per GPU for the life of the process: the chunk's
// survivor bitmap, its tile counts and offsets, the scan's segment sums and
// a two-word running survivor count.  Calls that overlap on one GPU share
// the array and race: the caller serializes them.
constexpr int64_t kNaChunk = 1 << 24;                        // rows a chunk
constexpr int64_t kNaTilesMax = kNaChunk / kProbeTile;       // 2048 tiles a chunk
constexpr int64_t kNaBitmap = 0;                             // kNaChunk / 64 words
constexpr int64_t kNaTiles = kNaBitmap + kNaChunk / 64;      // kNaTilesMax uint32 (half as many words)
constexpr int64_t kNaTileOff = kNaTiles + kNaTilesMax / 2;   // kNaTilesMax words
constexpr int64_t kNaPartials = kNaTileOff + kNaTilesMax;    // scan segment sums
constexpr int64_t kNaTotal = kNaPartials + (kNaTilesMax + kScanSegment - 1) / kScanSegment;
constexpr int64_t kNaWords = kNaTotal + 2;
static_assert(kNaChunk % kProbeTile == 0, "chunks are whole tiles");

static __device__ __align__(256) uint64_t scratch[kNaWords];

// The device address of `scratch` on a GPU, looked up once per device with
// that device current (each GPU holds its own copy of the array).
static uint64_t *noalloc_scratch(int device) {
  static std::mutex mu;
  static std::map<int, uint64_t *> ptrs;
  std::lock_guard<std::mutex> g(mu);
  auto it = ptrs.find(device);
  if (it != ptrs.end()) return it->second;
  int prev = 0;
  KG_HIP(hipGetDevice(&prev));
  KG_HIP(hipSetDevice(device));
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  KG_HIP(hipSetDevice(prev));
  if (e != hipSuccess || !p)
    fail(KG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  ptrs[device] = static_cast<uint64_t *>(p);
  return static_cast<uint64_t *>(p);
}

static void check_spark_key(const kg_column &c) {
  const int t = c.type;
  const int w = type_width(t);
  if (!(t == KG_T_INT8 || t == KG_T_INT16 || t == KG_T_INT32 || t == KG_T_DATE32 ||
        t == KG_T_INT64 || t == KG_T_TIMESTAMP || t == KG_T_DECIMAL64) || w <= 0)
    fail(KG_ERR_UNSUPPORTED, "Spark-layout Bloom keys of this type");
}

extern "C" {

uint64_t kg_bloom_bytes_for(uint64_t items, double fpp) {
  // bytes_for (bloom.rs:96-110): Parquet's -8n / ln(1 - fpp^(1/8)) bits in
  // whole 32-byte blocks, at least one
  if (fpp < 1e-9) fpp = 1e-9;
  if (fpp > 0.5) fpp = 0.5;
  double bits = -8.0 * (double)items / std::log(1.0 - std::pow(fpp, 1.0 / 8.0));
  uint64_t bytes = (uint64_t)std::ceil(bits / 8.0);
  uint64_t blocks = (bytes + 31) / 32;
  if (blocks < 1) blocks = 1;
  return blocks * 32;
}

uint64_t kg_spark_bloom_bits_for(uint64_t items, double fpp) {
  // optimalNumOfBits: (long)(-n ln p / (ln 2)^2)
  if (fpp <= 0.0 || fpp >= 1.0) return 0;
  const double l2 = std::log(2.0);
  return (uint64_t)(int64_t)(-(double)items * std::log(fpp) / (l2 * l2));
}

int32_t kg_spark_bloom_hashes_for(uint64_t items, uint64_t num_bits) {
  // optimalNumOfHashFunctions: max(1, round(m / n * ln 2))
  const double k = (double)num_bits / (double)(items ? items : 1) * std::log(2.0);
  const int32_t r = (int32_t)std::llround(k);
  return r < 1 ? 1 : r;
}

int kg_spark_bloom_create(kg_ctx *ctx, uint64_t num_bits, int32_t num_hashes, kg_bloom **out) {
  return guarded([&] {
    *out = nullptr;
    if (num_hashes < 1 || num_hashes > 64) fail(KG_ERR_UNSUPPORTED, "1 to 64 hash functions");
    auto b = std::make_unique<kg_bloom>();
    b->ctx = ctx;
    b->spark = true;
    b->k = num_hashes;
    b->nwords = (num_bits + 63) / 64;
    if (b->nwords < 1) b->nwords = 1;
    // Spark indexes bits with a non-negative int: more than 2^31 bits are unreachable
    if (b->nwords > (1ull << 26)) fail(KG_ERR_UNSUPPORTED, "Spark-layout Bloom of 2^32 bits or more");
    b->s = make_stream(ctx->device);
    StreamScope kg_scope(b->s);
    b->words.reset(ctx, b->nwords * 8);
    b->items.reset(ctx, 8);
    KG_HIP(hipMemsetAsync(b->words.get(), 0, b->nwords * 8, b->s));
    KG_HIP(hipMemsetAsync(b->items.get(), 0, 8, b->s));
    KG_HIP(hipStreamSynchronize(b->s));
    *out = b.release();
  });
}

int kg_bloom_create(kg_ctx *ctx, uint64_t nbytes, kg_bloom **out) {
  return guarded([&] {
    *out = nullptr;
    auto b = std::make_unique<kg_bloom>();
    b->ctx = ctx;
    b->nblocks = (nbytes + 31) / 32;
    if (b->nblocks < 1) b->nblocks = 1;
    if (b->nblocks > (1ull << 32)) fail(KG_ERR_UNSUPPORTED, "Bloom filter of 2^32 blocks or more");
    b->s = make_stream(ctx->device);
    StreamScope kg_scope(b->s);
    b->words.reset(ctx, b->nblocks * 32);
    b->items.reset(ctx, 8);
    KG_HIP(hipMemsetAsync(b->words.get(), 0, b->nblocks * 32, b->s));
    KG_HIP(hipMemsetAsync(b->items.get(), 0, 8, b->s));
    KG_HIP(hipStreamSynchronize(b->s));
    *out = b.release();
  });
}

int kg_bloom_insert(kg_bloom *b, const kg_column *key, const kg_column *select, int64_t nrows) {
  return guarded([&] {
    StreamScope kg_scope(b->s);
    if (nrows <= 0) return;
    check_column(*key, nrows, "Bloom key");
    Staged st(b->ctx, b->s);
    DCol k = st.add(*key, nrows);
    DCol sel{};
    if (select) {
      check_column(*select, nrows, "Bloom selection");
      sel = st.add(*select, nrows);
    }
    if (b->spark) {
      check_spark_key(*key);
      ProfScope prof("bloom_insert", b->s);
      k_spark_insert<<<stream_grid(nrows, 256), 256, 0, b->s>>>(
          k, sel, select != nullptr, nrows, b->words.get<unsigned long long>(), b->nwords, b->k,
          b->items.get<unsigned long long>());
      KG_HIP(hipGetLastError());
    } else {
      ProfScope prof("bloom_insert", b->s);
      k_bloom_insert<<<stream_grid(nrows, 256), 256, 0, b->s>>>(
          k, sel, select != nullptr, nrows, b->words.get<uint32_t>(), b->nblocks,
          b->items.get<unsigned long long>());
      KG_HIP(hipGetLastError());
    }
    if (st.staged_host()) KG_HIP(hipStreamSynchronize(b->s));
  });
}

double kg_bloom_expected_fpp(const kg_bloom *b, uint64_t items) {
  if (b->spark)  // Spark's expectedFpp: (1 - e^(-k n / m))^k
    return std::pow(1.0 - std::exp(-(double)b->k * (double)items / (64.0 * (double)b->nwords)),
                    (double)b->k);
  return kg::sbbf_expected_fpp(b->nblocks, items);
}

}  // extern "C"

namespace kg {
double sbbf_expected_fpp(uint64_t nblocks, uint64_t items) {
  // expected_fpp (bloom.rs:270-308): blocks receive Poisson(lambda) hashes
  double lambda = (double)items / (double)nblocks;
  if (lambda == 0.0) return 0.0;
  if (lambda > 1024.0) return std::pow(1.0 - std::pow(31.0 / 32.0, lambda), 8);
  double spread = 10.0 * std::sqrt(lambda) + 20.0;
  double lo = std::floor(std::fmax(lambda - spread, 0.0)), hi = lambda + spread;
  double ln_p = -lambda, k = 1.0;
  while (k <= lo) {
    ln_p += std::log(lambda) - std::log(k);
    k += 1.0;
  }
  double total = 0.0;
  k = lo;
  while (k <= hi) {
    double word = 1.0 - std::pow(31.0 / 32.0, k);
    double w2 = word * word, w4 = w2 * w2;
    total += std::exp(ln_p) * (w4 * w4);
    k += 1.0;
    ln_p += std::log(lambda) - std::log(k);
  }
  return total < 1.0 ? total : 1.0;
}
}  // namespace kg

extern "C" {

int kg_bloom_export(kg_bloom *b, uint8_t *bytes, uint64_t *len) {
  return guarded([&] {
    StreamScope kg_scope(b->s);
    uint64_t need = b->spark ? b->nwords * 8 : b->nblocks * 32;
    if (!bytes) {
      *len = need;
      return;
    }
    if (*len < need) fail(KG_ERR_INTERNAL, "export buffer too small");
    KG_HIP(hipMemcpyAsync(bytes, b->words.get(), need, hipMemcpyDeviceToHost, b->s));
    KG_HIP(hipStreamSynchronize(b->s));
    *len = need;
  });
}

int kg_bloom_probe(kg_bloom *b, const kg_column *key, int64_t nrows, int32_t *out_rows,
                   int64_t *out_count) {
  return guarded([&] {
    StreamScope kg_scope(b->s);
    *out_count = 0;
    if (nrows <= 0) return;
    if (nrows > INT32_MAX) fail(KG_ERR_UNSUPPORTED, "a probe batch of 2^31 rows or more");
    check_column(*key, nrows, "probe key");
    if (b->spark) check_spark_key(*key);
    Staged st(b->ctx, b->s);
    DCol k = st.add(*key, nrows);
    int64_t tiles = ceil_div(nrows, kProbeTile);
    b->bitmap.ensure(b->ctx, (uint64_t)ceil_div(nrows, 64) * 8 + 8);
    b->tiles.ensure(b->ctx, (uint64_t)tiles * 4);
    b->tile_off.ensure(b->ctx, (uint64_t)(tiles + 1) * 8);
    b->temp.ensure(b->ctx, (uint64_t)(scan_partials(tiles) + 2) * 8);
    uint64_t *total = b->temp.get<uint64_t>() + scan_partials(tiles);
    {
      ProfScope prof("bloom_probe", b->s);
      probe_bits(b, k, nrows, b->bitmap.get<uint64_t>(), b->tiles.get<uint32_t>());
    }
    exclusive_scan(b->tiles.get<uint32_t>(), b->tile_off.get<uint64_t>(), tiles,
                   b->temp.get<uint64_t>(), b->s, nullptr, total);
    {
      ProfScope prof("bloom_compact", b->s);
      k_bloom_probe_rows<<<(unsigned)tiles, kProbeBlock, 0, b->s>>>(
          b->bitmap.get<uint64_t>(), nrows, b->tile_off.get<uint64_t>(), 0, out_rows);
      KG_HIP(hipGetLastError());
    }
    uint64_t count = 0;
    KG_HIP(hipMemcpyAsync(&count, total, 8, hipMemcpyDeviceToHost, b->s));
    KG_HIP(hipStreamSynchronize(b->s));
    *out_count = (int64_t)count;
  });
}

int kg_bloom_probe_noalloc(kg_bloom *b, const kg_column *key, int64_t nrows, int32_t *out_rows,
                           int64_t *out_count) {
  return guarded([&] {
    *out_count = 0;
    if (nrows < 0) fail(KG_ERR_INTERNAL, "negative row count");
    if (nrows > INT32_MAX)
      fail(KG_ERR_UNSUPPORTED, "kg_bloom_probe_noalloc takes at most 2^31 - 1 rows a call");
    if (nrows == 0) return;
    if (!(key->flags & KG_COL_DEVICE))
      fail(KG_ERR_UNSUPPORTED, "kg_bloom_probe_noalloc takes device-resident key columns only");
    if (!out_rows) fail(KG_ERR_INTERNAL, "no output row buffer");
    check_column(*key, nrows, "probe key");
    if (b->spark) check_spark_key(*key);
    uint64_t *sc = noalloc_scratch(b->ctx->device);
    uint64_t *bitmap = sc + kNaBitmap, *partials = sc + kNaPartials, *total = sc + kNaTotal;
    uint32_t *tiles = reinterpret_cast<uint32_t *>(sc + kNaTiles);
    uint64_t *tile_off = sc + kNaTileOff;
    DCol k{};
    k.type = key->type;
    k.precision = key->precision;
    k.scale = key->scale;
    k.width = type_width(key->type);
    k.offset = key->offset;
    k.values = key->values;
    k.validity = key->validity;
    // the running survivor count starts at zero (a fill kernel, not a memset)
    k_fill_u64<<<1, 64, 0, b->s>>>(total, 2, 0);
    KG_HIP(hipGetLastError());
    int64_t c = 0;
    for (int64_t at = 0; at < nrows; at += kNaChunk, c++) {
      const int64_t m = std::min<int64_t>(kNaChunk, nrows - at);
      const int64_t tcount = ceil_div(m, kProbeTile);
      DCol kc = k;
      kc.offset += at;
      probe_bits(b, kc, m, bitmap, tiles);
      // survivors of the earlier chunks carry into this chunk's offsets
      exclusive_scan(tiles, tile_off, tcount, partials, b->s, total + (c & 1), total + ((c + 1) & 1));
      k_bloom_probe_rows<<<(unsigned)tcount, kProbeBlock, 0, b->s>>>(bitmap, m, tile_off, at,
                                                                      out_rows);
      KG_HIP(hipGetLastError());
    }
    uint64_t count = 0;
    KG_HIP(hipMemcpyAsync(&count, total + (c & 1), 8, hipMemcpyDeviceToHost, b->s));
    KG_HIP(hipStreamSynchronize(b->s));
    *out_count = (int64_t)count;
  });
}

void kg_bloom_destroy(kg_bloom *b) {
  if (!b) return;
  if (b->s) (void)hipStreamSynchronize(b->s);
  hipStream_t s = b->s;
  delete b;
  destroy_stream(s);
}

}  // extern "C"
