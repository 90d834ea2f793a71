// kg_exchange.hip -- the routing hash H_kg and the hash repartition behind
// exchanges (exchange::route + ExchangeSink's split,
// crates/krishiv-exec/src/exchange/mod.rs:159-212,447-501).
//
// Repartition is a stable counting sort by partition id in three launches:
//   1. k_route_hist   H_kg -> partition id per row (4 B/row written) and a
//                     per-block histogram (LDS atomics);
//   2. an exclusive scan of the partition-major histogram (kg_scan.hpp);
//   3. k_scatter      each wave replays its contiguous quarter of the
//                     block's rows; the rank of a row among the rows of its
//                     partition in the same 64-row round comes from a wave
//                     multisplit (one ballot per partition-id bit), so rows
//                     keep their input order inside a partition.
#include <hip/hip_runtime.h>


#include "kg_common.hpp"
#include "kg_device.hpp"
#include "kg_scan.hpp"

namespace kg {

constexpr int kMaxKeys = 8;
constexpr int kMaxCols = 32;
constexpr int kPartBlock = 256;  // 4 waves
constexpr uint32_t kMaxPartitions = 4096;

struct KeyCols {
  DCol c[kMaxKeys];
  int n;
};

struct CopyCols {
  DCol in[kMaxCols];
  void *out_values[kMaxCols];
  uint32_t *out_valid[kMaxCols];  // 32-bit words, zeroed; NULL when no validity
  int n;
};

__device__ __forceinline__ uint64_t hash_row(const KeyCols &k, int64_t i) {
  if (k.n == 1) return dev::value_hash(k.c[0], i);
  uint64_t h = dev::FOLD_INIT;
  for (int c = 0; c < k.n; c++) h = dev::fold(h ^ dev::value_hash(k.c[c], i), dev::FOLD_MUL);
  return h;
}

__global__ void k_hash(KeyCols k, int64_t n, uint64_t *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = hash_row(k, i);
}

__global__ void k_pids(KeyCols k, int64_t n, uint32_t parts, uint32_t *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = dev::partition_of(hash_row(k, i), parts);
}

// Rows [lo, hi) of block b: contiguous ranges of `per_block` rows.
__global__ void __launch_bounds__(kPartBlock) k_route_hist(KeyCols k, int64_t n, uint32_t parts,
                                                           int64_t per_block, uint32_t *pids,
                                                           uint32_t *hist) {
  extern __shared__ uint32_t s_hist[];
  for (uint32_t p = threadIdx.x; p < parts; p += blockDim.x) s_hist[p] = 0;
  __syncthreads();
  int64_t lo = blockIdx.x * per_block;
  int64_t hi = lo + per_block < n ? lo + per_block : n;
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    uint32_t p = dev::partition_of(hash_row(k, i), parts);
    pids[i] = p;
    atomicAdd(&s_hist[p], 1u);
  }
  __syncthreads();
  // partition-major layout: the scan then gives each (partition, block) its
  // first output row.
  for (uint32_t p = threadIdx.x; p < parts; p += blockDim.x)
    hist[(uint64_t)p * gridDim.x + blockIdx.x] = s_hist[p];
}

// The same histogram over partition ids given one byte a row (a join's
// grace partitions, Hashing::partition).
__global__ void __launch_bounds__(kPartBlock) k_pid_hist(const uint8_t *given, int64_t n,
                                                         uint32_t parts, int64_t per_block,
                                                         uint32_t *pids, uint32_t *hist) {
  extern __shared__ uint32_t s_hist[];
  for (uint32_t p = threadIdx.x; p < parts; p += blockDim.x) s_hist[p] = 0;
  __syncthreads();
  int64_t lo = blockIdx.x * per_block;
  int64_t hi = lo + per_block < n ? lo + per_block : n;
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const uint32_t p = given[i] < parts ? given[i] : parts - 1;
    pids[i] = p;
    atomicAdd(&s_hist[p], 1u);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < parts; p += blockDim.x)
    hist[(uint64_t)p * gridDim.x + blockIdx.x] = s_hist[p];
}

__device__ __forceinline__ void copy_value(const DCol &in, int64_t i, void *out, int64_t pos) {
  int64_t j = i + in.offset;
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
    default: {  // bit-packed BOOL
      if ((static_cast<const uint8_t *>(in.values)[j >> 3] >> (j & 7)) & 1)
        atomicOr(static_cast<uint32_t *>(out) + (pos >> 5), 1u << (pos & 31));
    }
  }
}

// One wave per contiguous quarter of the block's rows; LDS holds, per wave,
// the running output position of every partition.
__global__ void __launch_bounds__(kPartBlock) k_scatter(const uint32_t *pids, int64_t n,
                                                        uint32_t parts, int bits, int64_t per_block,
                                                        const uint64_t *base, CopyCols cols,
                                                        int64_t *out_rows) {
  extern __shared__ uint32_t s_cnt[];  // [4][parts] counts, then offsets (low 32 bits)
  __shared__ uint64_t s_base_hi;       // not used: offsets are kept relative
  (void)s_base_hi;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t p = threadIdx.x; p < 4 * parts; p += blockDim.x) s_cnt[p] = 0;
  __syncthreads();
  int64_t lo = blockIdx.x * per_block;
  int64_t hi = lo + per_block < n ? lo + per_block : n;
  int64_t q = (hi - lo + 3) / 4;
  int64_t wlo = lo + wave * q, whi = wlo + q < hi ? wlo + q : hi;
  uint32_t *cnt = s_cnt + wave * parts;
  // pass A: counts per partition of this wave's rows
  for (int64_t r = wlo; r < whi; r += 64) {
    int64_t i = r + lane;
    bool live = i < whi;
    uint32_t p = live ? pids[i] : 0;
    uint64_t peers = __ballot(live);
    for (int b = 0; b < bits; b++) {
      uint64_t m = __ballot(live && ((p >> b) & 1));
      peers &= ((p >> b) & 1) ? m : ~m;
    }
    if (live && (peers & lt) == 0) cnt[p] += (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // exclusive prefix across the four waves, per partition (relative to the
  // block's first row of that partition)
  for (uint32_t p = threadIdx.x; p < parts; p += blockDim.x) {
    uint32_t a = 0;
    for (int w = 0; w < 4; w++) {
      uint32_t c = s_cnt[w * parts + p];
      s_cnt[w * parts + p] = a;
      a += c;
    }
  }
  __syncthreads();
  // pass B: replay in order and scatter
  for (int64_t r = wlo; r < whi; r += 64) {
    int64_t i = r + lane;
    bool live = i < whi;
    uint32_t p = live ? pids[i] : 0;
    uint64_t peers = __ballot(live);
    for (int b = 0; b < bits; b++) {
      uint64_t m = __ballot(live && ((p >> b) & 1));
      peers &= ((p >> b) & 1) ? m : ~m;
    }
    if (live) {
      int64_t pos = (int64_t)base[(uint64_t)p * gridDim.x + blockIdx.x] + cnt[p] +
                    __popcll(peers & lt);
      for (int c = 0; c < cols.n; c++) {
        copy_value(cols.in[c], i, cols.out_values[c], pos);
        if (cols.out_valid[c] && dev::valid(cols.in[c], i))
          atomicOr(cols.out_valid[c] + (pos >> 5), 1u << (pos & 31));
      }
      if (out_rows) out_rows[pos] = i;
    }
    // every lane has read cnt[p] above before the leaders advance it
    __builtin_amdgcn_wave_barrier();
    if (live && (peers & lt) == 0) cnt[p] += (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
}

static KeyCols stage_keys(Staged &st, int32_t nkeys, const kg_column *keys, int64_t nrows) {
  if (nkeys < 1 || nkeys > kMaxKeys) fail(KG_ERR_UNSUPPORTED, "1 to 8 key columns");
  KeyCols k{};
  k.n = nkeys;
  for (int c = 0; c < nkeys; c++) {
    check_column(keys[c], nrows, "key");
    k.c[c] = st.add(keys[c], nrows);
  }
  return k;
}

static int ceil_log2(uint32_t v) {
  int b = 0;
  while ((1u << b) < v) b++;
  return b;
}

}  // namespace kg

using namespace kg;

namespace kg {
// out[p] = in[p * stride] (each partition's first-block scan value)
__global__ void k_stride_gather(const uint64_t *in, uint64_t stride, uint32_t n, uint64_t *out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) out[p] = in[(uint64_t)p * stride];
}
}  // namespace kg

namespace kg {
// The row layout's stored hash h over the key columns (layout.rs:487-553):
// h = fold(0, MIX) = 0; per column a NULL mixes NULL_HASH, a value its LE
// words (narrower integers zero-extended, floats normalized, decimal128 as
// two words); h = fold(h, FINISH).
__device__ __forceinline__ uint64_t layout_hash(const KeyCols &k, int64_t i) {
  uint64_t h = 0;
  for (int c = 0; c < k.n; c++) {
    const DCol &d = k.c[c];
    auto mix = [&](uint64_t v) { h = dev::fold(h ^ v, dev::L_MIX); };
    if (!dev::valid(d, i)) {
      mix(dev::NULL_HASH);
      continue;
    }
    const int64_t j = i + d.offset;
    switch (d.type) {
      case KG_T_FLOAT64: mix(dev::norm_f64_bits(static_cast<const double *>(d.values)[j])); break;
      case KG_T_FLOAT32: mix(dev::norm_f32_bits(static_cast<const float *>(d.values)[j])); break;
      case KG_T_DECIMAL128: {
        const uint64_t *w = static_cast<const uint64_t *>(d.values) + 2 * j;
        mix(w[0]);
        mix(w[1]);
        break;
      }
      case KG_T_BOOL: mix((uint64_t)dev::load_i64(d, i)); break;
      default: {
        uint64_t v = 0;
        switch (d.width) {
          case 1: v = static_cast<const uint8_t *>(d.values)[j]; break;
          case 2: v = static_cast<const uint16_t *>(d.values)[j]; break;
          case 4: v = static_cast<const uint32_t *>(d.values)[j]; break;
          default: v = static_cast<const uint64_t *>(d.values)[j];
        }
        mix(v);
      }
    }
  }
  return dev::fold(h, dev::L_FINISH);
}

__global__ void k_join_partition(KeyCols k, int64_t n, int radix, int depth, uint64_t seed,
                                 uint64_t *hashes, uint8_t *parts) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h = layout_hash(k, i);
    if (hashes) hashes[i] = h;
    uint32_t p;
    if (depth == 0) {
      p = (uint32_t)(h >> (64 - radix));
    } else {
      const int d = depth < 6 ? depth : 6;
      p = (uint32_t)((dev::finalize(h, seed) >> (48 - 3 * d)) & 7);
    }
    parts[i] = (uint8_t)p;
  }
}
}  // namespace kg

extern "C" {

int kg_hash_columns(kg_ctx *ctx, int32_t ncols, const kg_column *cols, int64_t nrows,
                    uint64_t *out_hashes, uint32_t out_flags) {
  return guarded([&] {
    if (nrows <= 0) return;
    hipStream_t s = ctx->stream;
    Staged st(ctx, s);
    KeyCols k = stage_keys(st, ncols, cols, nrows);
    DevBuf tmp;
    uint64_t *out = out_hashes;
    if (!(out_flags & KG_COL_DEVICE)) {
      tmp.reset(ctx, (uint64_t)nrows * 8);
      out = tmp.get<uint64_t>();
    }
    {
      ProfScope prof("hash", s);
      k_hash<<<stream_grid(nrows, 256), 256, 0, s>>>(k, nrows, out);
      KG_HIP(hipGetLastError());
    }
    if (!(out_flags & KG_COL_DEVICE))
      KG_HIP(hipMemcpyAsync(out_hashes, out, (size_t)nrows * 8, hipMemcpyDeviceToHost, s));
    KG_HIP(hipStreamSynchronize(s));
  });
}

int kg_join_partition(kg_ctx *ctx, int32_t nkeys, const kg_column *keys, int64_t nrows,
                      uint32_t radix_bits, uint32_t depth, uint64_t seed, uint64_t *out_hashes,
                      uint8_t *out_parts, uint32_t out_flags) {
  return guarded([&] {
    if (nrows <= 0) return;
    if (nkeys < 1) fail(KG_ERR_UNSUPPORTED, "at least one key column");
    const int radix = radix_bits < 3 ? 3 : radix_bits > 6 ? 6 : (int)radix_bits;
    hipStream_t s = ctx->stream;
    Staged st(ctx, s);
    KeyCols k = stage_keys(st, nkeys, keys, nrows);
    const bool dev_out = (out_flags & KG_COL_DEVICE) != 0;
    DevBuf th, tp;
    uint64_t *hh = out_hashes;
    uint8_t *pp = out_parts;
    if (!dev_out) {
      if (out_hashes) {
        th.reset(ctx, (uint64_t)nrows * 8);
        hh = th.get<uint64_t>();
      }
      tp.reset(ctx, (uint64_t)nrows);
      pp = tp.get<uint8_t>();
    }
    {
      ProfScope prof("join_partition", s);
      k_join_partition<<<stream_grid(nrows, 256), 256, 0, s>>>(k, nrows, radix, (int)depth, seed,
                                                               hh, pp);
      KG_HIP(hipGetLastError());
    }
    if (!dev_out) {
      if (out_hashes)
        KG_HIP(hipMemcpyAsync(out_hashes, hh, (size_t)nrows * 8, hipMemcpyDeviceToHost, s));
      KG_HIP(hipMemcpyAsync(out_parts, pp, (size_t)nrows, hipMemcpyDeviceToHost, s));
    }
    KG_HIP(hipStreamSynchronize(s));
  });
}

int kg_partition(kg_ctx *ctx, int32_t nkeys, const kg_column *keys, int64_t nrows,
                 uint32_t partitions, uint32_t *out_pids, uint32_t out_flags) {
  return guarded([&] {
    if (nrows <= 0) return;
    hipStream_t s = ctx->stream;
    DevBuf tmp;
    uint32_t *out = out_pids;
    if (!(out_flags & KG_COL_DEVICE)) {
      tmp.reset(ctx, (uint64_t)nrows * 4);
      out = tmp.get<uint32_t>();
    }
    if (partitions <= 1 || nkeys == 0) {
      KG_HIP(hipMemsetAsync(out, 0, (size_t)nrows * 4, s));
    } else {
      Staged st(ctx, s);
      KeyCols k = stage_keys(st, nkeys, keys, nrows);
      ProfScope prof("hash_partition", s);
      k_pids<<<stream_grid(nrows, 256), 256, 0, s>>>(k, nrows, partitions, out);
      KG_HIP(hipGetLastError());
    }
    if (!(out_flags & KG_COL_DEVICE))
      KG_HIP(hipMemcpyAsync(out_pids, out, (size_t)nrows * 4, hipMemcpyDeviceToHost, s));
    KG_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"

namespace kg {
// The rows of `cols` regrouped by partition, stable within a partition:
// partitions routed from the key columns (given NULL), or given one byte a
// row (device).
static std::unique_ptr<kg_batch> split_rows(kg_ctx *ctx, int32_t ncols, const kg_column *cols,
                                            int64_t nrows, int32_t nkeys, const int32_t *key_cols,
                                            const uint8_t *given, uint32_t partitions,
                                            int64_t *out_offsets) {
  {
    if (ncols < 1 || ncols > kMaxCols) fail(KG_ERR_UNSUPPORTED, "1 to 32 columns");
    if (partitions < 1) partitions = 1;
    if (partitions > kMaxPartitions)
      fail(KG_ERR_UNSUPPORTED, "more than 4096 partitions on the device repartition");
    hipStream_t s = ctx->stream;
    Staged st(ctx, s);
    std::vector<DCol> in(ncols);
    for (int c = 0; c < ncols; c++) {
      check_column(cols[c], nrows, "repartition column");
      in[c] = st.add(cols[c], nrows);
    }
    KeyCols k{};
    k.n = nkeys;
    if (nkeys < 0 || nkeys > kMaxKeys) fail(KG_ERR_UNSUPPORTED, "0 to 8 key columns");
    for (int c = 0; c < nkeys; c++) {
      if (key_cols[c] < 0 || key_cols[c] >= ncols) fail(KG_ERR_INTERNAL, "key column index");
      k.c[c] = in[key_cols[c]];
    }
    auto b = std::make_unique<kg_batch>();
    b->ctx = ctx;
    b->nrows = nrows;
    b->cols.resize(ncols);
    CopyCols cc{};
    cc.n = ncols;
    for (int c = 0; c < ncols; c++) {
      auto &oc = b->cols[c];
      oc.type = cols[c].type;
      oc.precision = cols[c].precision;
      oc.scale = cols[c].scale;
      int w = type_width(oc.type);
      uint64_t vbytes = w == 0 ? (uint64_t)ceil_div(nrows, 32) * 4 : (uint64_t)nrows * w;
      oc.values.reset(ctx, vbytes > 0 ? vbytes : 4);
      if (w == 0) KG_HIP(hipMemsetAsync(oc.values.get(), 0, oc.values.bytes(), s));
      cc.in[c] = in[c];
      cc.out_values[c] = oc.values.get();
      if (in[c].validity) {
        oc.validity.reset(ctx, (uint64_t)ceil_div(nrows, 32) * 4 + 4);
        KG_HIP(hipMemsetAsync(oc.validity.get(), 0, oc.validity.bytes(), s));
        cc.out_valid[c] = oc.validity.get<uint32_t>();
      }
    }
    if (nrows == 0 || partitions == 1 || (nkeys == 0 && !given)) {
      // one partition: the rows in order (route() puts every row in 0)
      for (uint32_t p = 0; p <= partitions; p++) out_offsets[p] = p == 0 ? 0 : nrows;
      if (nrows > 0) {
        std::vector<uint32_t> dummy;
        DevBuf pids(ctx, (uint64_t)nrows * 4);
        KG_HIP(hipMemsetAsync(pids.get(), 0, (size_t)nrows * 4, s));
        DevBuf base(ctx, 8);
        KG_HIP(hipMemsetAsync(base.get(), 0, 8, s));
        ProfScope prof("scatter", s);
        k_scatter<<<1, kPartBlock, 4 * sizeof(uint32_t), s>>>(pids.get<uint32_t>(), nrows, 1, 0,
                                                              nrows, base.get<uint64_t>(), cc,
                                                              nullptr);
        KG_HIP(hipGetLastError());
        KG_HIP(hipStreamSynchronize(s));
      }
      return b;
    }
    // block count: the histogram is partitions x blocks entries
    int64_t nb = ceil_div(nrows, 4096);
    int64_t cap = (int64_t)(16u << 20) / partitions;
    if (nb > cap) nb = cap;
    if (nb > 65535) nb = 65535;
    if (nb < 1) nb = 1;
    int64_t per_block = ceil_div(nrows, nb);
    nb = ceil_div(nrows, per_block);
    DevBuf pids(ctx, (uint64_t)nrows * 4);
    uint64_t hn = (uint64_t)partitions * nb;
    DevBuf hist(ctx, hn * 4), base(ctx, hn * 8);
    {
      ProfScope prof("hash_partition", s);
      if (given)
        k_pid_hist<<<(unsigned)nb, kPartBlock, partitions * 4, s>>>(
            given, nrows, partitions, per_block, pids.get<uint32_t>(), hist.get<uint32_t>());
      else
        k_route_hist<<<(unsigned)nb, kPartBlock, partitions * 4, s>>>(
            k, nrows, partitions, per_block, pids.get<uint32_t>(), hist.get<uint32_t>());
      KG_HIP(hipGetLastError());
    }
    DevBuf temp(ctx, (uint64_t)scan_partials((int64_t)hn) * 8);
    exclusive_scan(hist.get<uint32_t>(), base.get<uint64_t>(), (int64_t)hn, temp.get<uint64_t>(), s);
    {
      ProfScope prof("scatter", s);
      k_scatter<<<(unsigned)nb, kPartBlock, 4 * partitions * 4, s>>>(
          pids.get<uint32_t>(), nrows, partitions, ceil_log2(partitions), per_block,
          base.get<uint64_t>(), cc, nullptr);
      KG_HIP(hipGetLastError());
    }
    // partition offsets: the scanned value at each partition's first block
    std::vector<uint64_t> firsts(partitions);
    DevBuf fd(ctx, (uint64_t)partitions * 8);
    k_stride_gather<<<(unsigned)ceil_div(partitions, 256), 256, 0, s>>>(base.get<uint64_t>(), (uint64_t)nb,
                                                              partitions, fd.get<uint64_t>());
    KG_HIP(hipGetLastError());
    KG_HIP(hipMemcpyAsync(firsts.data(), fd.get(), (size_t)partitions * 8, hipMemcpyDeviceToHost, s));
    KG_HIP(hipStreamSynchronize(s));
    for (uint32_t p = 0; p < partitions; p++) out_offsets[p] = (int64_t)firsts[p];
    out_offsets[partitions] = nrows;
    return b;
  }
}
}  // namespace kg

extern "C" {

int kg_repartition(kg_ctx *ctx, int32_t ncols, const kg_column *cols, int64_t nrows,
                   int32_t nkeys, const int32_t *key_cols, uint32_t partitions,
                   int64_t *out_offsets, kg_batch **out) {
  return guarded([&] {
    *out = nullptr;
    *out = split_rows(ctx, ncols, cols, nrows, nkeys, key_cols, nullptr, partitions, out_offsets)
               .release();
  });
}

int kg_split_partitions(kg_ctx *ctx, int32_t ncols, const kg_column *cols, int64_t nrows,
                        const uint8_t *partition_of_row, uint32_t partitions,
                        int64_t *out_offsets, kg_batch **out) {
  return guarded([&] {
    *out = nullptr;
    if (!partition_of_row) fail(KG_ERR_INTERNAL, "no partition ids");
    if (partitions > 256) fail(KG_ERR_UNSUPPORTED, "one-byte partition ids: at most 256");
    *out = split_rows(ctx, ncols, cols, nrows, 0, nullptr, partition_of_row, partitions,
                      out_offsets)
               .release();
  });
}


}  // extern "C"
