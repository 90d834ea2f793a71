// kg_select.hip -- the scan's row stage: the range form of a runtime filter
// (runtime_filter/member.rs Range: [min, max] of the build keys, NULLs never
// pass) as a stable selection vector, and the take of a selection vector
// (operators/select.rs).
//
// Selection is two passes like the Bloom probe: pass 1 streams the column
// and writes a 64-row ballot word per wave and a per-tile count; pass 2
// scans the counts and expands the bitmap into row ids in input order.
#include <hip/hip_runtime.h>

#include "kg_scan.hpp"

#include "kg_common.hpp"
#include "kg_device.hpp"

namespace kg {

constexpr int kSelBlock = 256;
constexpr int kSelTile = kSelBlock * 32;
constexpr int kGatherCols = 32;

__global__ void __launch_bounds__(kSelBlock) k_range_bits(DCol c, int64_t n, int64_t lo, int64_t hi,
                                                          uint64_t *bitmap, uint32_t *tile_count) {
  __shared__ uint32_t s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kSelTile;
  uint32_t cnt = 0;
  const int lane = threadIdx.x & 63;
  for (int k = 0; k < kSelTile / kSelBlock; k++) {
    int64_t i = base + (int64_t)k * kSelBlock + threadIdx.x;
    bool pass = false;
    if (i < n && dev::valid(c, i)) {
      int64_t v = dev::load_i64(c, i);
      pass = v >= lo && v <= hi;
    }
    uint64_t m = __ballot(pass);
    int64_t w0 = base + (int64_t)k * kSelBlock + (threadIdx.x & ~63);
    if (lane == 0 && w0 < n) {
      bitmap[w0 >> 6] = m;
      cnt += (uint32_t)__popcll(m);
    }
  }
  if (lane == 0 && cnt) atomicAdd(&s_cnt, cnt);
  __syncthreads();
  if (threadIdx.x == 0) tile_count[blockIdx.x] = s_cnt;
}

__global__ void __launch_bounds__(kSelBlock) k_bits_to_rows(const uint64_t *bitmap, int64_t n,
                                                            const uint64_t *tile_off,
                                                            int32_t *out_rows) {
  __shared__ uint32_t s_w[kSelBlock / 64];
  int64_t base = (int64_t)blockIdx.x * kSelTile;
  int64_t r0 = base + (int64_t)threadIdx.x * 32;
  uint32_t bits = 0;
  if (r0 < n) {
    uint64_t w = bitmap[r0 >> 6];
    bits = (uint32_t)(((r0 >> 5) & 1) ? (w >> 32) : w);
  }
  uint32_t tot;
  const uint32_t cnt = (uint32_t)__popc(bits);
  const uint32_t off = block_exclusive_scan<kSelBlock>(cnt, s_w, &tot);
  int64_t pos = (int64_t)tile_off[blockIdx.x] + off;
  while (bits) {
    int b = __ffs(bits) - 1;
    bits &= bits - 1;
    out_rows[pos++] = (int32_t)(r0 + b);
  }
}

struct TakeCols {
  DCol src[kGatherCols];
  void *dst[kGatherCols];
  uint32_t *dst_valid[kGatherCols];
  int n;
};

// A wave takes 64 consecutive output rows (whole-wave blocks and strides):
// a nullable column's validity bits of those rows are two whole 32-bit
// words, one ballot stored by lanes 0 and 32.
__global__ void k_take(TakeCols G, const int32_t *rows, int64_t n) {
  const int lane = threadIdx.x & 63;
  for (int64_t b0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); b0 < n;
       b0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = b0 + lane;
    const bool in = o < n;
    const int64_t r = in ? rows[o] : 0;
    for (int c = 0; c < G.n; c++) {
      const DCol &s = G.src[c];
      int64_t j = r + s.offset;
      void *out = G.dst[c];
      if (G.dst_valid[c]) {
        const uint64_t m = __ballot(in && dev::valid(s, r));
        if (lane == 0) G.dst_valid[c][b0 >> 5] = (uint32_t)m;
        if (lane == 32 && b0 + 32 < n) G.dst_valid[c][(b0 >> 5) + 1] = (uint32_t)(m >> 32);
      }
      if (!in) continue;
      switch (s.width) {
        case 1: static_cast<uint8_t *>(out)[o] = static_cast<const uint8_t *>(s.values)[j]; break;
        case 2: static_cast<uint16_t *>(out)[o] = static_cast<const uint16_t *>(s.values)[j]; break;
        case 4: static_cast<uint32_t *>(out)[o] = static_cast<const uint32_t *>(s.values)[j]; break;
        case 8: static_cast<uint64_t *>(out)[o] = static_cast<const uint64_t *>(s.values)[j]; break;
        case 16: {
          const uint64_t *sv = static_cast<const uint64_t *>(s.values) + 2 * j;
          uint64_t *d = static_cast<uint64_t *>(out) + 2 * o;
          d[0] = sv[0];
          d[1] = sv[1];
          break;
        }
        default:
          if ((static_cast<const uint8_t *>(s.values)[j >> 3] >> (j & 7)) & 1)
            atomicOr(static_cast<uint32_t *>(out) + (o >> 5), 1u << (o & 31));
      }
    }
  }
}

}  // namespace kg

using namespace kg;

extern "C" {

int kg_select_range(kg_ctx *ctx, const kg_column *col, int64_t nrows, int64_t lo, int64_t hi,
                    int32_t *out_rows, int64_t *out_count) {
  return guarded([&] {
    *out_count = 0;
    if (nrows <= 0) return;
    if (nrows > INT32_MAX) fail(KG_ERR_UNSUPPORTED, "a batch of 2^31 rows or more");
    check_column(*col, nrows, "range filter column");
    if (!is_int_like(col->type)) fail(KG_ERR_UNSUPPORTED, "range filter on this type");
    hipStream_t s = ctx->stream;
    Staged st(ctx, s);
    DCol c = st.add(*col, nrows);
    int64_t tiles = ceil_div(nrows, kSelTile);
    DevBuf bitmap(ctx, (uint64_t)ceil_div(nrows, 64) * 8 + 8), counts(ctx, (uint64_t)tiles * 4),
        offs(ctx, (uint64_t)(tiles + 1) * 8);
    {
      ProfScope prof("select_range", s);
      k_range_bits<<<(unsigned)tiles, kSelBlock, 0, s>>>(c, nrows, lo, hi, bitmap.get<uint64_t>(),
                                                         counts.get<uint32_t>());
      KG_HIP(hipGetLastError());
    }
    DevBuf temp(ctx, (uint64_t)scan_partials(tiles) * 8);
    exclusive_scan(counts.get<uint32_t>(), offs.get<uint64_t>(), tiles, temp.get<uint64_t>(), s);
    k_bits_to_rows<<<(unsigned)tiles, kSelBlock, 0, s>>>(bitmap.get<uint64_t>(), nrows,
                                                         offs.get<uint64_t>(), out_rows);
    KG_HIP(hipGetLastError());
    uint64_t last_off = 0;
    uint32_t last_cnt = 0;
    KG_HIP(hipMemcpyAsync(&last_off, offs.get<uint64_t>() + tiles - 1, 8, hipMemcpyDeviceToHost, s));
    KG_HIP(hipMemcpyAsync(&last_cnt, counts.get<uint32_t>() + tiles - 1, 4, hipMemcpyDeviceToHost, s));
    KG_HIP(hipStreamSynchronize(s));
    *out_count = (int64_t)(last_off + last_cnt);
  });
}

int kg_gather(kg_ctx *ctx, int32_t ncols, const kg_column *cols, const int32_t *rows, int64_t nrows,
              kg_batch **out) {
  return guarded([&] {
    *out = nullptr;
    if (ncols < 1 || ncols > kGatherCols) fail(KG_ERR_UNSUPPORTED, "1 to 32 columns");
    hipStream_t s = ctx->stream;
    Staged st(ctx, s);
    auto b = std::make_unique<kg_batch>();
    b->ctx = ctx;
    b->nrows = nrows;
    b->cols.resize(ncols);
    TakeCols G{};
    G.n = ncols;
    uint64_t vb = (uint64_t)ceil_div(nrows, 32) * 4 + 4;
    for (int c = 0; c < ncols; c++) {
      check_column(cols[c], cols[c].length, "gather column");
      DCol d = st.add(cols[c], cols[c].length);
      auto &oc = b->cols[c];
      oc.type = d.type;
      oc.precision = d.precision;
      oc.scale = d.scale;
      int w = type_width(d.type);
      uint64_t bytes = w == 0 ? vb : (uint64_t)nrows * w + 16;
      oc.values.reset(ctx, bytes);
      if (w == 0) KG_HIP(hipMemsetAsync(oc.values.get(), 0, bytes, s));
      if (d.validity) {
        oc.validity.reset(ctx, vb);
        KG_HIP(hipMemsetAsync(oc.validity.get(), 0, vb, s));
      }
      G.src[c] = d;
      G.dst[c] = oc.values.get();
      G.dst_valid[c] = d.validity ? oc.validity.get<uint32_t>() : nullptr;
    }
    if (nrows > 0) {
      ProfScope prof("gather", s);
      k_take<<<stream_grid(nrows, 256), 256, 0, s>>>(G, rows, nrows);
      KG_HIP(hipGetLastError());
    }
    KG_HIP(hipStreamSynchronize(s));
    *out = b.release();
  });
}

}  // extern "C"
