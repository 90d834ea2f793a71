// kg_calib.hip -- measured ceilings for the random-access phases: the rate of
// random reads from a device table of a given size (what a hash-table or
// Bloom-filter probe can reach at that size) and of 64-bit global atomic
// adds over a given number of addresses (what a global-table update can
// reach, from all lanes on one word to every lane on its own line).  The
// probe kernels' rates are reported against these, the streaming kernels'
// against HBM bandwidth.
#include <hip/hip_runtime.h>

#include "kg_common.hpp"
#include "kg_device.hpp"

namespace kg {

constexpr int kCalBlock = 256;
constexpr int kCalU = 8;  // independent reads in flight per lane

// `ops` reads of WORDS x 8 bytes at hashed entries of t[0..mask]; the XOR of
// what was read goes to sink only in the (never seen) all-ones case, so the
// loads stay and the kernel writes nothing.
template <int WORDS>
__global__ void __launch_bounds__(kCalBlock) k_calib_read(const uint64_t *t, uint64_t mask,
                                                          int64_t ops, uint64_t salt,
                                                          uint64_t *sink) {
  uint64_t acc = 0;
  const int64_t stride = (int64_t)gridDim.x * kCalBlock * kCalU;
  for (int64_t i0 = ((int64_t)blockIdx.x * kCalBlock + threadIdx.x) * kCalU; i0 < ops;
       i0 += stride) {
    uint64_t v[kCalU];
#pragma unroll
    for (int u = 0; u < kCalU; u++) {
      const uint64_t e = dev::table_hash((uint64_t)(i0 + u), salt) & mask;
      const uint64_t *p = t + e * WORDS;
      if constexpr (WORDS == 1) {
        v[u] = p[0];
      } else if constexpr (WORDS == 2) {
        const ulonglong2 x = *reinterpret_cast<const ulonglong2 *>(p);
        v[u] = x.x ^ x.y;
      } else {
        const ulonglong2 x = *reinterpret_cast<const ulonglong2 *>(p);
        const ulonglong2 y = *reinterpret_cast<const ulonglong2 *>(p + 2);
        v[u] = x.x ^ x.y ^ y.x ^ y.y;
      }
    }
#pragma unroll
    for (int u = 0; u < kCalU; u++) acc ^= v[u];
  }
  if (acc == ~0ull) sink[0] = acc;
}

// `ops` no-return 64-bit atomic adds at hashed words of w[0..mask] (mask 0:
// every add on one word).
__global__ void __launch_bounds__(kCalBlock) k_calib_atomic(uint64_t *w, uint64_t mask,
                                                            int64_t ops, uint64_t salt) {
  const int64_t stride = (int64_t)gridDim.x * kCalBlock;
  for (int64_t i = (int64_t)blockIdx.x * kCalBlock + threadIdx.x; i < ops; i += stride) {
    const uint64_t e = dev::table_hash((uint64_t)i, salt) & mask;
    __hip_atomic_fetch_add(w + e, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Mean ms per launch of `reps` launches after one untimed launch.
template <class F>
static double time_launches(hipStream_t s, int reps, F &&launch) {
  launch();
  KG_HIP(hipGetLastError());
  hipEvent_t a, b;
  KG_HIP(hipEventCreate(&a));
  KG_HIP(hipEventCreate(&b));
  KG_HIP(hipEventRecord(a, s));
  for (int r = 0; r < reps; r++) launch();
  KG_HIP(hipEventRecord(b, s));
  KG_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  KG_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return (double)ms / reps;
}

static bool pow2(uint64_t v) { return v && !(v & (v - 1)); }

}  // namespace kg

using namespace kg;

extern "C" {

int kg_calib_random_read(kg_ctx *ctx, uint64_t table_bytes, int32_t read_bytes, int64_t ops,
                         int32_t reps, double *ms_per_launch) {
  return guarded([&] {
    if (read_bytes != 8 && read_bytes != 16 && read_bytes != 32)
      fail(KG_ERR_UNSUPPORTED, "read_bytes must be 8, 16 or 32");
    if (!pow2(table_bytes) || table_bytes < (uint64_t)read_bytes)
      fail(KG_ERR_UNSUPPORTED, "table_bytes must be a power of two of at least one read");
    if (ops < 1 || reps < 1) fail(KG_ERR_UNSUPPORTED, "ops and reps must be positive");
    StreamScope kg_scope(ctx->stream);
    ops = ceil_div(ops, kCalU) * kCalU;
    DevBuf t(ctx, table_bytes), sink(ctx, 64);
    KG_HIP(hipMemsetAsync(t.get(), 0x5A, table_bytes, ctx->stream));
    const uint64_t mask = table_bytes / read_bytes - 1;
    const unsigned grid = (unsigned)((ctx->num_cus > 0 ? ctx->num_cus : 256) * 8);
    const uint64_t *tp = t.get<uint64_t>();
    uint64_t *sp = sink.get<uint64_t>();
    uint64_t salt = 0x2545F4914F6CDD1DULL;
    *ms_per_launch = time_launches(ctx->stream, reps, [&] {
      salt += 0x9E3779B97F4A7C15ULL;  // other entries on each launch
      if (read_bytes == 8)
        k_calib_read<1><<<grid, kCalBlock, 0, ctx->stream>>>(tp, mask, ops, salt, sp);
      else if (read_bytes == 16)
        k_calib_read<2><<<grid, kCalBlock, 0, ctx->stream>>>(tp, mask, ops, salt, sp);
      else
        k_calib_read<4><<<grid, kCalBlock, 0, ctx->stream>>>(tp, mask, ops, salt, sp);
    });
  });
}

int kg_calib_atomic_add(kg_ctx *ctx, uint64_t addresses, int64_t ops, int32_t reps,
                        double *ms_per_launch) {
  return guarded([&] {
    if (!pow2(addresses)) fail(KG_ERR_UNSUPPORTED, "addresses must be a power of two");
    if (ops < 1 || reps < 1) fail(KG_ERR_UNSUPPORTED, "ops and reps must be positive");
    StreamScope kg_scope(ctx->stream);
    DevBuf w(ctx, addresses * 8);
    KG_HIP(hipMemsetAsync(w.get(), 0, addresses * 8, ctx->stream));
    const unsigned grid = (unsigned)((ctx->num_cus > 0 ? ctx->num_cus : 256) * 8);
    uint64_t *wp = w.get<uint64_t>();
    uint64_t salt = 0x6A09E667F3BCC909ULL;
    *ms_per_launch = time_launches(ctx->stream, reps, [&] {
      salt += 0x9E3779B97F4A7C15ULL;
      k_calib_atomic<<<grid, kCalBlock, 0, ctx->stream>>>(wp, addresses - 1, ops, salt);
    });
  });
}

}  // extern "C"
