// kg_gen.hip -- counter-based synthetic inputs generated in HBM (benchmarks
// and tests).  The integer streams (SplitMix64 output function over
// seed/stream/row counters, a Feistel permutation cycle-walked into range)
// are the same arithmetic as the CPU generators the tests check against;
// the Zipf draw reads a cumulative distribution built once on the host.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kg_common.hpp"
#include "kg_device.hpp"
#include "kg_scan.hpp"

namespace kg {

__host__ __device__ __forceinline__ uint64_t mix64_hd(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t stream_base(uint64_t seed, uint64_t stream) {
  return mix64_hd(seed ^ (stream * 0xD1B54A32D192ED03ULL));
}

// A bijection of [0, n): 4 Feistel rounds on the smallest even bit width
// covering n, cycle-walked back into range.
__device__ __forceinline__ uint64_t perm(uint64_t i, uint64_t n, uint64_t seed) {
  if (n <= 1) return 0;
  int bits = 2;
  while (bits < 64 && (1ULL << bits) < n) bits++;
  if (bits & 1) bits++;
  int half = bits / 2;
  uint64_t mask = (1ULL << half) - 1;
  uint64_t x = i;
  do {
    uint64_t L = x >> half, R = x & mask;
    for (uint64_t r = 0; r < 4; r++) {
      uint64_t f = mix64_hd(R ^ (seed * 0x9E3779B97F4A7C15ULL) ^ (r << 56)) & mask;
      uint64_t t = L ^ f;
      L = R;
      R = t;
    }
    x = (L << half) | R;
  } while (x >= n);
  return x;
}

__global__ void k_gen_join_build(uint64_t seed, int64_t lo, int64_t hi, int64_t n_total,
                                 int64_t *key, int64_t *val) {
  uint64_t b1 = stream_base(seed, 1);
  for (int64_t i = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hi;
       i += (int64_t)gridDim.x * blockDim.x) {
    key[i - lo] = (int64_t)perm((uint64_t)i, (uint64_t)n_total, seed);
    if (val) val[i - lo] = (int64_t)dev::below(dev::rand_at(b1, (uint64_t)i), 1000000);
  }
}

__global__ void k_gen_join_probe(uint64_t seed, int64_t lo, int64_t hi, int64_t key_range,
                                 uint32_t null_permille, int64_t *fk, uint32_t *fk_valid,
                                 int64_t *amt) {
  uint64_t b0 = stream_base(seed, 0), b1 = stream_base(seed, 1), b2 = stream_base(seed, 2);
  // one thread per 32 rows: it owns one validity word
  int64_t n = hi - lo;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w * 32 < n;
       w += (int64_t)gridDim.x * blockDim.x) {
    uint32_t bits = 0;
    for (int k = 0; k < 32; k++) {
      int64_t r = w * 32 + k;
      if (r >= n) break;
      uint64_t u = (uint64_t)(lo + r);
      if (fk) fk[r] = (int64_t)dev::below(dev::rand_at(b0, u), (uint64_t)key_range);
      if (dev::below(dev::rand_at(b1, u), 1000) >= null_permille) bits |= 1u << k;
      if (amt) amt[r] = 1 + (int64_t)dev::below(dev::rand_at(b2, u), 10000);
    }
    if (fk_valid) fk_valid[w] = bits;
  }
}

__device__ __forceinline__ uint64_t zipf_rank(const double *cdf, uint64_t n, double u) {
  uint64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] > u) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

__global__ void k_gen_agg(uint64_t seed, int64_t lo, int64_t hi, const double *cdf,
                          uint64_t ndistinct, int64_t *key, int64_t *dec, double *f, int32_t *i32,
                          int8_t *k8) {
  uint64_t b0 = stream_base(seed, 0), b1 = stream_base(seed, 1), b2 = stream_base(seed, 2),
           b3 = stream_base(seed, 3), b4 = stream_base(seed, 4);
  for (int64_t i = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hi;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t u = (uint64_t)i;
    if (key) {
      uint64_t r = zipf_rank(cdf, ndistinct, dev::unit(dev::rand_at(b0, u)));
      key[i - lo] = (int64_t)perm(r, ndistinct, seed ^ 0x5BD1E995ULL);
    }
    if (dec) dec[i - lo] = 1 + (int64_t)dev::below(dev::rand_at(b1, u), 9999999);
    if (f) f[i - lo] = dev::unit(dev::rand_at(b2, u));
    if (i32) i32[i - lo] = (int32_t)dev::below(dev::rand_at(b3, u), 51);
    if (k8) k8[i - lo] = (int8_t)dev::below(dev::rand_at(b4, u), 4);
  }
}

// Zipf keys whose rank-to-key permutation comes from key_seed: two columns
// with different row seeds and one key_seed draw from one distribution (the
// same keys are hot on both sides of a join).  key_seed == seed gives
// k_gen_agg's key column.
__global__ void k_gen_zipf_keys(uint64_t seed, uint64_t key_seed, int64_t lo, int64_t hi,
                                const double *cdf, uint64_t ndistinct, int64_t *key) {
  const uint64_t b0 = stream_base(seed, 0);
  for (int64_t i = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hi;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t r = zipf_rank(cdf, ndistinct, dev::unit(dev::rand_at(b0, (uint64_t)i)));
    key[i - lo] = (int64_t)perm(r, ndistinct, key_seed ^ 0x5BD1E995ULL);
  }
}

// The Zipf table on the device: w[k] = (k+1)^-s in place, an exclusive
// scan, then cdf[k] = (prefix + w[k]) / total with the last entry 1.
__global__ void k_zipf_w(double *w, uint64_t n, double s) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n;
       k += (uint64_t)gridDim.x * blockDim.x)
    w[k] = pow((double)(k + 1), -s);
}
__global__ void k_zipf_norm(double *cdf, uint64_t n, double s, const double *total) {
  const double inv = 1.0 / *total;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n;
       k += (uint64_t)gridDim.x * blockDim.x)
    cdf[k] = k + 1 == n ? 1.0 : (cdf[k] + pow((double)(k + 1), -s)) * inv;
}

}  // namespace kg

using namespace kg;

extern "C" {

int kg_zipf_cdf_device(kg_ctx *ctx, uint64_t n, double s, double *cdf_dev) {
  return guarded([&] {
    if (n == 0) fail(KG_ERR_INTERNAL, "empty distribution");
    StreamScope kg_scope(ctx->stream);
    DevBuf partials(ctx, (uint64_t)(scan_partials((int64_t)n) + 1) * 8);
    double *total = partials.get<double>() + scan_partials((int64_t)n);
    k_zipf_w<<<stream_grid((int64_t)n, 256), 256, 0, ctx->stream>>>(cdf_dev, n, s);
    KG_HIP(hipGetLastError());
    exclusive_scan(cdf_dev, cdf_dev, (int64_t)n, partials.get<double>(), ctx->stream,
                   nullptr, total);
    k_zipf_norm<<<stream_grid((int64_t)n, 256), 256, 0, ctx->stream>>>(cdf_dev, n, s, total);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int kg_gen_zipf_keys(kg_ctx *ctx, uint64_t seed, uint64_t key_seed, int64_t lo, int64_t hi,
                     const double *cdf_dev, uint64_t ndistinct, int64_t *key_dev) {
  return guarded([&] {
    if (hi <= lo) return;
    k_gen_zipf_keys<<<stream_grid(hi - lo, 256), 256, 0, ctx->stream>>>(
        seed, key_seed, lo, hi, cdf_dev, ndistinct, key_dev);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int kg_gen_join_build(kg_ctx *ctx, uint64_t seed, int64_t lo, int64_t hi, int64_t n_total,
                      int64_t *key_dev, int64_t *val_dev) {
  return guarded([&] {
    if (hi <= lo) return;
    k_gen_join_build<<<stream_grid(hi - lo, 256), 256, 0, ctx->stream>>>(seed, lo, hi, n_total,
                                                                       key_dev, val_dev);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int kg_gen_join_probe(kg_ctx *ctx, uint64_t seed, int64_t lo, int64_t hi, int64_t key_range,
                      uint32_t null_permille, int64_t *fk_dev, uint8_t *fk_validity_dev,
                      int64_t *amt_dev) {
  return guarded([&] {
    if (hi <= lo) return;
    if (fk_validity_dev && ((uintptr_t)fk_validity_dev & 3))
      fail(KG_ERR_INTERNAL, "validity buffer must be 4-byte aligned");
    k_gen_join_probe<<<stream_grid(ceil_div(hi - lo, 32), 256, 1), 256, 0, ctx->stream>>>(
        seed, lo, hi, key_range, null_permille, fk_dev, (uint32_t *)fk_validity_dev, amt_dev);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int kg_zipf_cdf(uint64_t n, double s, double *cdf) {
  return guarded([&] {
    if (n == 0) fail(KG_ERR_INTERNAL, "empty distribution");
    double acc = 0.0;
    for (uint64_t k = 0; k < n; k++) {
      acc += std::pow((double)(k + 1), -s);
      cdf[k] = acc;
    }
    double inv = 1.0 / acc;
    for (uint64_t k = 0; k < n; k++) cdf[k] *= inv;
    cdf[n - 1] = 1.0;
  });
}

int kg_gen_agg(kg_ctx *ctx, uint64_t seed, int64_t lo, int64_t hi, const double *cdf_dev,
               uint64_t ndistinct, int64_t *key_dev, int64_t *dec_dev, double *f_dev,
               int32_t *i32_dev, int8_t *k8_dev) {
  return guarded([&] {
    if (hi <= lo) return;
    k_gen_agg<<<stream_grid(hi - lo, 256), 256, 0, ctx->stream>>>(
        seed, lo, hi, cdf_dev, ndistinct, key_dev, dec_dev, f_dev, i32_dev, k8_dev);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
