// kg_device.hpp -- device functions shared by the kernels: column loads,
// the routing hash H_kg (XXH64 of normalized key bytes), the row-layout
// table hash h and its finalizer g, partition_of, and the SBBF mask.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kg_common.hpp"

namespace kg {
namespace dev {

// ---------------------------------------------------------------- columns
__host__ __device__ __forceinline__ bool is_f(int t) { return t == KG_T_FLOAT32 || t == KG_T_FLOAT64; }

__device__ __forceinline__ bool valid(const DCol &c, int64_t i) {
  if (!c.validity) return true;
  int64_t j = i + c.offset;
  return (c.validity[j >> 3] >> (j & 7)) & 1;
}

// The 64-bit word of an 8-byte column at row i, loaded with no condition.
// Streaming kernels clamp i into their row range and mask the result after
// the loads of all their rows are issued: a load behind a per-lane branch
// (a validity test, a range test) waits for the loads before it.
__device__ __forceinline__ uint64_t load_w64(const DCol &c, int64_t i) {
  return static_cast<const uint64_t *>(c.values)[i + c.offset];
}

// Integer-like value widened to int64 (bool, signed/unsigned ints, dates,
// timestamps, decimal64).  UINT64 keeps its bit pattern.
__device__ __forceinline__ int64_t load_i64(const DCol &c, int64_t i) {
  int64_t j = i + c.offset;
  switch (c.type) {
    case KG_T_BOOL: return (static_cast<const uint8_t *>(c.values)[j >> 3] >> (j & 7)) & 1;
    case KG_T_INT8: return static_cast<const int8_t *>(c.values)[j];
    case KG_T_UINT8: return static_cast<const uint8_t *>(c.values)[j];
    case KG_T_INT16: return static_cast<const int16_t *>(c.values)[j];
    case KG_T_UINT16: return static_cast<const uint16_t *>(c.values)[j];
    case KG_T_INT32:
    case KG_T_DATE32: return static_cast<const int32_t *>(c.values)[j];
    case KG_T_UINT32: return static_cast<const uint32_t *>(c.values)[j];
    case KG_T_DECIMAL128: return static_cast<const int64_t *>(c.values)[2 * j];
    default: return static_cast<const int64_t *>(c.values)[j];
  }
}

__device__ __forceinline__ double load_f64(const DCol &c, int64_t i) {
  int64_t j = i + c.offset;
  if (c.type == KG_T_FLOAT64) return static_cast<const double *>(c.values)[j];
  if (c.type == KG_T_FLOAT32) return static_cast<const float *>(c.values)[j];
  return (double)load_i64(c, i);
}

// 128-bit value (decimal128, or any integer-like widened with sign).
__device__ __forceinline__ void load_i128(const DCol &c, int64_t i, uint64_t &lo, int64_t &hi) {
  if (c.type == KG_T_DECIMAL128) {
    const uint64_t *w = static_cast<const uint64_t *>(c.values) + 2 * (i + c.offset);
    lo = w[0];
    hi = (int64_t)w[1];
  } else {
    int64_t v = load_i64(c, i);
    lo = (uint64_t)v;
    hi = c.type == KG_T_UINT64 ? 0 : (v >> 63);
  }
}

// Spark key normalization of floats: every NaN is one NaN, -0.0 is 0.0
// (rows/layout.rs:1147-1153, routing.rs:438-450).
__device__ __forceinline__ uint64_t norm_f64_bits(double f) {
  if (f != f) return 0x7FF8000000000000ULL;
  f = f + 0.0;
  return (uint64_t)__double_as_longlong(f);
}
__device__ __forceinline__ uint32_t norm_f32_bits(float f) {
  if (f != f) return 0x7FC00000u;
  f = f + 0.0f;
  return (uint32_t)__float_as_uint(f);
}

// ---------------------------------------------------------------- hashing
constexpr uint64_t P1 = 0x9E3779B185EBCA87ULL;
constexpr uint64_t P2 = 0xC2B2AE3D27D4EB4FULL;
constexpr uint64_t P3 = 0x165667B19E3779F9ULL;
constexpr uint64_t P4 = 0x85EBCA77C2B2AE63ULL;
constexpr uint64_t P5 = 0x27D4EB2F165667C5ULL;
constexpr uint64_t NULL_HASH = 0xA5A55A5AC3C33C3CULL;  // routing.rs:78, layout.rs:43
constexpr uint64_t FOLD_INIT = 0x243F6A8885A308D3ULL;  // routing.rs:81
constexpr uint64_t FOLD_MUL = 0xB7E151628AED2A6BULL;   // routing.rs:84
constexpr uint64_t L_MIX = 0x243F6A8885A308D3ULL;      // layout.rs:45
constexpr uint64_t L_FINISH = 0xB7E151628AED2A6BULL;   // layout.rs:46
constexpr uint64_t L_FIN = 0x6A09E667F3BCC909ULL;      // layout.rs:50

__host__ __device__ __forceinline__ uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

__host__ __device__ __forceinline__ uint64_t avalanche(uint64_t h) {
  h ^= h >> 33;
  h *= P2;
  h ^= h >> 29;
  h *= P3;
  return h ^ (h >> 32);
}

__host__ __device__ __forceinline__ uint64_t lane8(uint64_t h, uint64_t k) {
  return rotl(h ^ (rotl(k * P2, 31) * P1), 27) * P1 + P4;
}

// XXH64 seed 0 of `width` (1, 2, 4, 8, 16) little-endian bytes held in
// (lo, hi): the short-input path of routing.rs:112-142.
__host__ __device__ __forceinline__ uint64_t xxh64_fixed(uint64_t lo, uint64_t hi, int width) {
  uint64_t h = P5 + (uint64_t)width;
  if (width == 16) {
    h = lane8(lane8(h, lo), hi);
  } else if (width == 8) {
    h = lane8(h, lo);
  } else if (width == 4) {
    h = rotl(h ^ ((lo & 0xFFFFFFFFULL) * P1), 23) * P2 + P3;
  } else {
    for (int k = 0; k < width; k++) {
      h = rotl(h ^ ((lo & 0xFF) * P5), 11) * P1;
      lo >>= 8;
    }
  }
  return avalanche(h);
}

// XXH64 with a seed of a 4- or 8-byte value (Spark's xxhash64 of an int or
// a bigint uses seed 42): the same short-input path with h = seed + P5 + len.
__device__ __forceinline__ uint64_t xxh64_seeded(uint64_t v, int width, uint64_t seed) {
  uint64_t h = seed + P5 + (uint64_t)width;
  if (width == 8) {
    h = lane8(h, v);
  } else {
    h = rotl(h ^ ((v & 0xFFFFFFFFULL) * P1), 23) * P2 + P3;
  }
  return avalanche(h);
}

// Murmur3 x86_32 as Spark runs it (crates/krishiv-kernels/src/hash/murmur3.rs:9-52).
__device__ __forceinline__ uint32_t m3_rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
__device__ __forceinline__ uint32_t m3_mix_k1(uint32_t k1) {
  return m3_rotl(k1 * 0xcc9e2d51u, 15) * 0x1b873593u;
}
__device__ __forceinline__ uint32_t m3_mix_h1(uint32_t h1, uint32_t k1) {
  return m3_rotl(h1 ^ k1, 13) * 5u + 0xe6546b64u;
}
__device__ __forceinline__ uint32_t m3_fmix(uint32_t h1, uint32_t len) {
  uint32_t h = h1 ^ len;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  return h ^ (h >> 16);
}
// hash_long: the low, then the high 32 bits
__device__ __forceinline__ int32_t murmur3_long(uint64_t v, int32_t seed) {
  uint32_t h1 = m3_mix_h1((uint32_t)seed, m3_mix_k1((uint32_t)v));
  h1 = m3_mix_h1(h1, m3_mix_k1((uint32_t)(v >> 32)));
  return (int32_t)m3_fmix(h1, 8);
}

// The folded multiply (layout.rs:1131-1139; fold_column, routing.rs:172-183).
__device__ __forceinline__ uint64_t fold(uint64_t a, uint64_t b) {
  return (a * b) ^ __umul64hi(a, b);
}

// H_kg column hash of row i (routing.rs:332-450): NULL_HASH for a NULL.
__device__ __forceinline__ uint64_t value_hash(const DCol &c, int64_t i) {
  if (!valid(c, i)) return NULL_HASH;
  int64_t j = i + c.offset;
  switch (c.type) {
    case KG_T_BOOL:
    case KG_T_INT8:
    case KG_T_UINT8: return xxh64_fixed((uint64_t)load_i64(c, i) & 0xFF, 0, 1);
    case KG_T_INT16:
    case KG_T_UINT16: return xxh64_fixed(static_cast<const uint16_t *>(c.values)[j], 0, 2);
    case KG_T_INT32:
    case KG_T_UINT32:
    case KG_T_DATE32: return xxh64_fixed(static_cast<const uint32_t *>(c.values)[j], 0, 4);
    case KG_T_FLOAT32:
      return xxh64_fixed(norm_f32_bits(static_cast<const float *>(c.values)[j]), 0, 4);
    case KG_T_FLOAT64:
      return xxh64_fixed(norm_f64_bits(static_cast<const double *>(c.values)[j]), 0, 8);
    case KG_T_DECIMAL64: {
      int64_t v = static_cast<const int64_t *>(c.values)[j];
      return xxh64_fixed((uint64_t)v, (uint64_t)(v >> 63), 16);
    }
    case KG_T_DECIMAL128: {
      const uint64_t *w = static_cast<const uint64_t *>(c.values) + 2 * j;
      return xxh64_fixed(w[0], w[1], 16);
    }
    default: return xxh64_fixed(static_cast<const uint64_t *>(c.values)[j], 0, 8);
  }
}

// partition_of (routing.rs:217-231)
__device__ __forceinline__ uint32_t partition_of(uint64_t h, uint32_t n) {
  if (n <= 1) return 0;
  if (n <= 4096) return (uint32_t)(((h >> 52) * (uint64_t)n) >> 12);
  return (uint32_t)__umul64hi(h, (uint64_t)n);
}

// The row-layout stored hash h of one 64-bit key word (layout.rs:487-553:
// h = fold(fold(0 ^ v, MIX), FINISH); fold(0, MIX) = 0) and of a NULL.
__device__ __forceinline__ uint64_t row_hash1(uint64_t v) { return fold(fold(v, L_MIX), L_FINISH); }
__device__ __forceinline__ uint64_t row_hash_null() { return fold(fold(NULL_HASH, L_MIX), L_FINISH); }
// Two key words, in key order.
__device__ __forceinline__ uint64_t row_hash2(uint64_t a, uint64_t b) {
  return fold(fold(fold(a, L_MIX) ^ b, L_MIX), L_FINISH);
}
// finalize (layout.rs:1080-1082): g = fold(h ^ seed, FIN)
__device__ __forceinline__ uint64_t finalize(uint64_t h, uint64_t seed) { return fold(h ^ seed, L_FIN); }

// Placement hash of this library's own device tables (join directories,
// aggregate tables): the SplitMix64 finalizer (Steele, Lea and Flood,
// "Fast splittable pseudorandom number generators", 2014) of key ^ seed.
// Slot placement is internal to a table and never visible in results, so it
// need not be the reference's row hash; this one costs two 64-bit
// multiplies where h and g take three 64x64->128 products.
__device__ __forceinline__ uint64_t table_hash(uint64_t key, uint64_t seed) {
  uint64_t z = key ^ seed;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------- SBBF
__host__ __device__ __forceinline__ uint32_t sbbf_salt(int w) {
  // bloom.rs:15-24 (the Parquet specification's salts)
  switch (w) {
    case 0: return 0x47b6137bu;
    case 1: return 0x44974d91u;
    case 2: return 0x8824ad5bu;
    case 3: return 0xa2b7289du;
    case 4: return 0x705495c7u;
    case 5: return 0x2df1424bu;
    case 6: return 0x9efc4947u;
    default: return 0x5c6bfb31u;
  }
}

// RuntimeBloom::check_hash (bloom.rs:128-200): the block of h, (upper 32
// bits * blocks) >> 32, holds all eight bits of its mask.
__device__ __forceinline__ bool sbbf_test(const uint32_t *words, uint64_t nblocks, uint64_t h) {
  const uint4 *b = reinterpret_cast<const uint4 *>(words + (((h >> 32) * nblocks) >> 32) * 8);
  const uint4 x = b[0], y = b[1];
  const uint32_t k = (uint32_t)h;
  uint32_t miss = 0;
  miss |= (1u << ((k * sbbf_salt(0)) >> 27)) & ~x.x;
  miss |= (1u << ((k * sbbf_salt(1)) >> 27)) & ~x.y;
  miss |= (1u << ((k * sbbf_salt(2)) >> 27)) & ~x.z;
  miss |= (1u << ((k * sbbf_salt(3)) >> 27)) & ~x.w;
  miss |= (1u << ((k * sbbf_salt(4)) >> 27)) & ~y.x;
  miss |= (1u << ((k * sbbf_salt(5)) >> 27)) & ~y.y;
  miss |= (1u << ((k * sbbf_salt(6)) >> 27)) & ~y.z;
  miss |= (1u << ((k * sbbf_salt(7)) >> 27)) & ~y.w;
  return miss == 0;
}

// RuntimeBloom::insert_hash (bloom.rs:134-145) with atomics (OR is order-free).
__device__ __forceinline__ void sbbf_put(uint32_t *words, uint64_t nblocks, uint64_t h) {
  uint32_t *b = words + (((h >> 32) * nblocks) >> 32) * 8;
  const uint32_t k = (uint32_t)h;
#pragma unroll
  for (int w = 0; w < 8; w++) atomicOr(b + w, 1u << ((k * sbbf_salt(w)) >> 27));
}

// ---------------------------------------------------------------- generators
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t rand_at(uint64_t base, uint64_t i) {
  return mix64(base + (i + 1) * 0x9E3779B97F4A7C15ULL);
}
__device__ __forceinline__ uint64_t below(uint64_t r, uint64_t n) { return __umul64hi(r, n); }
__device__ __forceinline__ double unit(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

}  // namespace dev
}  // namespace kg
