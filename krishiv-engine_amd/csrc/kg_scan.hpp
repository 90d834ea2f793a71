// kg_scan.hpp -- prefix sums written for 64-lane wavefronts: an in-wave
// inclusive scan by shuffles, wave totals through LDS, and a device-wide
// exclusive scan in three launches (segment sums, one workgroup over the
// segment sums, segment scans) whose only scratch is a caller-provided
// array of segment sums -- so callers that must not allocate can pass
// memory they own.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kg_common.hpp"

namespace kg {
namespace {

// Exclusive scan of v over the workgroup (BLOCK lanes, a multiple of 64);
// s_waves holds BLOCK / 64 entries of LDS.  *total gets the workgroup sum.
// Every lane of the workgroup must call it.
template <int BLOCK, class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *s_waves, T *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) s_waves[wave] = x;
  __syncthreads();
  T base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / 64; w++) {
    const T t = s_waves[w];
    base += w < wave ? t : (T)0;
    tot += t;
  }
  __syncthreads();  // s_waves may be reused by the caller
  *total = tot;
  return base + x - v;
}

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int64_t kScanSegment = (int64_t)kScanBlock * kScanItems;  // 4096 inputs a workgroup

// Segment sums: partials[b] = sum of in[b * 4096, (b + 1) * 4096), added in
// TA (uint64 for counts; double for the Zipf table).
template <class TI, class TA>
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const TI *in, int64_t n,
                                                            TA *partials) {
  __shared__ TA s_w[kScanBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanSegment;
  TA acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
    acc += i < n ? (TA)in[i] : (TA)0;
  }
  TA tot;
  (void)block_exclusive_scan<kScanBlock>(acc, s_w, &tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// One 1024-lane workgroup turns the nb segment sums into exclusive offsets
// in place, starting from *carry_in (0 without one); *total_out (if given)
// gets the grand total including the carry.
template <class TA>
__global__ void __launch_bounds__(1024) k_scan_partials(TA *partials, int64_t nb,
                                                        const TA *carry_in, TA *total_out) {
  __shared__ TA s_w[16];
  const int64_t per = (nb + 1023) / 1024;
  const int64_t lo = (int64_t)threadIdx.x * per;
  const int64_t hi = lo + per < nb ? lo + per : nb;
  TA mine = 0;
  for (int64_t i = lo; i < hi; i++) mine += partials[i];
  TA tot;
  TA run = block_exclusive_scan<1024>(mine, s_w, &tot) + (carry_in ? *carry_in : (TA)0);
  for (int64_t i = lo; i < hi; i++) {
    const TA v = partials[i];
    partials[i] = run;
    run += v;
  }
  if (threadIdx.x == 0 && total_out) *total_out = tot + (carry_in ? *carry_in : (TA)0);
}

// Segment scans: out[i] = partials[segment] + the sum of the segment's
// inputs before i.  Each lane owns kScanItems consecutive inputs, read before
// any is written (so out may be in when TI is TA).
template <class TI, class TA>
__global__ void __launch_bounds__(kScanBlock) k_scan_down(const TI *in, int64_t n,
                                                          const TA *partials, TA *out) {
  __shared__ TA s_w[kScanBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanSegment + (int64_t)threadIdx.x * kScanItems;
  TA v[kScanItems];
  TA acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = base + k;
    v[k] = i < n ? (TA)in[i] : (TA)0;
    acc += v[k];
  }
  TA tot;
  TA run = block_exclusive_scan<kScanBlock>(acc, s_w, &tot) + partials[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

// A row bitmap (bit r % 64 of word r / 64: row r passes) expanded into the
// passing row ids, in order: tile t holds rows [t * BLOCK * 32, (t + 1) *
// BLOCK * 32), its first output at tile_off[t] (an exclusive scan of the
// tiles' counts); row ids are written as row_base + row.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_bitmap_rows(const uint64_t *bitmap, int64_t n,
                                                       const uint64_t *tile_off, int64_t row_base,
                                                       int32_t *out_rows) {
  __shared__ uint32_t s_w[BLOCK / 64];
  const int64_t r0 = (int64_t)blockIdx.x * BLOCK * 32 + (int64_t)threadIdx.x * 32;
  uint32_t bits = 0;
  if (r0 < n) {
    const uint64_t w = bitmap[r0 >> 6];
    bits = (uint32_t)(((r0 >> 5) & 1) ? (w >> 32) : w);
  }
  uint32_t tot;
  const uint32_t off = block_exclusive_scan<BLOCK>((uint32_t)__popc(bits), s_w, &tot);
  int64_t pos = (int64_t)tile_off[blockIdx.x] + off;
  while (bits) {
    const int b = __ffs(bits) - 1;
    bits &= bits - 1;
    out_rows[pos++] = (int32_t)(row_base + r0 + b);
  }
}

}  // namespace

// Segment sums a device-wide scan of n inputs needs (entries of uint64).
inline int64_t scan_partials(int64_t n) { return ceil_div(n > 0 ? n : 1, kScanSegment); }

// out[i] = carry + sum of in[0, i) for i < n (carry = *carry_in, or 0), the
// sums taken in TA, enqueued on s; partials must hold scan_partials(n)
// entries.  *total_out (device, optional) gets carry + the sum of all n
// inputs.
template <class T>
struct same_type {
  using type = T;
};  // keeps the carry arguments out of template deduction (nullptr works)

template <class TI, class TA>
inline void exclusive_scan(const TI *in, TA *out, int64_t n, TA *partials, hipStream_t s,
                           const typename same_type<TA>::type *carry_in = nullptr,
                           typename same_type<TA>::type *total_out = nullptr) {
  if (n <= 0) return;
  const int64_t nb = scan_partials(n);
  k_scan_reduce<TI, TA><<<(unsigned)nb, kScanBlock, 0, s>>>(in, n, partials);
  k_scan_partials<TA><<<1, 1024, 0, s>>>(partials, nb, carry_in, total_out);
  k_scan_down<TI, TA><<<(unsigned)nb, kScanBlock, 0, s>>>(in, n, partials, out);
  KG_HIP(hipGetLastError());
}

}  // namespace kg
