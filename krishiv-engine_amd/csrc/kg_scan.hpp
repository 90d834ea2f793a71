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

// Segment sums: partials[b] = sum of in[b * 4096, (b + 1) * 4096).
template <class TI>
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const TI *in, int64_t n,
                                                            uint64_t *partials) {
  __shared__ uint64_t s_w[kScanBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanSegment;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
    acc += i < n ? (uint64_t)in[i] : 0;
  }
  uint64_t tot;
  (void)block_exclusive_scan<kScanBlock>(acc, s_w, &tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// One 1024-lane workgroup turns the nb segment sums into exclusive offsets
// in place, starting from *carry_in (0 without one); *total_out (if given)
// gets the grand total including the carry.
__global__ void __launch_bounds__(1024) k_scan_partials(uint64_t *partials, int64_t nb,
                                                        const uint64_t *carry_in,
                                                        uint64_t *total_out) {
  __shared__ uint64_t s_w[16];
  const int64_t per = (nb + 1023) / 1024;
  const int64_t lo = (int64_t)threadIdx.x * per;
  const int64_t hi = lo + per < nb ? lo + per : nb;
  uint64_t mine = 0;
  for (int64_t i = lo; i < hi; i++) mine += partials[i];
  uint64_t tot;
  uint64_t run = block_exclusive_scan<1024>(mine, s_w, &tot) + (carry_in ? *carry_in : 0);
  for (int64_t i = lo; i < hi; i++) {
    const uint64_t v = partials[i];
    partials[i] = run;
    run += v;
  }
  if (threadIdx.x == 0 && total_out) *total_out = tot + (carry_in ? *carry_in : 0);
}

// Segment scans: out[i] = partials[segment] + the sum of the segment's
// inputs before i.  Each lane owns kScanItems consecutive inputs.
template <class TI>
__global__ void __launch_bounds__(kScanBlock) k_scan_down(const TI *in, int64_t n,
                                                          const uint64_t *partials, uint64_t *out) {
  __shared__ uint64_t s_w[kScanBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanSegment + (int64_t)threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = base + k;
    v[k] = i < n ? (uint64_t)in[i] : 0;
    acc += v[k];
  }
  uint64_t tot;
  uint64_t run = block_exclusive_scan<kScanBlock>(acc, s_w, &tot) + partials[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    const int64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

}  // namespace

// Segment sums a device-wide scan of n inputs needs (entries of uint64).
inline int64_t scan_partials(int64_t n) { return ceil_div(n > 0 ? n : 1, kScanSegment); }

// out[i] = carry + sum of in[0, i) for i < n (carry = *carry_in, or 0),
// enqueued on s; partials must hold scan_partials(n) entries.  *total_out
// (device, optional) gets carry + the sum of all n inputs.
template <class TI>
inline void exclusive_scan(const TI *in, uint64_t *out, int64_t n, uint64_t *partials,
                           hipStream_t s, const uint64_t *carry_in = nullptr,
                           uint64_t *total_out = nullptr) {
  if (n <= 0) return;
  const int64_t nb = scan_partials(n);
  k_scan_reduce<TI><<<(unsigned)nb, kScanBlock, 0, s>>>(in, n, partials);
  k_scan_partials<<<1, 1024, 0, s>>>(partials, nb, carry_in, total_out);
  k_scan_down<TI><<<(unsigned)nb, kScanBlock, 0, s>>>(in, n, partials, out);
  KG_HIP(hipGetLastError());
}

}  // namespace kg
