// Device-wide exclusive prefix sums in two launches: per-block sums, then a
// down-sweep in which every block first adds up the sums of the blocks
// before it (at most a few thousand values, read from L2) and scans its own
// elements.  256 threads x 16 items = 4096 elements per block.  Element type
// is int32 or Int2 (two independent int32 sums scanned together).
#pragma once

#include <algorithm>
#include <cstdint>

#include "gsmi_device.h"

namespace gsmi {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

struct Int2 {
  int32_t x, y;
};
__device__ __forceinline__ Int2 operator+(Int2 a, Int2 b) { return Int2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ Int2 shfl_up(Int2 v, int d) { return Int2{__shfl_up(v.x, d), __shfl_up(v.y, d)}; }
__device__ __forceinline__ int32_t shfl_up(int32_t v, int d) { return __shfl_up(v, d); }
template <typename T>
__device__ __forceinline__ T zero_of();
template <>
__device__ __forceinline__ int32_t zero_of<int32_t>() {
  return 0;
}
template <>
__device__ __forceinline__ Int2 zero_of<Int2>() {
  return Int2{0, 0};
}

// Inclusive wave scan, then block-exclusive prefix of the per-thread value.
// Returns the thread's exclusive prefix; *total receives the block sum.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* s_wave, T* total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  T inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = shfl_up(inc, d);
    if (lane >= d) inc = inc + o;
  }
  if (lane == 63) s_wave[w] = inc;
  __syncthreads();
  T pre = zero_of<T>();
  T tot = zero_of<T>();
#pragma unroll
  for (int i = 0; i < kScanThreads / 64; ++i) {
    if (i < w) pre = pre + s_wave[i];
    tot = tot + s_wave[i];
  }
  __syncthreads();
  *total = tot;
  T excl = shfl_up(inc, 1);
  if (lane == 0) excl = zero_of<T>();
  return pre + excl;
}

// Per-block sums; whole, 16-byte aligned tiles are read as one 16-byte
// vector per thread and step.
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T* __restrict__ in, int64_t n,
                                                              T* __restrict__ partial) {
  constexpr int kPer = 16 / (int)sizeof(T);  // elements per vector
  __shared__ T s_wave[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  T s = zero_of<T>();
  if (base + kScanTile <= n && (reinterpret_cast<uintptr_t>(in) & 15) == 0) {
    const uint4* in4 = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
    for (int k = 0; k < kScanTile / kPer / kScanThreads; ++k) {
      const uint4 raw = in4[threadIdx.x + k * kScanThreads];
      T v[kPer];
      __builtin_memcpy(v, &raw, sizeof(raw));
#pragma unroll
      for (int i = 0; i < kPer; ++i) s = s + v[i];
    }
  } else {
    for (int i = threadIdx.x; i < kScanTile; i += kScanThreads) {
      const int64_t p = base + i;
      if (p < n) s = s + in[p];
    }
  }
  T tot;
  block_exclusive_scan<T>(s, s_wave, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single block: exclusive scan of the partials in place (for launches of
// more than kScanSumBlocks blocks, where per-block summing would be slow).
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(T* __restrict__ partial, int64_t nb) {
  __shared__ T s_wave[kScanThreads / 64];
  T carry = zero_of<T>();
  for (int64_t c = 0; c < nb; c += kScanThreads) {
    const int64_t p = c + threadIdx.x;
    const T v = p < nb ? partial[p] : zero_of<T>();
    T tot;
    const T ex = block_exclusive_scan<T>(v, s_wave, &tot);
    if (p < nb) partial[p] = carry + ex;
    carry = carry + tot;
  }
}

constexpr int64_t kScanSumBlocks = 4096;

// Each thread owns kScanItems consecutive elements of the block's tile, moved
// as 16-byte vectors when the tile is whole and both arrays are 16-byte
// aligned.  Block b's offset is the sum of partial[0..b-1] (partial[b] when
// the partials are already scanned); the last block writes the grand total.
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const T* in, int64_t n, const T* __restrict__ partial,
                                                            bool scanned, T* out, T* __restrict__ total) {
  static_assert(sizeof(T) % 4 == 0 && (kScanItems * sizeof(T)) % 16 == 0, "vector width");
  constexpr int kVec = kScanItems * (int)sizeof(T) / 16;
  __shared__ T s_wave[kScanThreads / 64];
  const int64_t b = blockIdx.x;
  T offset;
  if (scanned) {
    offset = partial[b];
  } else {
    T pre = zero_of<T>();
    for (int64_t i = threadIdx.x; i < b; i += kScanThreads) pre = pre + partial[i];
    block_exclusive_scan<T>(pre, s_wave, &offset);
  }

  const int64_t base = b * kScanTile + (int64_t)threadIdx.x * kScanItems;
  const bool vec =
      (b + 1) * kScanTile <= n && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
  T v[kScanItems];
  if (vec) {
    uint4 raw[kVec];
#pragma unroll
    for (int k = 0; k < kVec; ++k) raw[k] = reinterpret_cast<const uint4*>(in + base)[k];
    __builtin_memcpy(v, raw, sizeof(raw));
  } else {
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) v[i] = (base + i < n) ? in[base + i] : zero_of<T>();
  }
  T s = zero_of<T>();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) s = s + v[i];
  T tot;
  T run = block_exclusive_scan<T>(s, s_wave, &tot) + offset;
  if (total && b == (int64_t)gridDim.x - 1 && threadIdx.x == 0) *total = offset + tot;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const T x = v[i];
    v[i] = run;
    run = run + x;
  }
  if (vec) {
    uint4 raw[kVec];
    __builtin_memcpy(raw, v, sizeof(raw));
#pragma unroll
    for (int k = 0; k < kVec; ++k) reinterpret_cast<uint4*>(out + base)[k] = raw[k];
  } else {
#pragma unroll
    for (int i = 0; i < kScanItems; ++i)
      if (base + i < n) out[base + i] = v[i];
  }
}

inline int64_t scan_blocks(int64_t n) { return (n + kScanTile - 1) / kScanTile; }
inline size_t scan_workspace_bytes(int64_t n, size_t elem) {
  return (size_t)(scan_blocks(n) + 1) * elem + 256;
}

// Exclusive scan of n elements of `in` into `out` (may alias), total into
// `total` (device, may be null).  ws holds the block partials.
template <typename T>
inline int exclusive_scan(const T* in, T* out, int64_t n, T* total, void* ws, hipStream_t st) {
  const int64_t nb = scan_blocks(n);
  T* partial = reinterpret_cast<T*>(ws);
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n, partial);
    GSMI_CHECK_LAUNCH();
  }
  const bool scanned = nb > kScanSumBlocks;
  if (scanned) {
    hipLaunchKernelGGL(k_scan_partials<T>, dim3(1), dim3(kScanThreads), 0, st, partial, nb);
    GSMI_CHECK_LAUNCH();
  }
  // n == 0: one block with no elements writes the zero total
  hipLaunchKernelGGL(k_scan_down<T>, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(kScanThreads), 0, st, in, n,
                     partial, scanned, out, total);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

}  // namespace gsmi
