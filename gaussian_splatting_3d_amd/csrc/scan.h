// Device-wide exclusive prefix sums (reduce, scan-of-partials, down-sweep).
// 256 threads x 16 items = 4096 elements per block.  Element type is int32
// or Int2 (two independent int32 sums scanned together).
#pragma once

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

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T* __restrict__ in, int64_t n,
                                                              T* __restrict__ partial) {
  __shared__ T s_wave[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  T s = zero_of<T>();
  for (int i = threadIdx.x; i < kScanTile; i += kScanThreads) {
    const int64_t p = base + i;
    if (p < n) s = s + in[p];
  }
  T tot;
  block_exclusive_scan<T>(s, s_wave, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single block: exclusive scan of the partials in place, total to *total.
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(T* __restrict__ partial, int64_t nb,
                                                                T* __restrict__ total) {
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
  if (threadIdx.x == 0 && total) *total = carry;
}

// Each thread owns kScanItems consecutive elements of the block's tile.
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const T* __restrict__ in, int64_t n,
                                                            const T* __restrict__ partial, T* __restrict__ out) {
  __shared__ T s_wave[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = zero_of<T>();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : zero_of<T>();
    s = s + v[i];
  }
  T tot;
  T run = block_exclusive_scan<T>(s, s_wave, &tot) + partial[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run = run + v[i];
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
  hipLaunchKernelGGL(k_scan_partials<T>, dim3(1), dim3(kScanThreads), 0, st, partial, nb, total);
  GSMI_CHECK_LAUNCH();
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_down<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n, partial, out);
    GSMI_CHECK_LAUNCH();
  }
  return GSMI_OK;
}

}  // namespace gsmi
