// Tile binning on gfx950: pair emission, stable LSD radix sort of 64-bit
// (tile << 32 | depth bits) keys, per-tile ranges.
//
// Radix sort: 8-bit digits over only the key bits that can be non-zero
// (32 depth bits + ceil(log2 n_tiles) tile bits), three kernels per pass:
//   hist    — per-block digit histogram (4096 keys per block, per-wave LDS
//             counters), written digit-major;
//   scan    — exclusive prefix over [digit][block];
//   scatter — stable in-block ranking (wave64 ballot match on the 8 digit
//             bits, per-wave running counters, block prefix over digits),
//             keys/values permuted through LDS so that global writes are runs
//             of consecutive addresses.
// Stability plus index-order emission makes the result a deterministic
// function of the input (the reference's atomic-slot emission,
// aabb_culling.h:32, is not).
#include "scan.h"

namespace gsmi {

constexpr int kRadixThreads = 256;
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kRadixThreads * kRadixItems;  // 4096 keys per block
constexpr int kRadixBins = 256;

__global__ __launch_bounds__(kRadixThreads) void k_radix_hist(const uint64_t* __restrict__ keys, int64_t P,
                                                              int shift, uint32_t* __restrict__ hist, int nblk) {
  __shared__ uint32_t s_cnt[4][kRadixBins];
  const int t = threadIdx.x, w = t >> 6;
#pragma unroll
  for (int i = 0; i < 4; ++i) s_cnt[i][t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
#pragma unroll 4
  for (int i = 0; i < kRadixItems; ++i) {
    const int64_t p = base + t + (int64_t)i * kRadixThreads;
    if (p < P) {
      const uint32_t d = (uint32_t)(keys[p] >> shift) & 255u;
      atomicAdd(&s_cnt[w][d], 1u);
    }
  }
  __syncthreads();
  hist[(int64_t)t * nblk + blockIdx.x] = s_cnt[0][t] + s_cnt[1][t] + s_cnt[2][t] + s_cnt[3][t];
}

__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(const uint64_t* __restrict__ kin,
                                                                 const int32_t* __restrict__ vin,
                                                                 uint64_t* __restrict__ kout,
                                                                 int32_t* __restrict__ vout, int64_t P, int shift,
                                                                 const uint32_t* __restrict__ offs, int nblk) {
  __shared__ uint32_t s_cnt[4][kRadixBins];
  __shared__ uint32_t s_base[4][kRadixBins];
  __shared__ uint32_t s_dstart[kRadixBins];
  __shared__ uint32_t s_goff[kRadixBins];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint64_t s_keys[kRadixTile];
  __shared__ int32_t s_vals[kRadixTile];

  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
#pragma unroll
  for (int i = 0; i < 4; ++i) s_cnt[w][lane + 64 * i] = 0;
  s_goff[t] = offs[(int64_t)t * nblk + blockIdx.x];
  __syncthreads();

  const uint64_t lt_mask = (1ull << lane) - 1ull;
  uint64_t key[kRadixItems];
  int32_t val[kRadixItems];
  uint32_t rank[kRadixItems];
  // Wave w ranks items [w*1024, (w+1)*1024) of the block, 64 per round, in
  // position order, so ranks are stable.
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const int64_t p = base + w * (64 * kRadixItems) + r * 64 + lane;
    const bool valid = p < P;
    key[r] = valid ? kin[p] : ~0ull;
    val[r] = valid ? vin[p] : 0;
    const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
    uint64_t peers = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t lower = __popcll(peers & lt_mask);
    const uint32_t old = s_cnt[w][d];
    __builtin_amdgcn_wave_barrier();
    if (lower == 0) s_cnt[w][d] = old + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rank[r] = old + lower;
  }
  __syncthreads();
  // Digit t: per-wave counts -> wave offsets; block prefix over digits.
  const uint32_t c0 = s_cnt[0][t], c1 = s_cnt[1][t], c2 = s_cnt[2][t], c3 = s_cnt[3][t];
  const uint32_t tot = c0 + c1 + c2 + c3;
  uint32_t inc = tot;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_wsum[w] = inc;
  __syncthreads();
  uint32_t pre = 0;
  for (int i = 0; i < w; ++i) pre += s_wsum[i];
  const uint32_t dstart = pre + inc - tot;
  s_dstart[t] = dstart;
  s_base[0][t] = dstart;
  s_base[1][t] = dstart + c0;
  s_base[2][t] = dstart + c0 + c1;
  s_base[3][t] = dstart + c0 + c1 + c2;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
    const uint32_t lp = s_base[w][d] + rank[r];
    s_keys[lp] = key[r];
    s_vals[lp] = val[r];
  }
  __syncthreads();
  const int64_t nvalid = min((int64_t)kRadixTile, P - base);
#pragma unroll 4
  for (int i = t; i < kRadixTile; i += kRadixThreads) {
    if (i < nvalid) {
      const uint64_t k = s_keys[i];
      const uint32_t d = (uint32_t)(k >> shift) & 255u;
      const int64_t gp = (int64_t)s_goff[d] + (i - (int64_t)s_dstart[d]);
      kout[gp] = k;
      vout[gp] = s_vals[i];
    }
  }
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// aabb_culling.h:70-103: start = first index of a tile's run, end = last+1.
__global__ void k_tile_ranges(const uint64_t* __restrict__ keys, int64_t P, int32_t* __restrict__ start,
                              int32_t* __restrict__ end, uint32_t n_tiles) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const uint32_t tile = (uint32_t)(keys[i] >> 32);
  if (tile >= n_tiles) return;
  if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != tile) start[tile] = (int32_t)i;
  if (i == P - 1 || (uint32_t)(keys[i + 1] >> 32) != tile) end[tile] = (int32_t)(i + 1);
}

// Pair count per Gaussian from its tile bounds (gs/culling.py:33-35).
__global__ void k_count_pairs(const int32_t* __restrict__ tl, const int32_t* __restrict__ br, int64_t V,
                              int32_t* __restrict__ cnt) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int2 a = reinterpret_cast<const int2*>(tl)[v];
  const int2 b = reinterpret_cast<const int2*>(br)[v];
  cnt[v] = (b.x - a.x + 1) * (b.y - a.y + 1);
}

// aabb_culling.h:15-41 emission, x outer, y inner, at precomputed offsets.
__device__ __forceinline__ void emit_pairs(int tlx, int tly, int brx, int bry, uint32_t dbits, int32_t value,
                                           int64_t off, int64_t P, int n_tiles_w, uint64_t* keys, int32_t* values) {
  for (int x = tlx; x <= brx; ++x) {
    for (int y = tly; y <= bry; ++y) {
      if (off < 0 || off >= P) return;
      const uint64_t tile = (uint64_t)(uint32_t)(y * n_tiles_w + x);
      keys[off] = (tile << 32) | (uint64_t)dbits;
      values[off] = value;
      ++off;
    }
  }
}

__global__ void k_emit_pairs(const int32_t* __restrict__ tl, const int32_t* __restrict__ br,
                             const float* __restrict__ depth, const int32_t* __restrict__ off, int64_t V, int64_t P,
                             int n_tiles_w, uint64_t* __restrict__ keys, int32_t* __restrict__ values) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int2 a = reinterpret_cast<const int2*>(tl)[v];
  const int2 b = reinterpret_cast<const int2*>(br)[v];
  emit_pairs(a.x, a.y, b.x, b.y, __float_as_uint(depth[v]), (int32_t)v, off[v], P, n_tiles_w, keys, values);
}

static int bits_for(int64_t n) {
  int b = 0;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

struct SortWs {
  uint64_t* keys_alt;
  int32_t* vals_alt;
  uint32_t* hist;
  void* scan_ws;
};

static size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

static size_t sort_ws_bytes(int64_t P) {
  const int64_t nblk = (P + kRadixTile - 1) / kRadixTile;
  const int64_t nh = nblk * kRadixBins;
  return align_up(P * 8) + align_up(P * 4) + align_up(nh * 4) + align_up(scan_workspace_bytes(nh, 4));
}

static SortWs carve_sort_ws(void* ws, int64_t P) {
  const int64_t nblk = (P + kRadixTile - 1) / kRadixTile;
  const int64_t nh = nblk * kRadixBins;
  char* p = reinterpret_cast<char*>(ws);
  SortWs s;
  s.keys_alt = reinterpret_cast<uint64_t*>(p);
  p += align_up(P * 8);
  s.vals_alt = reinterpret_cast<int32_t*>(p);
  p += align_up(P * 4);
  s.hist = reinterpret_cast<uint32_t*>(p);
  p += align_up(nh * 4);
  s.scan_ws = p;
  return s;
}

static int sort_pairs_impl(int64_t P, uint64_t* keys, int32_t* values, int end_bit, void* ws, hipStream_t st) {
  if (P <= 1) return GSMI_OK;
  const int nblk = (int)((P + kRadixTile - 1) / kRadixTile);
  SortWs s = carve_sort_ws(ws, P);
  uint64_t* kin = keys;
  int32_t* vin = values;
  uint64_t* kout = s.keys_alt;
  int32_t* vout = s.vals_alt;
  const int npass = (end_bit + 7) / 8;
  for (int pass = 0; pass < npass; ++pass) {
    const int shift = 8 * pass;
    hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(kRadixThreads), 0, st, kin, P, shift, s.hist, nblk);
    GSMI_CHECK_LAUNCH();
    int rc = exclusive_scan<int32_t>(reinterpret_cast<int32_t*>(s.hist), reinterpret_cast<int32_t*>(s.hist),
                                     (int64_t)nblk * kRadixBins, nullptr, s.scan_ws, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(kRadixThreads), 0, st, kin, vin, kout, vout, P, shift,
                       s.hist, nblk);
    GSMI_CHECK_LAUNCH();
    uint64_t* tk = kin;
    kin = kout;
    kout = tk;
    int32_t* tv = vin;
    vin = vout;
    vout = tv;
  }
  if (kin != keys) {
    if (hipMemcpyAsync(keys, kin, P * 8, hipMemcpyDeviceToDevice, st) != hipSuccess) return GSMI_ERR_LAUNCH;
    if (hipMemcpyAsync(values, vin, P * 4, hipMemcpyDeviceToDevice, st) != hipSuccess) return GSMI_ERR_LAUNCH;
  }
  return GSMI_OK;
}

static int ranges_impl(int64_t P, const uint64_t* sorted_keys, int n_tiles, int32_t* start, int32_t* end,
                       hipStream_t st) {
  const int nb = (n_tiles + 255) / 256;
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, st, start, (int64_t)n_tiles, -1);
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, st, end, (int64_t)n_tiles, -1);
  GSMI_CHECK_LAUNCH();
  if (P > 0) {
    hipLaunchKernelGGL(k_tile_ranges, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, sorted_keys, P, start,
                       end, (uint32_t)n_tiles);
    GSMI_CHECK_LAUNCH();
  }
  return GSMI_OK;
}

}  // namespace gsmi

using namespace gsmi;

extern "C" size_t gsmi_sort_workspace_size(int64_t P) { return sort_ws_bytes(P < 0 ? 0 : P); }

extern "C" int gsmi_sort_pairs(int64_t P, uint64_t* keys, int32_t* values, int end_bit, void* workspace,
                               size_t workspace_bytes, void* stream) {
  if (P < 0 || end_bit < 0 || end_bit > 64) return GSMI_ERR_ARG;
  if (P > 0 && (!keys || !values)) return GSMI_ERR_ARG;
  if (P >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  if (P > 1 && (!workspace || workspace_bytes < sort_ws_bytes(P))) return GSMI_ERR_WORKSPACE;
  return sort_pairs_impl(P, keys, values, end_bit, workspace, (hipStream_t)stream);
}

extern "C" int gsmi_tile_ranges(int64_t P, const uint64_t* sorted_keys, int n_tiles, int32_t* start, int32_t* end,
                                void* stream) {
  if (P < 0 || n_tiles <= 0 || !start || !end || (P > 0 && !sorted_keys)) return GSMI_ERR_ARG;
  return ranges_impl(P, sorted_keys, n_tiles, start, end, (hipStream_t)stream);
}

// Workspace: counts/offsets [V] + scan partials + total + keys [P] + sort ws.
extern "C" size_t gsmi_binning_workspace_size(int64_t V, int64_t P, int n_tiles) {
  (void)n_tiles;
  if (V < 0) V = 0;
  if (P < 0) P = 0;
  return align_up(V * 4) + align_up(scan_workspace_bytes(V, 4)) + 256 + align_up(P * 8) + sort_ws_bytes(P);
}

extern "C" int gsmi_tile_culling_aabb_start_end(int64_t V, int64_t P, const int32_t* tl, const int32_t* br,
                                                const float* depth, int n_tiles_h, int n_tiles_w,
                                                int32_t* gaussian_ids, int32_t* start, int32_t* end,
                                                uint64_t* sorted_keys, void* workspace, size_t workspace_bytes,
                                                void* stream) {
  if (V < 0 || P < 0 || n_tiles_h <= 0 || n_tiles_w <= 0 || !start || !end) return GSMI_ERR_ARG;
  if (V > 0 && (!tl || !br || !depth)) return GSMI_ERR_ARG;
  if (((uintptr_t)tl & 7) || ((uintptr_t)br & 7)) return GSMI_ERR_ARG;
  if (P > 0 && !gaussian_ids) return GSMI_ERR_ARG;
  if (P >= (int64_t(1) << 31) || V >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  const int n_tiles = n_tiles_h * n_tiles_w;
  if (!workspace || workspace_bytes < gsmi_binning_workspace_size(V, P, n_tiles)) return GSMI_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* p = reinterpret_cast<char*>(workspace);
  int32_t* off = reinterpret_cast<int32_t*>(p);
  p += align_up(V * 4);
  void* scan_ws = p;
  p += align_up(scan_workspace_bytes(V, 4));
  int32_t* d_total = reinterpret_cast<int32_t*>(p);
  p += 256;
  uint64_t* keys = reinterpret_cast<uint64_t*>(p);
  p += align_up(P * 8);
  void* sort_ws = p;

  if (V > 0) {
    hipLaunchKernelGGL(k_count_pairs, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, st, tl, br, V, off);
    GSMI_CHECK_LAUNCH();
  }
  int rc = exclusive_scan<int32_t>(off, off, V, d_total, scan_ws, st);
  if (rc) return rc;
  // The caller sized gaussian_ids with its own count; check it before any
  // pair is written (the reference asserts the same, aabb_culling.h:226-228).
  int32_t h_total = 0;
  if (hipMemcpyAsync(&h_total, d_total, 4, hipMemcpyDeviceToHost, st) != hipSuccess) return GSMI_ERR_LAUNCH;
  if (hipStreamSynchronize(st) != hipSuccess) return GSMI_ERR_LAUNCH;
  if ((int64_t)h_total != P) return GSMI_ERR_COUNT;
  if (V > 0) {
    hipLaunchKernelGGL(k_emit_pairs, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, st, tl, br, depth, off, V, P,
                       n_tiles_w, keys, gaussian_ids);
    GSMI_CHECK_LAUNCH();
  }
  const int end_bit = 32 + bits_for(n_tiles);
  rc = sort_pairs_impl(P, keys, gaussian_ids, end_bit, sort_ws, st);
  if (rc) return rc;
  rc = ranges_impl(P, keys, n_tiles, start, end, st);
  if (rc) return rc;
  if (sorted_keys && P > 0) {
    if (hipMemcpyAsync(sorted_keys, keys, P * 8, hipMemcpyDeviceToDevice, st) != hipSuccess) return GSMI_ERR_LAUNCH;
  }
  return GSMI_OK;
}
