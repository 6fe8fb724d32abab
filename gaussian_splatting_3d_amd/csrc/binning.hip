// Tile binning on gfx950: depth-first pair binning, stable LSD radix sorts,
// per-tile ranges.
//
// The reference emits one (tile << 32 | depth bits, index) pair per covered
// tile and sorts all 64 key bits (aabb_culling.h:15-41, 235-241).  The same
// order is produced here with far less traffic:
//   1. stable radix sort of the V visible Gaussians by depth bits (32-bit
//      keys, values = index), so ties keep index order;
//   2. pair counts gathered in that order, exclusive scan -> offsets;
//   3. load-balanced expansion: each output slot finds its Gaussian by binary
//      search over the offsets (in LDS) and writes (tile, index), Gaussians in
//      depth order and each Gaussian's tiles x-outer, y-inner;
//   4. stable radix sort of the pairs on the tile id alone
//      (ceil(log2 n_tiles) bits, two passes for <= 16384 tiles).
// Within a tile the entries are then ordered by (depth bits, index), i.e.
// exactly the order of a stable sort of the 64-bit keys emitted in index
// order.  The 64-bit keys are materialised only when a caller asks for them.
//
// Radix pass = hist (per-block digit counts, per-wave LDS counters) + scan
// over [digit][block] + scatter (stable in-block ranking with wave64 ballot
// matching on the digit bits, keys/values permuted through LDS so global
// writes are runs of consecutive addresses).
#include <algorithm>
#include <utility>

#include "scan.h"

namespace gsmi {

constexpr int kRadixThreads = 256;
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kRadixThreads * kRadixItems;  // 4096 keys per block

template <typename K, int RB>
__global__ __launch_bounds__(kRadixThreads) void k_radix_hist(const K* __restrict__ keys, int64_t P, int shift,
                                                              uint32_t* __restrict__ hist, int nblk) {
  constexpr int BINS = 1 << RB;
  __shared__ uint32_t s_cnt[4][BINS];
  const int t = threadIdx.x, w = t >> 6;
  for (int i = t; i < 4 * BINS; i += kRadixThreads) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
#pragma unroll 4
  for (int i = 0; i < kRadixItems; ++i) {
    const int64_t p = base + t + (int64_t)i * kRadixThreads;
    if (p < P) {
      const uint32_t d = (uint32_t)(keys[p] >> shift) & (BINS - 1);
      atomicAdd(&s_cnt[w][d], 1u);
    }
  }
  __syncthreads();
  if (t < BINS) hist[(int64_t)t * nblk + blockIdx.x] = s_cnt[0][t] + s_cnt[1][t] + s_cnt[2][t] + s_cnt[3][t];
}

template <typename K, int RB>
__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(const K* __restrict__ kin,
                                                                 const int32_t* __restrict__ vin,
                                                                 K* __restrict__ kout, int32_t* __restrict__ vout,
                                                                 int64_t P, int shift,
                                                                 const uint32_t* __restrict__ offs, int nblk) {
  constexpr int BINS = 1 << RB;
  static_assert(BINS <= kRadixThreads, "one digit per thread in the block prefix");
  __shared__ uint32_t s_cnt[4][BINS];
  __shared__ uint32_t s_base[4][BINS];
  __shared__ uint32_t s_dstart[BINS];
  __shared__ uint32_t s_goff[BINS];
  __shared__ uint32_t s_wsum[4];
  __shared__ K s_keys[kRadixTile];
  __shared__ int32_t s_vals[kRadixTile];

  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  for (int i = lane; i < BINS; i += 64) s_cnt[w][i] = 0;
  if (t < BINS) s_goff[t] = offs[(int64_t)t * nblk + blockIdx.x];
  __syncthreads();

  const uint64_t lt_mask = (1ull << lane) - 1ull;
  K key[kRadixItems];
  int32_t val[kRadixItems];
  uint32_t rank[kRadixItems];
  // Wave w ranks items [w*1024, (w+1)*1024) of the block, 64 per round, in
  // position order, so ranks are stable.
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const int64_t p = base + w * (64 * kRadixItems) + r * 64 + lane;
    const bool valid = p < P;
    key[r] = valid ? kin[p] : (K)~(K)0;
    val[r] = valid ? vin[p] : 0;
    const uint32_t d = (uint32_t)(key[r] >> shift) & (BINS - 1);
    uint64_t peers = ~0ull;
#pragma unroll
    for (int b = 0; b < RB; ++b) {
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
  uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  if (t < BINS) {
    c0 = s_cnt[0][t];
    c1 = s_cnt[1][t];
    c2 = s_cnt[2][t];
    c3 = s_cnt[3][t];
  }
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
  if (t < BINS) {
    const uint32_t dstart = pre + inc - tot;
    s_dstart[t] = dstart;
    s_base[0][t] = dstart;
    s_base[1][t] = dstart + c0;
    s_base[2][t] = dstart + c0 + c1;
    s_base[3][t] = dstart + c0 + c1 + c2;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const uint32_t d = (uint32_t)(key[r] >> shift) & (BINS - 1);
    const uint32_t lp = s_base[w][d] + rank[r];
    s_keys[lp] = key[r];
    s_vals[lp] = val[r];
  }
  __syncthreads();
  const int64_t nvalid = min((int64_t)kRadixTile, P - base);
#pragma unroll 4
  for (int i = t; i < kRadixTile; i += kRadixThreads) {
    if (i < nvalid) {
      const K k = s_keys[i];
      const uint32_t d = (uint32_t)(k >> shift) & (BINS - 1);
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
// Tile of entry i = key[i] >> SHIFT (32 for 64-bit keys, 0 for tile keys).
template <typename K, int SHIFT>
__global__ void k_tile_ranges(const K* __restrict__ keys, int64_t P, int32_t* __restrict__ start,
                              int32_t* __restrict__ end, uint32_t n_tiles) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const uint32_t tile = (uint32_t)(keys[i] >> SHIFT);
  if (tile >= n_tiles) return;
  if (i == 0 || (uint32_t)(keys[i - 1] >> SHIFT) != tile) start[tile] = (int32_t)i;
  if (i == P - 1 || (uint32_t)(keys[i + 1] >> SHIFT) != tile) end[tile] = (int32_t)(i + 1);
}

// Depth keys and identity values for the V visible Gaussians.
__global__ void k_depth_keys(const float* __restrict__ depth, int64_t V, uint32_t* __restrict__ dkey,
                             int32_t* __restrict__ dval) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  dkey[v] = __float_as_uint(depth[v]);
  dval[v] = (int32_t)v;
}

// Pair count of the Gaussian at depth rank r (gs/culling.py:33-35).
__global__ void k_sorted_counts(const int32_t* __restrict__ dval, const int32_t* __restrict__ tl,
                                const int32_t* __restrict__ br, int64_t V, int32_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V) return;
  const int v = dval[r];
  const int2 a = reinterpret_cast<const int2*>(tl)[v];
  const int2 b = reinterpret_cast<const int2*>(br)[v];
  cnt[r] = (b.x - a.x + 1) * (b.y - a.y + 1);
}

// Load-balanced expansion of the depth-ordered Gaussians into (tile, index)
// pairs.  Block b owns output slots [b*kExpandSlots, (b+1)*kExpandSlots); the
// Gaussians overlapping that range are staged in LDS (every visible Gaussian
// has at least one pair, so at most kExpandSlots + 1 of them), and each thread
// locates the Gaussian of its slots by binary search.
constexpr int kExpandSlots = 2048;
constexpr int kExpandPerThread = kExpandSlots / 256;

__global__ __launch_bounds__(256) void k_expand(const int32_t* __restrict__ off, const int32_t* __restrict__ dval,
                                                const int32_t* __restrict__ tl, const int32_t* __restrict__ br,
                                                int64_t V, int64_t P, int n_tiles_w, uint32_t* __restrict__ tile_out,
                                                int32_t* __restrict__ val_out) {
  __shared__ int32_t s_off[kExpandSlots + 2];
  __shared__ int32_t s_v[kExpandSlots + 2];
  __shared__ int32_t s_range[2];
  const int t = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * kExpandSlots;
  const int64_t p1 = min(p0 + kExpandSlots, P);  // exclusive
  if (t < 2) {
    // largest r with off[r] <= target
    const int64_t target = t == 0 ? p0 : p1 - 1;
    int64_t lo = 0, hi = V - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if ((int64_t)off[mid] <= target) lo = mid;
      else hi = mid - 1;
    }
    s_range[t] = (int32_t)lo;
  }
  __syncthreads();
  const int r0 = s_range[0];
  // <= kExpandSlots + 1 when every Gaussian has at least one pair; clamped so
  // inconsistent bounds cannot index past the LDS arrays.
  const int n = max(1, min(s_range[1] - r0 + 1, kExpandSlots + 1));
  for (int i = t; i < n; i += 256) {
    s_off[i] = off[r0 + i];
    s_v[i] = dval[r0 + i];
  }
  __syncthreads();
  int64_t p = p0 + (int64_t)t * kExpandPerThread;
  if (p >= p1 || V <= 0) return;
  // binary search for the first slot, then walk forward
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int64_t)s_off[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  int j = lo;
  int v = s_v[j];
  int2 a = reinterpret_cast<const int2*>(tl)[v];
  int2 b = reinterpret_cast<const int2*>(br)[v];
#pragma unroll
  for (int k = 0; k < kExpandPerThread; ++k, ++p) {
    if (p >= p1) break;
    while (j + 1 < n && (int64_t)s_off[j + 1] <= p) {
      ++j;
      v = s_v[j];
      a = reinterpret_cast<const int2*>(tl)[v];
      b = reinterpret_cast<const int2*>(br)[v];
    }
    const int local = (int)(p - s_off[j]);
    const int ny = b.y - a.y + 1;
    const int x = a.x + local / ny;
    const int y = a.y + local - (local / ny) * ny;
    tile_out[p] = (uint32_t)(y * n_tiles_w + x);
    val_out[p] = v;
  }
}

// 64-bit keys of the sorted pairs: (tile << 32) | depth bits of the Gaussian.
__global__ void k_make_keys(const uint32_t* __restrict__ tile, const int32_t* __restrict__ ids,
                            const float* __restrict__ depth, int64_t P, uint64_t* __restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  keys[i] = ((uint64_t)tile[i] << 32) | (uint64_t)__float_as_uint(depth[ids[i]]);
}

static int bits_for(int64_t n) {
  int b = 0;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

static size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Radix sort of P (K, int32) pairs over key bits [0, end_bit) in passes of
// at most RB bits, ping-ponging with (kalt, valt); the result lands in
// (keys, values).
template <typename K, int RB>
static size_t radix_ws_bytes(int64_t P) {
  const int64_t nblk = (P + kRadixTile - 1) / kRadixTile;
  const int64_t nh = nblk * (1 << RB);
  return align_up(P * sizeof(K)) + align_up(P * 4) + align_up(nh * 4) + align_up(scan_workspace_bytes(nh, 4));
}

template <typename K, int RB>
static int radix_sort(int64_t P, K* keys, int32_t* values, int end_bit, void* ws, hipStream_t st) {
  if (P <= 1 || end_bit <= 0) return GSMI_OK;
  const int nblk = (int)((P + kRadixTile - 1) / kRadixTile);
  const int64_t nh = (int64_t)nblk * (1 << RB);
  char* w = reinterpret_cast<char*>(ws);
  K* kalt = reinterpret_cast<K*>(w);
  w += align_up(P * sizeof(K));
  int32_t* valt = reinterpret_cast<int32_t*>(w);
  w += align_up(P * 4);
  uint32_t* hist = reinterpret_cast<uint32_t*>(w);
  w += align_up(nh * 4);
  void* scan_ws = w;
  const int npass = (end_bit + RB - 1) / RB;
  K* kin = keys;
  int32_t* vin = values;
  K* kout = kalt;
  int32_t* vout = valt;
  for (int pass = 0; pass < npass; ++pass) {
    const int shift = RB * pass;
    hipLaunchKernelGGL((k_radix_hist<K, RB>), dim3(nblk), dim3(kRadixThreads), 0, st, kin, P, shift, hist, nblk);
    GSMI_CHECK_LAUNCH();
    int rc = exclusive_scan<int32_t>(reinterpret_cast<int32_t*>(hist), reinterpret_cast<int32_t*>(hist), nh,
                                     nullptr, scan_ws, st);
    if (rc) return rc;
    hipLaunchKernelGGL((k_radix_scatter<K, RB>), dim3(nblk), dim3(kRadixThreads), 0, st, kin, vin, kout, vout, P,
                       shift, hist, nblk);
    GSMI_CHECK_LAUNCH();
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  if (kin != keys) {
    if (hipMemcpyAsync(keys, kin, P * sizeof(K), hipMemcpyDeviceToDevice, st) != hipSuccess) return GSMI_ERR_LAUNCH;
    if (hipMemcpyAsync(values, vin, P * 4, hipMemcpyDeviceToDevice, st) != hipSuccess) return GSMI_ERR_LAUNCH;
  }
  return GSMI_OK;
}

template <typename K, int SHIFT>
static int ranges_impl(int64_t P, const K* sorted_keys, int n_tiles, int32_t* start, int32_t* end, hipStream_t st) {
  const int nb = (n_tiles + 255) / 256;
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, st, start, (int64_t)n_tiles, -1);
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, st, end, (int64_t)n_tiles, -1);
  GSMI_CHECK_LAUNCH();
  if (P > 0) {
    hipLaunchKernelGGL((k_tile_ranges<K, SHIFT>), dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, sorted_keys,
                       P, start, end, (uint32_t)n_tiles);
    GSMI_CHECK_LAUNCH();
  }
  return GSMI_OK;
}

// Workspace layout of the depth-first binning.
struct BinWs {
  uint32_t* dkey;
  int32_t* dval;
  int32_t* cnt;
  int32_t* total;
  uint32_t* tkey;
  void* scan_ws;
  void* sort_ws;
};

static size_t bin_ws_bytes(int64_t V, int64_t P) {
  return align_up(V * 4) * 3 + 256 + align_up(P * 4) + align_up(scan_workspace_bytes(V, 4)) +
         std::max(radix_ws_bytes<uint32_t, 8>(V), radix_ws_bytes<uint32_t, 8>(P));
}

static BinWs carve_bin_ws(void* ws, int64_t V, int64_t P) {
  char* p = reinterpret_cast<char*>(ws);
  BinWs b;
  b.dkey = reinterpret_cast<uint32_t*>(p);
  p += align_up(V * 4);
  b.dval = reinterpret_cast<int32_t*>(p);
  p += align_up(V * 4);
  b.cnt = reinterpret_cast<int32_t*>(p);
  p += align_up(V * 4);
  b.total = reinterpret_cast<int32_t*>(p);
  p += 256;
  b.tkey = reinterpret_cast<uint32_t*>(p);
  p += align_up(P * 4);
  b.scan_ws = p;
  p += align_up(scan_workspace_bytes(V, 4));
  b.sort_ws = p;
  return b;
}

}  // namespace gsmi

using namespace gsmi;

extern "C" size_t gsmi_sort_workspace_size(int64_t P) { return radix_ws_bytes<uint64_t, 8>(P < 0 ? 0 : P); }

extern "C" int gsmi_sort_pairs(int64_t P, uint64_t* keys, int32_t* values, int end_bit, void* workspace,
                               size_t workspace_bytes, void* stream) {
  if (P < 0 || end_bit < 0 || end_bit > 64) return GSMI_ERR_ARG;
  if (P > 0 && (!keys || !values)) return GSMI_ERR_ARG;
  if (P >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  if (P > 1 && (!workspace || workspace_bytes < radix_ws_bytes<uint64_t, 8>(P))) return GSMI_ERR_WORKSPACE;
  // 8-bit digits at fixed 8-bit steps
  const int eb = ((end_bit + 7) / 8) * 8;
  return radix_sort<uint64_t, 8>(P, keys, values, eb, workspace, (hipStream_t)stream);
}

extern "C" int gsmi_tile_ranges(int64_t P, const uint64_t* sorted_keys, int n_tiles, int32_t* start, int32_t* end,
                                void* stream) {
  if (P < 0 || n_tiles <= 0 || !start || !end || (P > 0 && !sorted_keys)) return GSMI_ERR_ARG;
  return ranges_impl<uint64_t, 32>(P, sorted_keys, n_tiles, start, end, (hipStream_t)stream);
}

extern "C" size_t gsmi_binning_workspace_size(int64_t V, int64_t P, int n_tiles) {
  (void)n_tiles;
  return bin_ws_bytes(V < 0 ? 0 : V, P < 0 ? 0 : P);
}

extern "C" int gsmi_tile_culling_aabb_start_end(int64_t V, int64_t P, const int32_t* tl, const int32_t* br,
                                                const float* depth, int n_tiles_h, int n_tiles_w,
                                                int32_t* gaussian_ids, int32_t* start, int32_t* end,
                                                uint64_t* sorted_keys, void* workspace, size_t workspace_bytes,
                                                int flags, void* stream) {
  if (V < 0 || P < 0 || n_tiles_h <= 0 || n_tiles_w <= 0 || !start || !end) return GSMI_ERR_ARG;
  if (V > 0 && (!tl || !br || !depth)) return GSMI_ERR_ARG;
  if (((uintptr_t)tl & 7) || ((uintptr_t)br & 7)) return GSMI_ERR_ARG;
  if (P > 0 && !gaussian_ids) return GSMI_ERR_ARG;
  if (P >= (int64_t(1) << 31) || V >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  const int n_tiles = n_tiles_h * n_tiles_w;
  if (n_tiles > (1 << 24)) return GSMI_ERR_UNSUPPORTED;
  if (!workspace || workspace_bytes < bin_ws_bytes(V, P)) return GSMI_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  BinWs w = carve_bin_ws(workspace, V, P);
  int rc;
  if (V > 0) {
    const unsigned g = (unsigned)((V + 255) / 256);
    hipLaunchKernelGGL(k_depth_keys, dim3(g), dim3(256), 0, st, depth, V, w.dkey, w.dval);
    GSMI_CHECK_LAUNCH();
    rc = radix_sort<uint32_t, 8>(V, w.dkey, w.dval, 32, w.sort_ws, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sorted_counts, dim3(g), dim3(256), 0, st, w.dval, tl, br, V, w.cnt);
    GSMI_CHECK_LAUNCH();
  }
  rc = exclusive_scan<int32_t>(w.cnt, w.cnt, V, w.total, w.scan_ws, st);
  if (rc) return rc;
  if (flags & 1) {
    // The caller sized gaussian_ids with its own count; check it before any
    // pair is written (the reference asserts the same, aabb_culling.h:226-228).
    int32_t h_total = 0;
    if (hipMemcpyAsync(&h_total, w.total, 4, hipMemcpyDeviceToHost, st) != hipSuccess) return GSMI_ERR_LAUNCH;
    if (hipStreamSynchronize(st) != hipSuccess) return GSMI_ERR_LAUNCH;
    if ((int64_t)h_total != P) return GSMI_ERR_COUNT;
  }
  if (P > 0 && V > 0) {
    hipLaunchKernelGGL(k_expand, dim3((unsigned)((P + kExpandSlots - 1) / kExpandSlots)), dim3(256), 0, st, w.cnt,
                       w.dval, tl, br, V, P, n_tiles_w, w.tkey, gaussian_ids);
    GSMI_CHECK_LAUNCH();
    rc = radix_sort<uint32_t, 8>(P, w.tkey, gaussian_ids, bits_for(n_tiles), w.sort_ws, st);
    if (rc) return rc;
  }
  rc = ranges_impl<uint32_t, 0>(P, w.tkey, n_tiles, start, end, st);
  if (rc) return rc;
  if (sorted_keys && P > 0) {
    hipLaunchKernelGGL(k_make_keys, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, w.tkey, gaussian_ids, depth,
                       P, sorted_keys);
    GSMI_CHECK_LAUNCH();
  }
  return GSMI_OK;
}
