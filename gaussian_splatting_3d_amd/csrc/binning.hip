// Tile binning on gfx950: depth-first pair binning, stable LSD radix sorts,
// per-tile ranges.
//
// The reference emits one (tile << 32 | depth bits, index) pair per covered
// tile and sorts all 64 key bits (aabb_culling.h:15-41, 235-241).  The same
// order is produced here with far less traffic:
//   1. stable radix sort of the V visible Gaussians by depth bits (32-bit
//      keys, values = index), so ties keep index order;
//   2. pair counts and tile rectangles gathered in that order, exclusive
//      scan -> offsets;
//   3. the pairs are never written in depth order: the expansion (each output
//      slot -> its Gaussian and tile, Gaussians in depth order, each
//      Gaussian's tiles x-outer, y-inner) runs twice, once to count the low
//      7 tile-id bits per 2048-slot block and once to rank the block's pairs
//      on those bits and scatter them as the first radix pass;
//   4. the remaining tile-id bits in passes of at most 8; the last pass also
//      writes each tile's first and last position from the digit offsets at
//      the boundaries of the low-bit buckets (block_ranges), so no separate
//      range pass reads the keys.
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
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kRadixThreads * kRadixItems;  // 2048 keys per block
// Radix and expansion kernels map blockIdx to data blocks with xcd_tile, so
// each XCD works on a contiguous range: the [digit][block] histogram entries
// and the digit runs that neighbouring blocks write are then adjacent in one
// L2 (measured at 58.7M pairs: final scatter 284 -> 171 us).

template <typename K, int RB>
__global__ __launch_bounds__(kRadixThreads) void k_radix_hist(const K* __restrict__ keys, int64_t P, int shift,
                                                              uint32_t* __restrict__ hist, int nblk) {
  constexpr int BINS = 1 << RB;
  __shared__ uint32_t s_cnt[4][BINS];
  const int blk = xcd_tile(blockIdx.x, nblk);
  if (blk >= nblk) return;
  const int t = threadIdx.x, w = t >> 6;
  for (int i = t; i < 4 * BINS; i += kRadixThreads) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blk * kRadixTile;
  // all loads in flight before the first counter update
  K k[kRadixItems];
#pragma unroll
  for (int i = 0; i < kRadixItems; ++i) {
    const int64_t p = base + t + (int64_t)i * kRadixThreads;
    k[i] = p < P ? keys[p] : (K)0;
  }
#pragma unroll
  for (int i = 0; i < kRadixItems; ++i)
    if (base + t + (int64_t)i * kRadixThreads < P) atomicAdd(&s_cnt[w][(uint32_t)(k[i] >> shift) & (BINS - 1)], 1u);
  __syncthreads();
  if (t < BINS) hist[(int64_t)t * nblk + blk] = s_cnt[0][t] + s_cnt[1][t] + s_cnt[2][t] + s_cnt[3][t];
}

template <typename K, int RB, int TILE>
struct ScatterLds {
  uint32_t cnt[4][1 << RB];
  uint32_t base[4][1 << RB];
  uint32_t dstart[1 << RB];
  uint32_t goff[1 << RB];
  uint32_t pref[1 << RB];  // ranges: digit counts before a bucket boundary
  uint32_t wsum[4];
  K wlast[4];              // ranges: last key of each wave's items
  K kfirst, klast;         // ranges: keys at block positions 0 and nvalid - 1
  int nextb;
  uint32_t lnew, lold;
  K keys[TILE];
  alignas(16) int32_t vals[TILE];
};

constexpr uint32_t kNoBucket = 0xffffffffu;

// Tile ranges of the last pass without atomics.  Before the last pass (shift
// = its first bit) the pairs are sorted by the low tile-id bits, so a block's
// "buckets" (runs of equal low bits) start and end at known positions, and
// for digit h the number of pairs with tile < (h, l) is
//   X(h) = offs[h][block] + (digit-h pairs of the block before the position
//          where bucket l starts),
// since earlier blocks and earlier positions hold only lower buckets.  So at
// every bucket boundary (low bits change between positions pb - 1 and pb;
// also the array's first position and its end) X(h) is the start of tile
// (h, l_new) and the end of tile (h, l_old), for every digit h.  Only blocks
// holding a boundary (~1% at 1080p) do this; each tile's start and end are
// written once with plain stores.  Tiles whose two values are equal (empty)
// or never written are set to -1 afterwards (k_fix_ranges).
template <typename K, int RB, int ITEMS>
__device__ __forceinline__ void block_ranges(ScatterLds<K, RB, kRadixThreads * ITEMS>& s, const K (&key)[ITEMS],
                                             int nvalid, bool is_last, int shift, K prev_key, bool has_prev,
                                             int32_t* __restrict__ start, int32_t* __restrict__ end,
                                             uint32_t n_tiles) {
  static_assert(sizeof(K) == 4, "ranges are built on 32-bit tile ids");
  constexpr int BINS = 1 << RB;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const uint32_t lowmask = shift >= 32 ? ~0u : (1u << shift) - 1u;
  const uint32_t lf = (uint32_t)s.kfirst & lowmask, ll = (uint32_t)s.klast & lowmask;
  const uint32_t lprev = (uint32_t)prev_key & lowmask;
  const bool starts = !has_prev || lprev != lf;
  if (!(starts || lf != ll || is_last)) return;  // block-uniform
  // boundary flags: the low bits differ from the previous position's
  bool isb[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const int p = w * (64 * ITEMS) + r * 64 + lane;
    const K up = (K)__shfl_up((uint32_t)key[r], 1);
    K wrap;
    if (r > 0) wrap = (K)__shfl((uint32_t)key[r - 1], 63);
    else wrap = w > 0 ? s.wlast[w - 1] : prev_key;
    const K pk = lane > 0 ? up : wrap;
    const bool has_pred = p > 0 || has_prev;
    isb[r] = p < nvalid && (!has_pred || (((uint32_t)pk & lowmask) != ((uint32_t)key[r] & lowmask)));
  }
  int cur = -1;
  for (;;) {
    if (t == 0) {
      s.nextb = 0x7fffffff;
      s.lnew = kNoBucket;
      s.lold = kNoBucket;
    }
    for (int i = t; i < BINS; i += kRadixThreads) s.pref[i] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const int p = w * (64 * ITEMS) + r * 64 + lane;
      if (isb[r] && p > cur) atomicMin(&s.nextb, p);
    }
    __syncthreads();
    int pb = s.nextb;
    if (pb == 0x7fffffff) {
      if (!(is_last && cur < nvalid)) break;  // block-uniform
      pb = nvalid;                            // the end of the array
    }
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const int p = w * (64 * ITEMS) + r * 64 + lane;
      const uint32_t lo = (uint32_t)key[r] & lowmask;
      if (p < pb) atomicAdd(&s.pref[(uint32_t)(key[r] >> shift) & (BINS - 1)], 1u);
      if (p == pb && p < nvalid) s.lnew = lo;  // none at the end of the array
      if (p == pb - 1) s.lold = lo;
    }
    if (pb == 0 && has_prev && t == 0) s.lold = lprev;
    __syncthreads();
    const uint32_t lnew = s.lnew, lold = s.lold;
    for (int h = t; h < BINS; h += kRadixThreads) {
      const int32_t X = (int32_t)(s.goff[h] + s.pref[h]);
      const uint32_t hi = (uint32_t)h << shift;
      if (lnew != kNoBucket && (hi | lnew) < n_tiles) start[hi | lnew] = X;
      if (lold != kNoBucket && (hi | lold) < n_tiles) end[hi | lold] = X;
    }
    cur = pb;
    __syncthreads();
  }
}

// Stable scatter of one block of 256 * ITEMS items.  Item r of lane `lane` in
// wave w sits at block position w*64*ITEMS + r*64 + lane (key[r], val[r]; keys of
// invalid positions are all-ones).  Ranks by digit (key >> shift) & (2^RB-1)
// with wave64 ballot matching, permutes through LDS, writes each digit's run
// at offs[digit][block] (exclusive-scanned histogram).  kout may be null
// (keys not needed).  With RANGES this is the last pass over tile ids and
// block_ranges writes the tile ranges; prev_key is the key at position
// base - 1 (when has_prev).
template <typename K, int RB, bool RANGES, int ITEMS>
__device__ __forceinline__ void block_scatter(ScatterLds<K, RB, kRadixThreads * ITEMS>& s, K (&key)[ITEMS],
                                              int32_t (&val)[ITEMS], int64_t base, int64_t P, int shift,
                                              const uint32_t* __restrict__ offs, int nblk, K* __restrict__ kout,
                                              int32_t* __restrict__ vout, int32_t* __restrict__ start,
                                              int32_t* __restrict__ end, uint32_t n_tiles, int blk, K prev_key,
                                              bool has_prev) {
  constexpr int BINS = 1 << RB;
  constexpr int TILE = kRadixThreads * ITEMS;
  static_assert(BINS <= kRadixThreads, "one digit per thread in the block prefix");
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int nvalid = (int)min((int64_t)TILE, P - base);
  for (int i = lane; i < BINS; i += 64) s.cnt[w][i] = 0;
  if (t < BINS) s.goff[t] = offs[(int64_t)t * nblk + blk];
  if constexpr (RANGES) {
    if (lane == 63) s.wlast[w] = key[ITEMS - 1];
    if (t == 0) s.kfirst = key[0];
    const int pl = nvalid - 1;  // the last valid position's owner publishes its key
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      if (w * (64 * ITEMS) + r * 64 + lane == pl) s.klast = key[r];
  }
  __syncthreads();

  uint32_t rank[ITEMS];
  // Wave w ranks its 64 * ITEMS items 64 per round in position order, so
  // ranks are stable.  Lanes with the same digit ("peers") are found one
  // digit bit at a time: m is the lane's bit as an all-zeros/all-ones word,
  // and a lane keeps the ballot bits that agree with it, ~(ballot ^ m), in
  // each 32-lane half.
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint32_t d = (uint32_t)(key[r] >> shift) & (BINS - 1);
    uint32_t plo = ~0u, phi = ~0u;
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const uint32_t m = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
      const uint64_t bal = __ballot(m != 0u);
      plo &= ~((uint32_t)bal ^ m);
      phi &= ~((uint32_t)(bal >> 32) ^ m);
    }
    // peers in lanes below this one, and in the whole wave
    const uint32_t lower = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
    const uint32_t old = s.cnt[w][d];
    __builtin_amdgcn_wave_barrier();
    if (lower == 0) s.cnt[w][d] = old + (uint32_t)(__popc(plo) + __popc(phi));
    __builtin_amdgcn_wave_barrier();
    rank[r] = old + lower;
  }
  __syncthreads();
  // Digit t: per-wave counts -> wave offsets; block prefix over digits.
  uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  if (t < BINS) {
    c0 = s.cnt[0][t];
    c1 = s.cnt[1][t];
    c2 = s.cnt[2][t];
    c3 = s.cnt[3][t];
  }
  const uint32_t tot = c0 + c1 + c2 + c3;
  uint32_t inc = tot;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s.wsum[w] = inc;
  __syncthreads();
  uint32_t pre = 0;
  for (int i = 0; i < w; ++i) pre += s.wsum[i];
  if (t < BINS) {
    const uint32_t dstart = pre + inc - tot;
    s.dstart[t] = dstart;
    s.base[0][t] = dstart;
    s.base[1][t] = dstart + c0;
    s.base[2][t] = dstart + c0 + c1;
    s.base[3][t] = dstart + c0 + c1 + c2;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint32_t d = (uint32_t)(key[r] >> shift) & (BINS - 1);
    const uint32_t lp = s.base[w][d] + rank[r];
    s.keys[lp] = key[r];
    s.vals[lp] = val[r];
  }
  __syncthreads();
  if constexpr (RANGES)
    block_ranges<K, RB, ITEMS>(s, key, nvalid, base + nvalid >= P, shift, prev_key, has_prev, start, end, n_tiles);
#pragma unroll 4
  for (int i = t; i < TILE; i += kRadixThreads) {
    if (i < nvalid) {
      const K k = s.keys[i];
      const uint32_t d = (uint32_t)(k >> shift) & (BINS - 1);
      const int64_t gp = (int64_t)s.goff[d] + (i - (int64_t)s.dstart[d]);
      if (kout) kout[gp] = k;
      vout[gp] = s.vals[i];
    }
  }
}

template <typename K, int RB, bool RANGES>
__global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(const K* __restrict__ kin,
                                                                 const int32_t* __restrict__ vin,
                                                                 K* __restrict__ kout, int32_t* __restrict__ vout,
                                                                 int64_t P, int shift,
                                                                 const uint32_t* __restrict__ offs, int nblk,
                                                                 int32_t* __restrict__ start,
                                                                 int32_t* __restrict__ end, uint32_t n_tiles) {
  __shared__ ScatterLds<K, RB, kRadixTile> s;
  const int blk = xcd_tile(blockIdx.x, nblk);
  if (blk >= nblk) return;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int64_t base = (int64_t)blk * kRadixTile;
  K key[kRadixItems];
  int32_t val[kRadixItems];
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const int64_t p = base + w * (64 * kRadixItems) + r * 64 + lane;
    const bool valid = p < P;
    key[r] = valid ? kin[p] : (K)~(K)0;
    val[r] = valid ? vin[p] : 0;
  }
  const K prev_key = RANGES && base > 0 ? kin[base - 1] : (K)0;
  block_scatter<K, RB, RANGES, kRadixItems>(s, key, val, base, P, shift, offs, nblk, kout, vout, start, end,
                                            n_tiles, blk, prev_key, base > 0);
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

// Depth keys and identity values for the V visible Gaussians, and their
// tile bounds packed as one 16-byte record {tl.x, tl.y, br.x, br.y}, so the
// gather in depth order below touches one cache line per Gaussian.
// shift > 0: the bounds are binned per block of 2^shift x 2^shift tiles
// (bounds are >= 0, so >> is the floor division).
__global__ void k_depth_keys(const float* __restrict__ depth, const int32_t* __restrict__ tl,
                             const int32_t* __restrict__ br, int64_t V, int shift, uint32_t* __restrict__ dkey,
                             int32_t* __restrict__ dval, int4* __restrict__ tlbr) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  dkey[v] = __float_as_uint(depth[v]);
  dval[v] = (int32_t)v;
  const int2 a = reinterpret_cast<const int2*>(tl)[v];
  const int2 b = reinterpret_cast<const int2*>(br)[v];
  tlbr[v] = make_int4(a.x >> shift, a.y >> shift, b.x >> shift, b.y >> shift);
}

// Pair count and tile rectangle of the Gaussian at depth rank r
// (gs/culling.py:33-35): rect = {first tile id, tiles per row, rows, index};
// the pair of local index l is tile first + (l % rows) * per_row + l / rows
// (x outer, y inner).  Multi-view: view[r] selects the record's camera and
// its tile_base (first tile id of the view in the batch numbering).
__global__ void k_sorted_rects(const int32_t* __restrict__ dval, const uint32_t* __restrict__ view,
                               const gsmi_view* __restrict__ views, int n_tiles_w, const int4* __restrict__ tlbr,
                               int64_t V, int32_t* __restrict__ cnt, int4* __restrict__ rect) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V) return;
  const int v = dval[r];
  int base = 0, ntw = n_tiles_w;
  if (views) {
    const gsmi_view& cam = views[view ? view[r] : 0];
    base = cam.tile_base;
    ntw = cam.n_tiles_w;
  }
  const int4 q = tlbr[v];
  const int2 a = make_int2(q.x, q.y), b = make_int2(q.z, q.w);
  cnt[r] = (b.x - a.x + 1) * (b.y - a.y + 1);
  rect[r] = make_int4(base + a.y * ntw + a.x, ntw, b.y - a.y + 1, v);
}

// View of visible entry dval[r]: the largest b with bases[2b] <= dval[r]
// (record bases staged in LDS; at most kMaxViewsLds views).
constexpr int kMaxViewsLds = 1024;

__global__ __launch_bounds__(256) void k_view_keys(const int32_t* __restrict__ dval, int64_t V,
                                                   const int64_t* __restrict__ bases, int B,
                                                   uint32_t* __restrict__ view) {
  __shared__ int32_t s_base[kMaxViewsLds];
  for (int b = threadIdx.x; b < B && b < kMaxViewsLds; b += blockDim.x) s_base[b] = (int32_t)bases[2 * b];
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V) return;
  const int32_t v = dval[r];
  int lo = 0, hi = min(B, kMaxViewsLds) - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s_base[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  view[r] = (uint32_t)lo;
}

// Expansion of the depth-ordered Gaussians into (tile, index) pairs.  Block b
// owns output slots [b*kExpandTile, (b+1)*kExpandTile) (2048: half the radix
// tile, so that the fused scatter needs 27 KB of LDS and 5 workgroups fit a
// CU).  The offsets of the Gaussians overlapping that range are staged in LDS
// (every visible Gaussian has at least one pair, so at most kExpandTile + 1
// of them); thread t takes the kExpandItems consecutive slots starting at
// t*kExpandItems, finds the Gaussian of the first by binary search and walks
// forward.
constexpr int kExpandDigitBits = 7;
constexpr int kExpandItems = 8;
constexpr int kExpandTile = kRadixThreads * kExpandItems;

// Tile rectangles of the first kRectCache staged Gaussians are staged in LDS
// too (a block holds ~100 Gaussians at 21 pairs each), so the expansion
// reads them from LDS instead of a dependent global load per Gaussian.
constexpr int kRectCache = 512;

struct ExpandLds {
  int32_t off[kExpandTile + 2];
};

// first[b] = depth rank of the Gaussian whose pair range holds slot
// b * kExpandTile.  Every visible Gaussian has at least one pair, so every
// block start lies in exactly one range; entries a malformed count leaves
// unset keep the fill value -1 and are clamped by the reader.
__global__ void k_block_first(const int32_t* __restrict__ off, int64_t V, int64_t P, int32_t* __restrict__ first,
                              int nblk) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V) return;
  const int64_t lo = off[r];
  const int64_t hi = r + 1 < V ? (int64_t)off[r + 1] : P;
  for (int64_t b = (lo + kExpandTile - 1) / kExpandTile; b * kExpandTile < hi && b < nblk; ++b)
    if (b >= 0) first[b] = (int32_t)r;
}

// Stages the offsets of the Gaussians from first[b] to first[b + 1]
// (inclusive: the last may start exactly at the next block) and the
// rectangles of the first kRectCache of them; returns (r0, n) = first depth
// rank and count of staged Gaussians.
__device__ __forceinline__ int2 expand_stage(ExpandLds& s, int4* s_rect, const int32_t* __restrict__ off,
                                             const int4* __restrict__ rect, const int32_t* __restrict__ first,
                                             int64_t V, int nblk, int b) {
  const int t = threadIdx.x;
  const int hiV = (int)(V - 1);
  const int r0 = min(max(first[b], 0), hiV);
  const int r1 = b + 1 < nblk ? min(max(first[b + 1], r0), hiV) : hiV;
  // <= kExpandTile + 2 when every Gaussian has at least one pair; clamped so
  // inconsistent bounds cannot index past the LDS array.
  const int n = max(1, min(r1 - r0 + 1, kExpandTile + 2));
  for (int i = t; i < n; i += kRadixThreads) s.off[i] = off[r0 + i];
  for (int i = t; i < min(n, kRectCache); i += kRadixThreads) s_rect[i] = rect[r0 + i];
  __syncthreads();
  return make_int2(r0, n);
}

// Tiles and indices of this thread's slots p0 + t*kExpandItems + k; invalid
// slots (past P) get tile 0xffffffff.
__device__ __forceinline__ void expand_slots(const ExpandLds& s, const int4* s_rect, int2 rn,
                                             const int4* __restrict__ rect, int64_t P, int blk,
                                             uint32_t (&tile)[kExpandItems], int32_t (&val)[kExpandItems]) {
  const int t = threadIdx.x;
  const int r0 = rn.x, n = rn.y;
  int64_t p = (int64_t)blk * kExpandTile + (int64_t)t * kExpandItems;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int64_t)s.off[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  int j = lo;
  int4 rc = j < kRectCache ? s_rect[j] : rect[r0 + j];
  // Slot `local` of a Gaussian is tile (row, col) = (local % ny, local / ny)
  // of its rectangle.  One division for the first slot; later slots of the
  // same Gaussian step the row.  A Gaussian the loop moves to starts at
  // slot p exactly (offsets are non-decreasing and the previous slot lay
  // before off[j]), so it restarts at row = col = 0; with inconsistent
  // counts the tiles are wrong but the same in both expansion kernels.
  int ny = max(rc.z, 1);
  const int local0 = (int)(p - s.off[j]);
  int col = local0 / ny;
  int row = local0 - col * ny;
#pragma unroll
  for (int k = 0; k < kExpandItems; ++k, ++p) {
    tile[k] = 0xffffffffu;
    val[k] = 0;
    if (p < P) {
      bool moved = false;
      while (j + 1 < n && (int64_t)s.off[j + 1] <= p) {
        ++j;
        moved = true;
      }
      if (moved) {
        rc = j < kRectCache ? s_rect[j] : rect[r0 + j];
        ny = max(rc.z, 1);
        row = (int)(p - s.off[j]);
        col = 0;
      }
      tile[k] = (uint32_t)(rc.x + row * rc.y + col);
      val[k] = rc.w;
      if (++row == ny) {
        row = 0;
        ++col;
      }
    }
  }
}

// Pass-1 histogram of the expansion: counts of the low kExpandDigitBits
// tile-id bits per block, hist[digit][block].
__global__ __launch_bounds__(kRadixThreads) void k_expand_hist(const int32_t* __restrict__ off,
                                                               const int32_t* __restrict__ first,
                                                               const int4* __restrict__ rect, int64_t V, int64_t P,
                                                               uint32_t* __restrict__ hist,
                                                               int nblk) {
  constexpr int BINS = 1 << kExpandDigitBits;
  __shared__ ExpandLds s;
  __shared__ int4 s_rect[kRectCache];
  __shared__ uint32_t s_cnt[4][BINS];
  const int blk = xcd_tile(blockIdx.x, nblk);
  if (blk >= nblk) return;
  const int t = threadIdx.x, w = t >> 6;
  for (int i = t; i < 4 * BINS; i += kRadixThreads) (&s_cnt[0][0])[i] = 0;
  const int2 rn = expand_stage(s, s_rect, off, rect, first, V, nblk, blk);
  uint32_t tile[kExpandItems];
  int32_t val[kExpandItems];
  expand_slots(s, s_rect, rn, rect, P, blk, tile, val);
#pragma unroll
  for (int k = 0; k < kExpandItems; ++k)
    if (tile[k] != 0xffffffffu) atomicAdd(&s_cnt[w][tile[k] & (BINS - 1)], 1u);
  __syncthreads();
  if (t < BINS) hist[(int64_t)t * nblk + blk] = s_cnt[0][t] + s_cnt[1][t] + s_cnt[2][t] + s_cnt[3][t];
}

// Block position q of the staged expansion -> LDS word (64-word XOR
// swizzle: both the per-thread consecutive writes and the lane-interleaved
// reads are free of bank conflicts).
__device__ __forceinline__ int swz(int q) { return q ^ ((q >> 6) & 63); }

// Pass 1 of the pair sort fused with the expansion: the block's pairs are
// generated again, staged in LDS in slot order, and scattered stably on the
// low kExpandDigitBits tile-id bits (offsets from k_expand_hist, scanned).
template <bool RANGES>
__global__ __launch_bounds__(kRadixThreads) void k_expand_scatter(const int32_t* __restrict__ off,
                                                                  const int32_t* __restrict__ first,
                                                                  const int4* __restrict__ rect, int64_t V,
                                                                  int64_t P,
                                                                  const uint32_t* __restrict__ offs, int nblk,
                                                                  uint32_t* __restrict__ kout,
                                                                  int32_t* __restrict__ vout,
                                                                  int32_t* __restrict__ start,
                                                                  int32_t* __restrict__ end, uint32_t n_tiles) {
  __shared__ ExpandLds se;
  __shared__ ScatterLds<uint32_t, kExpandDigitBits, kExpandTile> s;
  const int blk = xcd_tile(blockIdx.x, nblk);
  if (blk >= nblk) return;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  int4* s_rect = reinterpret_cast<int4*>(s.vals);  // free until the pairs are staged below
  static_assert(sizeof(s.vals) >= kRectCache * sizeof(int4), "rect cache does not fit");
  const int2 rn = expand_stage(se, s_rect, off, rect, first, V, nblk, blk);
  {
    uint32_t tile[kExpandItems];
    int32_t val[kExpandItems];
    expand_slots(se, s_rect, rn, rect, P, blk, tile, val);
    __syncthreads();  // every rect read before s.vals is overwritten
#pragma unroll
    for (int k = 0; k < kExpandItems; ++k) {
      const int q = swz(t * kExpandItems + k);
      s.keys[q] = tile[k];
      s.vals[q] = val[k];
    }
  }
  __syncthreads();
  uint32_t key[kExpandItems];
  int32_t val[kExpandItems];
#pragma unroll
  for (int r = 0; r < kExpandItems; ++r) {
    const int q = swz(w * (64 * kExpandItems) + r * 64 + lane);
    key[r] = s.keys[q];
    val[r] = s.vals[q];
  }
  // block_scatter's first barrier orders these reads before its LDS writes;
  // as the only pass (shift 0) every pair is in one bucket, so the key
  // before the block is irrelevant
  block_scatter<uint32_t, kExpandDigitBits, RANGES, kExpandItems>(s, key, val, (int64_t)blk * kExpandTile, P, 0,
                                                                  offs, nblk, kout, vout, start, end, n_tiles, blk,
                                                                  0u, blk > 0);
}

// Tiles with start == end (empty, or never written: both -1) -> -1, -1.
__global__ void k_fix_ranges(int32_t* start, int32_t* end, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && start[i] >= end[i]) {
    start[i] = -1;
    end[i] = -1;
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
    hipLaunchKernelGGL((k_radix_hist<K, RB>), dim3(xcd_grid(nblk)), dim3(kRadixThreads), 0, st, kin, P, shift, hist,
                       nblk);
    GSMI_CHECK_LAUNCH();
    int rc = exclusive_scan<int32_t>(reinterpret_cast<int32_t*>(hist), reinterpret_cast<int32_t*>(hist), nh,
                                     nullptr, scan_ws, st);
    if (rc) return rc;
    hipLaunchKernelGGL((k_radix_scatter<K, RB, false>), dim3(xcd_grid(nblk)), dim3(kRadixThreads), 0, st, kin, vin,
                       kout, vout, P, shift, hist, nblk, nullptr, nullptr, 0u);
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
  uint32_t* dkey;  // [V] depth keys (radix-sorted in place)
  int32_t* dval;   // [V] indices in depth order
  int32_t* cnt;    // [V] pair counts in depth order -> exclusive offsets
  int4* rect;      // [V] tile rectangles in depth order
  int4* tlbr;      // [V] {tl, br} in index order
  uint32_t* view;  // [V] view of each entry in depth order (multi-view)
  int32_t* total;
  uint32_t *ka, *kb;  // [P] tile ids, ping-pong
  int32_t *va, *vb;   // [P] indices, ping-pong
  uint32_t* hist;     // [nblk * 256]
  int32_t* first;     // [expansion blocks] first Gaussian of each block
  void* scan_ws;
  void* sort_ws;  // depth sort
};

static int64_t pair_blocks(int64_t P) { return (P + kRadixTile - 1) / kRadixTile; }

static int64_t expand_blocks(int64_t P) { return (P + kExpandTile - 1) / kExpandTile; }

static size_t bin_ws_bytes(int64_t V, int64_t P) {
  const int64_t nh = pair_blocks(P) * 256;
  return align_up(V * 4) * 4 + align_up(V * 16) * 2 + 256 + align_up(P * 4) * 4 + align_up(nh * 4) +
         align_up((expand_blocks(P) + 1) * 4) +
         align_up(scan_workspace_bytes(std::max(V, nh), 4)) + radix_ws_bytes<uint32_t, 8>(V);
}

static BinWs carve_bin_ws(void* ws, int64_t V, int64_t P) {
  char* p = reinterpret_cast<char*>(ws);
  auto take = [&p](size_t bytes) {
    char* r = p;
    p += align_up(bytes);
    return r;
  };
  BinWs b;
  b.dkey = reinterpret_cast<uint32_t*>(take(V * 4));
  b.dval = reinterpret_cast<int32_t*>(take(V * 4));
  b.cnt = reinterpret_cast<int32_t*>(take(V * 4));
  b.rect = reinterpret_cast<int4*>(take(V * 16));
  b.tlbr = reinterpret_cast<int4*>(take(V * 16));
  b.view = reinterpret_cast<uint32_t*>(take(V * 4));
  b.total = reinterpret_cast<int32_t*>(take(256));
  b.ka = reinterpret_cast<uint32_t*>(take(P * 4));
  b.kb = reinterpret_cast<uint32_t*>(take(P * 4));
  b.va = reinterpret_cast<int32_t*>(take(P * 4));
  b.vb = reinterpret_cast<int32_t*>(take(P * 4));
  b.hist = reinterpret_cast<uint32_t*>(take(pair_blocks(P) * 256 * 4));
  b.first = reinterpret_cast<int32_t*>(take((expand_blocks(P) + 1) * 4));
  b.scan_ws = take(scan_workspace_bytes(std::max(V, pair_blocks(P) * 256), 4));
  b.sort_ws = p;
  return b;
}

// One radix pass over the pair list on tile-id bits [shift, shift + RB).
template <int RB>
static int pair_pass(const BinWs& w, int64_t P, int nblk, int shift, bool last, uint32_t* kin, int32_t* vin,
                     uint32_t* kout, int32_t* vout, int32_t* start, int32_t* end, unsigned nt, hipStream_t st) {
  const dim3 grid(xcd_grid(nblk)), block(kRadixThreads);
  hipLaunchKernelGGL((k_radix_hist<uint32_t, RB>), grid, block, 0, st, kin, P, shift, w.hist, nblk);
  GSMI_CHECK_LAUNCH();
  int rc = exclusive_scan<int32_t>(reinterpret_cast<int32_t*>(w.hist), reinterpret_cast<int32_t*>(w.hist),
                                   (int64_t)nblk << RB, nullptr, w.scan_ws, st);
  if (rc) return rc;
  if (last) {
    hipLaunchKernelGGL((k_radix_scatter<uint32_t, RB, true>), grid, block, 0, st, kin, vin, kout, vout, P, shift,
                       w.hist, nblk, start, end, nt);
  } else {
    hipLaunchKernelGGL((k_radix_scatter<uint32_t, RB, false>), grid, block, 0, st, kin, vin, kout, vout, P, shift,
                       w.hist, nblk, nullptr, nullptr, nt);
  }
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

// Pair sort on the tile id: pass 1 fused with the expansion (low 7 bits),
// then the remaining bits in passes of at most 8, split evenly (13 bits at
// 1920x1080: one 6-bit pass, 64 runs per block instead of 256); the last
// pass writes ids (and tiles when asked) and the ranges.  Returns the sorted
// tile ids in *tiles (null unless want_tiles).
static int pair_sort(const BinWs& w, int64_t V, int64_t P, int n_tiles, int32_t* ids,
                     int32_t* start, int32_t* end, bool want_tiles, uint32_t** tiles, hipStream_t st) {
  const int nblk = (int)pair_blocks(P);
  const int nblk1 = (int)expand_blocks(P);  // expansion blocks (pass 1)
  const int bits = std::max(bits_for(n_tiles), 1);
  const unsigned nt = (unsigned)n_tiles;
  *tiles = nullptr;
  static_assert((kRadixTile / kExpandTile) * (1 << kExpandDigitBits) <= 256, "pass-1 histogram exceeds the buffer");
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((nblk1 + 255) / 256)), dim3(256), 0, st, w.first, (int64_t)nblk1, -1);
  hipLaunchKernelGGL(k_block_first, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, st, w.cnt, V, P, w.first, nblk1);
  const dim3 grid1(xcd_grid(nblk1)), block1(kRadixThreads);
  hipLaunchKernelGGL(k_expand_hist, grid1, block1, 0, st, w.cnt, w.first, w.rect, V, P, w.hist, nblk1);
  GSMI_CHECK_LAUNCH();
  int rc = exclusive_scan<int32_t>(reinterpret_cast<int32_t*>(w.hist), reinterpret_cast<int32_t*>(w.hist),
                                   (int64_t)nblk1 << kExpandDigitBits, nullptr, w.scan_ws, st);
  if (rc) return rc;
  if (bits <= kExpandDigitBits) {
    uint32_t* ko = want_tiles ? w.ka : nullptr;
    hipLaunchKernelGGL((k_expand_scatter<true>), grid1, block1, 0, st, w.cnt, w.first, w.rect, V, P, w.hist, nblk1,
                       ko, ids, start, end, nt);
    GSMI_CHECK_LAUNCH();
    *tiles = ko;
    return GSMI_OK;
  }
  hipLaunchKernelGGL((k_expand_scatter<false>), grid1, block1, 0, st, w.cnt, w.first, w.rect, V, P, w.hist, nblk1,
                     w.ka, w.va, nullptr, nullptr, nt);
  GSMI_CHECK_LAUNCH();
  uint32_t *kin = w.ka, *kout = w.kb;
  int32_t *vin = w.va, *vout = w.vb;
  const int rem = bits - kExpandDigitBits;
  const int npass = (rem + 7) / 8;
  int shift = kExpandDigitBits;
  for (int pass = 0; pass < npass; ++pass) {
    const bool last = pass == npass - 1;
    const int left = bits - shift;
    const int rb = (left + (npass - pass) - 1) / (npass - pass);  // ceil(left / passes left), <= 8
    uint32_t* ko = last ? (want_tiles ? kout : nullptr) : kout;
    int32_t* vo = last ? ids : vout;
    // Fewer than 4 bits run as 4: the extra digit bits are above the tile
    // id's width (zero) or belong to the next pass, which re-sorts them stably.
    switch (rb) {
      case 5: rc = pair_pass<5>(w, P, nblk, shift, last, kin, vin, ko, vo, start, end, nt, st); break;
      case 6: rc = pair_pass<6>(w, P, nblk, shift, last, kin, vin, ko, vo, start, end, nt, st); break;
      case 7: rc = pair_pass<7>(w, P, nblk, shift, last, kin, vin, ko, vo, start, end, nt, st); break;
      case 8: rc = pair_pass<8>(w, P, nblk, shift, last, kin, vin, ko, vo, start, end, nt, st); break;
      default: rc = pair_pass<4>(w, P, nblk, shift, last, kin, vin, ko, vo, start, end, nt, st); break;
    }
    if (rc) return rc;
    if (last) {
      *tiles = ko;
    } else {
      std::swap(kin, kout);
      std::swap(vin, vout);
    }
    shift += rb;
  }
  return GSMI_OK;
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

// Depth-first binning of V records (tile rects tl/br, depth) into sorted
// (tile, record) lists.  Single view: views == null, tiles numbered
// y * n_tiles_w + x.  Multi-view: records are view-major with bases[2b] the
// first record of view b; tiles are numbered with each view's tile_base.
static int bin_impl(int64_t V, int64_t P, const int32_t* tl, const int32_t* br, const float* depth,
                    const gsmi_view* views, const int64_t* bases, int B, int n_tiles, int n_tiles_w,
                    int32_t* gaussian_ids, int32_t* start, int32_t* end, uint64_t* sorted_keys, void* workspace,
                    int flags, hipStream_t st) {
  BinWs w = carve_bin_ws(workspace, V, P);
  int rc;
  if (V > 0) {
    const unsigned g = (unsigned)((V + 255) / 256);
    hipLaunchKernelGGL(k_depth_keys, dim3(g), dim3(256), 0, st, depth, tl, br, V, (flags >> 8) & 7, w.dkey,
                       w.dval, w.tlbr);
    GSMI_CHECK_LAUNCH();
    rc = radix_sort<uint32_t, 8>(V, w.dkey, w.dval, 32, w.sort_ws, st);
    if (rc) return rc;
    const uint32_t* view = nullptr;
    if (views && B > 1) {
      // stable pass on the view id: (view, depth bits, record) order
      hipLaunchKernelGGL(k_view_keys, dim3(g), dim3(256), 0, st, w.dval, V, bases, B, w.view);
      GSMI_CHECK_LAUNCH();
      rc = radix_sort<uint32_t, 8>(V, w.view, w.dval, bits_for(B), w.sort_ws, st);
      if (rc) return rc;
      view = w.view;
    }
    hipLaunchKernelGGL(k_sorted_rects, dim3(g), dim3(256), 0, st, w.dval, view, views, n_tiles_w, w.tlbr, V, w.cnt,
                       w.rect);
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
  const int nb = (n_tiles + 255) / 256;
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, st, start, (int64_t)n_tiles, -1);
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, st, end, (int64_t)n_tiles, -1);
  GSMI_CHECK_LAUNCH();
  uint32_t* tiles = nullptr;
  if (P > 0 && V > 0) {
    rc = pair_sort(w, V, P, n_tiles, gaussian_ids, start, end, sorted_keys != nullptr, &tiles, st);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_fix_ranges, dim3(nb), dim3(256), 0, st, start, end, n_tiles);
  GSMI_CHECK_LAUNCH();
  if (sorted_keys && P > 0 && tiles) {
    hipLaunchKernelGGL(k_make_keys, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, tiles, gaussian_ids, depth,
                       P, sorted_keys);
    GSMI_CHECK_LAUNCH();
  }
  return GSMI_OK;
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
  if (((flags >> 8) & 7) > 4 || (flags & ~0x701)) return GSMI_ERR_ARG;  // bit 0 and bits 8-10 only
  const int64_t n_tiles = (int64_t)n_tiles_h * n_tiles_w;
  if (n_tiles > (1 << 24)) return GSMI_ERR_UNSUPPORTED;
  if (!workspace || workspace_bytes < bin_ws_bytes(V, P)) return GSMI_ERR_WORKSPACE;
  return bin_impl(V, P, tl, br, depth, nullptr, nullptr, 1, (int)n_tiles, n_tiles_w, gaussian_ids, start, end,
                  sorted_keys, workspace, flags, (hipStream_t)stream);
}

extern "C" int gsmi_bin_views(int64_t V, int64_t P, int B, const gsmi_view* views, const int64_t* bases,
                              int n_tiles_total, const int32_t* tl, const int32_t* br, const float* depth,
                              int32_t* gaussian_ids, int32_t* start, int32_t* end, void* workspace,
                              size_t workspace_bytes, void* stream) {
  if (V < 0 || P < 0 || B <= 0 || !views || !bases || n_tiles_total <= 0 || !start || !end) return GSMI_ERR_ARG;
  if (V > 0 && (!tl || !br || !depth)) return GSMI_ERR_ARG;
  if (((uintptr_t)tl & 7) || ((uintptr_t)br & 7) || ((uintptr_t)views & 7)) return GSMI_ERR_ARG;
  if (P > 0 && !gaussian_ids) return GSMI_ERR_ARG;
  if (P >= (int64_t(1) << 31) || V >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  if (n_tiles_total > (1 << 24) || B > kMaxViewsLds) return GSMI_ERR_UNSUPPORTED;
  if (!workspace || workspace_bytes < bin_ws_bytes(V, P)) return GSMI_ERR_WORKSPACE;
  return bin_impl(V, P, tl, br, depth, views, bases, B, n_tiles_total, 0, gaussian_ids, start, end, nullptr,
                  workspace, 0, (hipStream_t)stream);
}
