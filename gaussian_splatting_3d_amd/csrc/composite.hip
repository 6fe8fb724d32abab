// Per-tile alpha compositing on gfx950: forward and reverse-order adjoint.
//
// One 256-thread workgroup (4 wave64) per 16x16 tile; wave w owns the 8x8
// quadrant (w & 1, w >> 1) of the tile (squarer than rows of 16, so fewer
// Gaussians reach a wave's pixels).  Gaussians of the tile's sorted list are staged in LDS in
// batches; every wave walks the entries of the batch that can reach one of
// its pixels (a relevance mask built once per batch, one entry per lane) in
// lockstep (wave-uniform loop) and leaves the batch as soon as a ballot
// shows all 64 of its pixels saturated; the workgroup stops loading batches
// once no pixel is alive.
//
// Decision exactness.  The skip test a*G < 1/255 and the transmittance
// break use the reference's operation sequence (kernels.h:172-193,
// vol_render_sh.h:120-168) with the specified exponential, so they agree bit
// for bit with the oracle.  A per-Gaussian bound rlim lets most far-away
// pixels be rejected from q * (1/det) alone: for r > rlim the product a*G is
// below 1/255 by a relative margin of 5e-4, far above the few-ulp gap between
// q*(1/det) and q/det.  Everything else takes the exact path.
#include <algorithm>
#include <mutex>

#include "gsmi_device.h"
#include "scan.h"
#include "serve_internal.h"

namespace gsmi {

// Gaussians staged per LDS batch.  64 (not more) keeps the SH-4 forward's
// LDS at 15 KB so that 8 wave64 per SIMD fit; measured 1.29 -> 1.10 ms at
// 3M Gaussians / 1080p against 128.
#ifndef GSMI_FWD_BATCH
#define GSMI_FWD_BATCH 64
#endif
constexpr int kBatch = GSMI_FWD_BATCH;

struct RasterArgs {
  const float* mean;  // [V,2]
  const float* cov;   // [V,4]
  const float* sh;    // SH rows (or RGB colour [V,3] for the RGB kernels)
  const int32_t* sh_index;
  const float* alpha;  // [V]
  const int32_t* start;
  const int32_t* end;
  const int32_t* ids;
  const float* topleft;
  const float* c2w;
  const float* bg;
  float* out;
  float* T_final;
  int32_t* n_done;
  float* pre_bg;
  // backward only
  const float* out_saved;
  const float* T_saved;
  const int32_t* n_done_saved;
  const float* pre_bg_saved;
  const float* grad_out;
  float* grad_mean;
  float* grad_cov;
  float* grad_sh;
  float* grad_alpha;
  int64_t sh_sg, sh_sc;
  float psx, psy, thresh;
  int n_tiles_h, n_tiles_w, H, W;
  int write_empty;
  // Multi-view batch: tiles of all views in one launch (views == null: the
  // single view described by topleft/c2w/psx/psy/H/W/n_tiles_*).
  const gsmi_view* views;
  int n_views;
  int n_tiles_total;
  // Optional launch order: block b works on tile tile_order[b] (a permutation
  // of the tiles); null: the XCD-contiguous mapping.
  const int32_t* tile_order;
  // Two-level lists (forward): block-of-tiles lists sstart/send/sids (blocks
  // of 2^super_shift tiles a side), tile bounds rtl/rbr of the records; with
  // fids the forward writes each tile's filtered list to fids from
  // foff[tile] on and its end to fend[tile].  Multi-view: bviews[b] holds
  // view b's block grid (n_tiles_w/h, tile_base in block units).
  const gsmi_view* bviews;
  const int32_t* sstart;
  const int32_t* send;
  const int32_t* sids;
  const int32_t* rtl;
  const int32_t* rbr;
  const int32_t* foff;
  int32_t* fids;
  int32_t* fend;
  int super_shift, n_super_w;
  // Multi-view forward of mixed scenes (render server): view b reads its SH
  // rows from view_sh[b] (null: sh) and writes its [H, W, 3] image to
  // view_out[b] (null: out at the view's pix_base).
  const float* const* view_sh;
  float* const* view_out;
  // Multi-view lists kept per view (the batched training path's recorded
  // two-level lists): view b's list entries live at view_ids[b] (read) and the
  // forward records them to view_fids[b], with start/end / foff/fend
  // relative to that view's array; null: the shared ids / fids.
  const int32_t* const* view_ids;
  int32_t* const* view_fids;
  // Work queue (forward): null, one tile per workgroup; else a persistent
  // launch whose workgroups take tiles from 8 counters work[x], one per band
  // of ceil(n_tiles_total / 8) consecutive tiles, starting with band
  // blockIdx.x % 8 and then the others; the counters start at 0.
  int* work;
};

// Pixel of this thread in the block's tile.
struct TileCtx {
  int tile;  // global tile id (index into start/end)
  int view;  // view of the tile (0 for a single view)
  int gx, gy;
  bool inside;
  int64_t pix;  // flat pixel index into out / T_final / n_done / grad_out
  int64_t lpix;  // pixel index within the view (gy * W + gx)
  float px, py;
  float cw[9];  // first 9 floats of the view's c2w buffer (they feed the ray)
};

__device__ __forceinline__ bool tile_setup_at(const RasterArgs& a, TileCtx& c, int tile);

__device__ __forceinline__ bool tile_setup(const RasterArgs& a, TileCtx& c) {
  int tile;
  if (a.tile_order) {
    if ((int)blockIdx.x >= a.n_tiles_total) return false;
    tile = a.tile_order[blockIdx.x];
  } else {
    tile = xcd_tile(blockIdx.x, a.n_tiles_total);
  }
  return tile_setup_at(a, c, tile);
}

// Setup for a given tile id (the work-queue launch takes tiles directly).
__device__ __forceinline__ bool tile_setup_at(const RasterArgs& a, TileCtx& c, int tile) {
  if (tile < 0 || tile >= a.n_tiles_total) return false;
  int local = tile, ntw = a.n_tiles_w, W = a.W, H = a.H;
  float tlx, tly, psx = a.psx, psy = a.psy;
  int64_t pix_base = 0;
  c.view = 0;
  if (a.views) {
    c.view = view_of_tile(a.views, a.n_views, tile);
    const gsmi_view& v = a.views[c.view];
    local = tile - v.tile_base;
    ntw = v.n_tiles_w;
    W = v.W;
    H = v.H;
    tlx = v.topleft_x;
    tly = v.topleft_y;
    psx = v.pixel_size_x;
    psy = v.pixel_size_y;
    pix_base = v.pix_base;
#pragma unroll
    for (int k = 0; k < 9; ++k) c.cw[k] = v.c2w[k];
  } else {
    tlx = a.topleft[0];
    tly = a.topleft[1];
    if (a.c2w) {
#pragma unroll
      for (int k = 0; k < 9; ++k) c.cw[k] = a.c2w[k];
    }
  }
  const int tx = local % ntw, ty = local / ntw;
  const int t = threadIdx.x;
  c.tile = tile;
  const int w = t >> 6, l = t & 63;  // wave w: 8x8 quadrant, lane l: row l >> 3, column l & 7
  c.gx = tx * kTile + ((w & 1) << 3) + (l & 7);
  c.gy = ty * kTile + ((w >> 1) << 3) + (l >> 3);
  c.inside = c.gx < W && c.gy < H;
  c.lpix = (int64_t)c.gy * W + c.gx;
  c.pix = pix_base + c.lpix;
  c.px = tlx + (float)c.gx * psx;
  c.py = tly + (float)c.gy * psy;
  return true;
}

// Per-Gaussian LDS record: {mx, my, c0, c1}, {c2, c3, det, 1/det}, {a, rlim}.
struct alignas(16) GeoA {
  float mx, my, c0, c1;
};
struct alignas(16) GeoB {
  float c2, c3, det, inv_det;
};

// Rejection radius: for q/det > rlim the Gaussian is certainly skipped.
// a*exp(-r/2) < 1/255  <=>  r > 2 ln(255 a); inflated by 1e-3 (relative and
// absolute) so that approximate log and the q*(1/det) shortcut stay safe.
__device__ __forceinline__ float reject_radius(float a) {
  const float t = a * 255.0f;
  if (!(t > 0.0f)) return -__builtin_huge_valf();
  const float r = 2.0f * __logf(t);
  return r + 1e-3f * (1.0f + fabsf(r));
}

// Whether some pixel of the rectangle [x0,x1] x [y0,y1] (one wave's pixel
// positions) can pass the rejection test rq <= rlim of a staged Gaussian.
// With q = A dx^2 + B dx dy + C dy^2 (A = c3, B = -(c1 + c2), C = c0), the
// minimum of q over the rectangle is 0 when it holds the centre, else the
// least of the four edge minima (each edge's minimiser clamped to the
// edge).  The answer is false only when that minimum exceeds rlim by 2% +
// 1e-3 and the form is positive definite with condition number below 2000
// and off-diagonals bounded by sqrt(A C), so the rounding of a pixel's q
// (relative error below 2e-3 then) cannot bring it under rlim.  Non-finite
// positions or NaN covariances keep it true.  A wave that skips such an
// entry skips only pixels that the reference skips too.
__device__ __forceinline__ bool may_reach(const GeoA& ga, const GeoB& gb, float rlim, float x0, float x1, float y0,
                                          float y1) {
  const float A = gb.c3, C = ga.c0, B = -(ga.c1 + gb.c2);
  const float dq = A * C - 0.25f * (B * B);
  const float tr = A + C;
  const float off = fabsf(ga.c1) + fabsf(gb.c2);
  const float dx0 = x0 - ga.mx, dx1 = x1 - ga.mx, dy0 = y0 - ga.my, dy1 = y1 - ga.my;
  // minimisers: y = kx x on x-edges, x = ky y on y-edges.  The hardware
  // reciprocal suffices: q at a point off the minimiser by a relative e
  // exceeds the edge minimum by at most e^2 B^2 / (4 dq) < 500 e^2 of it
  // under the conditions above (clamping moves the point no further), far
  // inside the 2% margin, and any point of an edge bounds its minimum.
  const float kx = (-0.5f * B) * __builtin_amdgcn_rcpf(C), ky = (-0.5f * B) * __builtin_amdgcn_rcpf(A);
  const float ya = fminf(fmaxf(kx * dx0, dy0), dy1), yb = fminf(fmaxf(kx * dx1, dy0), dy1);
  const float xa = fminf(fmaxf(ky * dy0, dx0), dx1), xb = fminf(fmaxf(ky * dy1, dx0), dx1);
  const float fa = (A * dx0) * dx0 + (B * dx0) * ya + (C * ya) * ya;
  const float fb = (A * dx1) * dx1 + (B * dx1) * yb + (C * yb) * yb;
  const float fc = (A * xa) * xa + (B * xa) * dy0 + (C * dy0) * dy0;
  const float fd = (A * xb) * xb + (B * xb) * dy1 + (C * dy1) * dy1;
  const bool centre = dx0 <= 0.0f && dx1 >= 0.0f && dy0 <= 0.0f && dy1 >= 0.0f;
  const float m = centre ? 0.0f : fminf(fminf(fa, fb), fminf(fc, fd));
  const float big = 0x1p100f;
  const bool well = gb.det > 0.0f && A > 0.0f && C > 0.0f && dq > 0.0f && tr * tr < 2000.0f * dq &&
                    off * off <= 4.0f * (A * C) && fabsf(dx0) < big && fabsf(dx1) < big && fabsf(dy0) < big &&
                    fabsf(dy1) < big;
  const float thr = rlim + 0.02f * fabsf(rlim) + 1e-3f;
  return !(well && m * gb.inv_det > thr);
}

// This wave's pixel rectangle: positions are monotone in the pixel row and
// column, so lanes 0, 7 (first row) and 56 (last row) hold the extremes.  The
// values are wave-uniform and kept in scalar registers.
__device__ __forceinline__ float uniform_f(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ float4 wave_pixel_rect(float px, float py) {
  const float xa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(px), 0));
  const float xb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(px), 7));
  const float ya = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(py), 0));
  const float yb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(py), 56));
  return make_float4(uniform_f(fminf(xa, xb)), uniform_f(fmaxf(xa, xb)), uniform_f(fminf(ya, yb)),
                     uniform_f(fmaxf(ya, yb)));
}

// Exact Gaussian value from the shared quadratic form (reference order).
__device__ __forceinline__ float gauss_exact(float q, float det) {
  float radial = q / det;
  if (radial < 0.0f) radial = 1000.0f;
  return exp_spec_nonpos(-0.5f * radial);  // radial >= 0 or NaN
}

// The same exp(-radial / 2) for any radial without a constant move: a
// negative radial (taken as 1000) selects the exponent -inf, whose
// exponential is 0 exactly as exp(-500) is; -0 and NaN pass unchanged.  The
// test is on radial itself: -radial/2 of the smallest negative subnormal
// rounds to +0.  (The backward keeps gauss_exact: this form holds -inf in a
// register, which it has none for.)
__device__ __forceinline__ float gauss_of_radial(float radial) {
  return exp_spec_nonpos(radial < 0.0f ? -__builtin_huge_valf() : -0.5f * radial);
}

// The quotient q / det of gauss_exact from the staged reciprocal y =
// staged_recip(det) and q0 = q * y: div_staged, or the IEEE division where
// that gives NaN or |q| < 2^-100 (wave-uniform branch; the empty asm keeps
// the compiler from executing the division unconditionally and selecting).
// Bit for bit the IEEE quotient.
__device__ __forceinline__ float radial_staged(float q, float det, float y, float q0) {
  float radial = div_staged(q, det, y, q0);
  const bool ieee = radial != radial || fabsf(q) < 0x1p-100f;
  if (wave_any(ieee)) {
    asm volatile("");
    if (ieee) radial = q / det;
  }
  return radial;
}

// gauss_exact from the staged reciprocal: the same value.
__device__ __forceinline__ float gauss_staged(float q, float det, float y, float q0) {
  return gauss_of_radial(radial_staged(q, det, y, q0));
}

// Gaussian value for the backward pass: the hardware exponential of -r/2 with
// r = q/det from the staged reciprocal (1000 for r < 0, as the reference).
// Values feed gradients only; decisions use gauss_needs_exact.
__device__ __forceinline__ float gauss_fast(float rq) {
  // exp(-r/2) = 2^(-r log2(e) / 2); a negative rq (taken as 1000) selects
  // the exponent -inf: 0 either way, without a constant move.  The test is
  // on rq itself: the smallest negative values of rq scale to +0.
  return __builtin_amdgcn_exp2f(rq < 0.0f ? -__builtin_huge_valf() : rq * -0.72134752044448170f);
}

// Where the fast value cannot take the forward's 1/255 skip decision: a*G
// within a relative 2^-12 of the threshold (far wider than the hardware
// exponential's error), not a number, or q == 0; there the exact sequence
// (gauss_exact) decides, so the decision always equals the forward's.  One
// comparison covers the window and NaN (which fails >=).
__device__ __forceinline__ bool gauss_needs_exact(float aG, float q) {
  return !(fabsf(aG - kMinRenderAlpha) >= kMinRenderAlpha * 0x1p-12f) || q == 0.0f;
}

// The wave's lanes where gauss_needs_exact holds, as a combination of the
// ballots of its single comparisons: a ballot of one comparison is that
// comparison's own lane mask, while a ballot of a combined condition goes
// through a vector register (two more instructions).
__device__ __forceinline__ uint64_t ballot_needs_exact(float aG, float q) {
  return __builtin_amdgcn_ballot_w64(!(fabsf(aG - kMinRenderAlpha) >= kMinRenderAlpha * 0x1p-12f)) |
         __builtin_amdgcn_ballot_w64(q == 0.0f);
}

// The reference's kernel_gaussian_2d (double) from float inputs.
__device__ __forceinline__ float gauss_exact_d(float dx, float dy, float c0, float c1, float c2, float c3) {
  const double C0 = c0, C1 = c1, C2 = c2, C3 = c3;
  const double det = C0 * C3 - C1 * C2;
  const double x = dx, y = dy;
  const double tmpx = x * C3 - y * C2;
  const double tmpy = -x * C1 + y * C0;
  double radial = tmpx * x + tmpy * y;
  radial /= det;
  if (radial < 0.0) radial = 1000.0;
  return (float)exp_spec_d(-0.5 * radial);
}

template <int N>
__device__ __forceinline__ float dot_exact(const float* a, const float* b) {
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < N; ++i) s = s + a[i] * b[i];
  return s;
}

template <int N>
__device__ __forceinline__ float dot_fast(const float* a, const float* b) {
  float s = a[0] * b[0];
#pragma unroll
  for (int i = 1; i < N; ++i) s = fmaf(a[i], b[i], s);
  return s;
}

// First half of stage_batch_from: ids, SH row indices (when s_row is given)
// and geometry of the batch's entries.  STAGED_RECIP: 1/det as staged_recip
// (for gauss_staged; NaN outside its range, which the relevance and distance
// tests read as "may reach").
__device__ __forceinline__ void put_gc(float2* p, int t, float al, float r) { p[t] = make_float2(al, r); }
__device__ __forceinline__ void put_gc(float4* p, int t, float al, float r) { p[t] = make_float4(al, r, 0.f, 0.f); }

// GC: float2 {opacity, rejection radius}, or float4 with two unused words
// (the stride of GeoA and GeoB, so that one address serves all three).
// Geometry of Gaussian g into batch position t; returns its opacity as
// composited (clamped at 0.99).
template <bool STAGED_RECIP = false>
__device__ __forceinline__ float stage_geo(const RasterArgs& a, int t, int g, GeoA* s_ga, GeoB* s_gb) {
  const float2 m = reinterpret_cast<const float2*>(a.mean)[g];
  const float4 c = reinterpret_cast<const float4*>(a.cov)[g];
  const float det = c.x * c.w - c.y * c.z;
  const float al = fminf(a.alpha[g], 0.99f);
  s_ga[t] = GeoA{m.x, m.y, c.x, c.y};
  s_gb[t] = GeoB{c.z, c.w, det, STAGED_RECIP ? staged_recip(det) : 1.0f / det};
  return al;
}

template <bool IDS_IN_LDS = false, bool STAGED_RECIP = false, typename GC = float2>
__device__ __forceinline__ void stage_geometry(const RasterArgs& a, const int32_t* __restrict__ idsb, int s0,
                                               int base, int cnt, int* s_id, GeoA* s_ga, GeoB* s_gb, GC* s_gc,
                                               int* s_row) {
  const int t = threadIdx.x;
  if (t < cnt) {
    const int g = IDS_IN_LDS ? s_id[t] : idsb[s0 + base + t];
    s_id[t] = g;
    if (s_row) s_row[t] = a.sh_index ? a.sh_index[g] : g;  // the entry's SH row
    const float al = stage_geo<STAGED_RECIP>(a, t, g, s_ga, s_gb);
    put_gc(s_gc, t, al, reject_radius(al));
  }
}

// Second half of a batch's staging: the SH rows of its entries, row_of(j)
// giving entry j's row index, once a barrier has published what row_of
// reads.  NEG_LOG2E: the coefficients are staged times -log2(e), so that a
// colour's dot product is the argument of sigmoid_scaled (fast colours only).
template <int SHN, int CC, bool NEG_LOG2E = false, typename RowOf>
__device__ __forceinline__ void stage_sh_rows_of(const RasterArgs& a, const float* __restrict__ shb, int cnt,
                                                 RowOf row_of, float* s_sh) {
  const int t = threadIdx.x;
  if constexpr (SHN % 4 == 0) {
    // Rows of SHN contiguous floats (the stored order is the rendered one):
    // 16-byte pieces, SHN / 4 per entry.  Block-uniform test.
    if (a.sh_sc == CC && a.sh_sg == SHN && (reinterpret_cast<uintptr_t>(shb) & 15) == 0) {
      constexpr int Q = SHN / 4;
      const float4* sh4 = reinterpret_cast<const float4*>(shb);
      float4* d4 = reinterpret_cast<float4*>(s_sh);
      for (int e = t; e < cnt * Q; e += kTilePixels) {
        const int j = e / Q;
        const int64_t row = row_of(j);
        float4 v = sh4[row * Q + (e - j * Q)];
        if constexpr (NEG_LOG2E) {
          v.x *= kNegLog2e;
          v.y *= kNegLog2e;
          v.z *= kNegLog2e;
          v.w *= kNegLog2e;
        }
        d4[e] = v;
      }
      return;
    }
  }
  for (int e = t; e < cnt * SHN; e += kTilePixels) {
    const int j = e / SHN;
    const int r = e - j * SHN;
    const int ch = r / CC;
    const int k = r - ch * CC;
    const int64_t row = row_of(j);
    const float v = shb[row * a.sh_sg + ch * a.sh_sc + k];
    s_sh[e] = NEG_LOG2E ? v * kNegLog2e : v;
  }
}

// stage_sh_rows_of with the rows recorded in s_row by stage_geometry, or
// looked up from the ids in s_id (the same loop written out: through the
// row functor the forward compiles to a slower schedule).
template <int SHN, int CC, bool NEG_LOG2E = false>
__device__ __forceinline__ void stage_sh_rows(const RasterArgs& a, const float* __restrict__ shb, int cnt,
                                              const int* s_id, float* s_sh, const int* s_row) {
  const int t = threadIdx.x;
  if constexpr (SHN % 4 == 0) {
    if (a.sh_sc == CC && a.sh_sg == SHN && (reinterpret_cast<uintptr_t>(shb) & 15) == 0) {
      constexpr int Q = SHN / 4;
      const float4* sh4 = reinterpret_cast<const float4*>(shb);
      float4* d4 = reinterpret_cast<float4*>(s_sh);
      for (int e = t; e < cnt * Q; e += kTilePixels) {
        const int j = e / Q;
        int64_t row;
        if (s_row) {
          row = s_row[j];  // recorded with the geometry, before the barrier
        } else {
          const int g = s_id[j];
          row = a.sh_index ? (int64_t)a.sh_index[g] : (int64_t)g;
        }
        float4 v = sh4[row * Q + (e - j * Q)];
        if constexpr (NEG_LOG2E) {
          v.x *= kNegLog2e;
          v.y *= kNegLog2e;
          v.z *= kNegLog2e;
          v.w *= kNegLog2e;
        }
        d4[e] = v;
      }
      return;
    }
  }
  for (int e = t; e < cnt * SHN; e += kTilePixels) {
    const int j = e / SHN;
    const int r = e - j * SHN;
    const int ch = r / CC;
    const int k = r - ch * CC;
    const int g = s_id[j];
    const int64_t row = a.sh_index ? (int64_t)a.sh_index[g] : (int64_t)g;
    const float v = shb[row * a.sh_sg + ch * a.sh_sc + k];
    s_sh[e] = NEG_LOG2E ? v * kNegLog2e : v;
  }
}


// Stage entries [s0+base, s0+base+cnt) of the tile list into LDS (without
// WITH_SH only the geometry and opacity).  NEG_LOG2E: the SH coefficients are
// staged times -log2(e), so that a colour's dot product is the argument of
// sigmoid_scaled (fast colours only).  STAGED_RECIP as stage_geometry.
template <int SHN, int CC, bool IDS_IN_LDS = false, bool WITH_SH = true, bool NEG_LOG2E = false,
          bool STAGED_RECIP = false, typename GC = float2>
__device__ __forceinline__ void stage_batch_from(const RasterArgs& a, const float* __restrict__ shb,
                                                 const int32_t* __restrict__ idsb, int s0, int base, int cnt,
                                                 int* s_id, GeoA* s_ga, GeoB* s_gb, GC* s_gc, float* s_sh,
                                                 int* s_row = nullptr) {
  stage_geometry<IDS_IN_LDS, STAGED_RECIP>(a, idsb, s0, base, cnt, s_id, s_ga, s_gb, s_gc, s_row);
  if constexpr (!WITH_SH) return;
  __syncthreads();
  stage_sh_rows<SHN, CC, NEG_LOG2E>(a, shb, cnt, s_id, s_sh, s_row);
}

template <int SHN, int CC, bool IDS_IN_LDS = false, bool WITH_SH = true, typename GC = float2>
__device__ __forceinline__ void stage_batch(const RasterArgs& a, int s0, int base, int cnt, int* s_id, GeoA* s_ga,
                                            GeoB* s_gb, GC* s_gc, float* s_sh) {
  stage_batch_from<SHN, CC, IDS_IN_LDS, WITH_SH>(a, a.sh, a.ids, s0, base, cnt, s_id, s_ga, s_gb, s_gc, s_sh);
}

// Two-level lists: the next (up to kBatch) entries of the block-of-tiles list
// [cur, e0) whose tile bounds hold tile (tx, ty), in list order, into
// s_id[0..n).  Each round tests 256 candidates and appends the hits, ranked
// by ballot, to the ring s_q (head qh, qn queued); rounds run until a batch
// is queued or the list ends, and the batch is taken from the head.  A
// depth-ordered list filtered this way is exactly the tile's own list.  cur,
// qh, qn and the returned n are block-uniform and live in registers; s_wh
// alternates between two halves so one barrier per round suffices.
constexpr int kFilterQueue = 512;  // >= kBatch - 1 + kTilePixels queued entries
static_assert((kFilterQueue & (kFilterQueue - 1)) == 0 && kFilterQueue >= kBatch - 1 + kTilePixels, "ring size");

__device__ __forceinline__ int filter_batch(const RasterArgs& a, int e0, int tx, int ty, int* s_id, int* s_q,
                                            int* s_wh, int& cur, int& qh, int& qn, int& par) {
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  while (qn < kBatch && cur < e0) {
    const int c = cur + t;
    bool hit = false;
    int id = 0;
    if (c < e0) {
      id = a.sids[c];
      const int2 lo = reinterpret_cast<const int2*>(a.rtl)[id];
      const int2 hi = reinterpret_cast<const int2*>(a.rbr)[id];
      hit = lo.x <= tx && tx <= hi.x && lo.y <= ty && ty <= hi.y;
    }
    const uint64_t bal = __ballot(hit);
    int* wh = s_wh + 4 * par;
    if (lane == 0) wh[w] = __popcll(bal);
    __syncthreads();
    const int below =
        (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
    int pre = 0;
    for (int v = 0; v < w; ++v) pre += wh[v];
    const int total = wh[0] + wh[1] + wh[2] + wh[3];
    if (hit) s_q[(qh + qn + pre + below) & (kFilterQueue - 1)] = id;
    qn += total;
    cur += kTilePixels;
    par ^= 1;
  }
  __syncthreads();  // the rounds' queue writes
  const int n = min(kBatch, qn);
  if (t < n) s_id[t] = s_q[(qh + t) & (kFilterQueue - 1)];
  qh = (qh + n) & (kFilterQueue - 1);
  qn -= n;
  return n;
}

// ---------------------------------------------------------------------------
// Forward, SH colour.  vol_render_sh.h:171-248 (+ vol_render_bg.h:12-100).
// ---------------------------------------------------------------------------
// TWO: the tile's list is filtered on the fly from its block's list (see
// filter_batch) and recorded for the backward; single view, no background.
// PV: per-view SH rows, output image, list entries and recorded lists
// (view_sh / view_out / view_ids / view_fids; each null falls back to the
// shared array); without it the shared arrays, and no registers for them.
template <int C, bool BG, bool EXACT, bool TWO, bool PV>
__device__ __forceinline__ void fwd_sh_tile(const RasterArgs& a, TileCtx& tc) {
  constexpr int CC = C * C;
  constexpr int SHN = 3 * CC;
  static_assert(!(TWO && BG), "two-level lists without background");
  __shared__ GeoA s_ga[kBatch];
  __shared__ GeoB s_gb[kBatch];
  __shared__ float4 s_gc[kBatch];  // {opacity, rejection radius, -, -}: the stride of s_ga / s_gb
  __shared__ int s_id[kBatch];
  __shared__ __align__(16) float s_sh[kBatch * SHN];
  __shared__ int s_row[kBatch];  // SH row of each batch entry
  __shared__ int s_fq[TWO ? kFilterQueue : 1];
  __shared__ int s_fwh[8];

  const int tile = tc.tile;
  const bool inside = tc.inside;
  const int64_t pix = tc.pix;
  // the view's SH rows, output image and lists (PV), else the shared ones
  const float* shb = (PV && a.view_sh) ? a.view_sh[tc.view] : a.sh;
  float* outb = (PV && a.view_out) ? a.view_out[tc.view] : a.out;
  const int64_t opix = (PV && a.view_out) ? tc.lpix : pix;
  const int32_t* idsb = (PV && a.view_ids) ? a.view_ids[tc.view] : a.ids;
  int32_t* fidsb = (PV && a.view_fids) ? a.view_fids[tc.view] : a.fids;  // TWO: where the tile's list is recorded
  const int tx = tc.gx / kTile, ty = tc.gy / kTile;  // the view's own tile coordinates (TWO only)
  int s0, e0;
  if constexpr (TWO) {
    int sb;
    if (a.views) {
      const gsmi_view& bv = a.bviews[tc.view];
      sb = bv.tile_base + (ty >> a.super_shift) * bv.n_tiles_w + (tx >> a.super_shift);
    } else {
      sb = (ty >> a.super_shift) * a.n_super_w + (tx >> a.super_shift);
    }
    s0 = a.sstart[sb];
    e0 = a.send[sb];
  } else {
    s0 = a.start[tile];
    e0 = a.end[tile];
  }
  if (s0 == -1 || e0 - s0 == 0) {
    if constexpr (TWO) {
      if (a.fend && threadIdx.x == 0) a.fend[tile] = a.foff[tile];
    }
    if (inside) {
      if (BG && s0 == -1) {
        outb[3 * opix + 0] = a.bg[0];
        outb[3 * opix + 1] = a.bg[1];
        outb[3 * opix + 2] = a.bg[2];
      } else if (a.write_empty) {
        outb[3 * opix + 0] = 0.0f;
        outb[3 * opix + 1] = 0.0f;
        outb[3 * opix + 2] = 0.0f;
      }
      if (a.T_final) a.T_final[pix] = 1.0f;
      if (a.n_done) a.n_done[pix] = 0;
      if (BG && a.pre_bg) a.pre_bg[3 * pix] = a.pre_bg[3 * pix + 1] = a.pre_bg[3 * pix + 2] = 0.0f;
    }
    return;
  }
  const float px = tc.px;
  const float py = tc.py;
  float cw[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) cw[k] = tc.cw[k];
  float dir[3];
  pixel_direction(cw, px, py, dir);
  float Y[CC];
  sh_basis<C>(dir[0], dir[1], dir[2], Y);

  const float thresh = a.thresh;
  float T = 1.0f, o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
  bool done = !inside;
  const int n = e0 - s0;      // TWO: the block's list; the tile's length is known at the end
  int ndone = TWO ? -1 : n;
  if (inside && T < thresh) {
    done = true;
    ndone = 0;
  }
  static_assert(kBatch <= 64, "one batch entry per lane in the relevance mask");
  const float4 wr = wave_pixel_rect(px, py);
  const int lane = threadIdx.x & 63;
  // A lane is done when outside the image or once T < thresh (T stays put
  // from then on), so the wave is done when these two ballots of single
  // comparisons cover it; a ballot of `done` itself would go through a
  // vector register each entry.
  const uint64_t outside = __builtin_amdgcn_ballot_w64(!inside);
  int fcur = s0, fqh = 0, fqn = 0, fpar = 0;  // filter state (TWO), block-uniform
  int base = 0;
  for (;;) {
    int cnt;
    __syncthreads();
    if constexpr (TWO) {
      cnt = filter_batch(a, e0, tx, ty, s_id, s_fq, s_fwh, fcur, fqh, fqn, fpar);
      if (cnt == 0) break;
      if (fidsb && (int)threadIdx.x < cnt) fidsb[a.foff[tile] + base + threadIdx.x] = s_id[threadIdx.x];
      stage_batch_from<SHN, CC, true, true, !EXACT, true>(a, shb, idsb, s0, base, cnt, s_id, s_ga, s_gb, s_gc,
                                                           s_sh, s_row);
    } else {
      if (base >= n) break;
      cnt = min(kBatch, n - base);
      stage_batch_from<SHN, CC, false, true, !EXACT, true>(a, shb, idsb, s0, base, cnt, s_id, s_ga, s_gb, s_gc,
                                                            s_sh, s_row);
    }
    __syncthreads();
    // entries of the batch that can reach a pixel of this wave, in order
    const int le = lane < cnt ? lane : 0;
    uint64_t rel = __ballot(lane < cnt && may_reach(s_ga[le], s_gb[le], s_gc[le].y, wr.x, wr.y, wr.z, wr.w));
    // entries examined when entry j saturates a pixel: base + j + 1, with
    // base (block-uniform) in a scalar register, so that the per-entry
    // update is one select
    const int nd1 = __builtin_amdgcn_readfirstlane(base) + 1;
    while (rel) {
      const int j = __builtin_ctzll(rel);
      rel &= rel - 1;
      if ((__builtin_amdgcn_ballot_w64(T < thresh) | outside) == ~0ull) break;
      if (done) continue;
      const GeoA ga = s_ga[j];
      const GeoB gb = s_gb[j];
      const float4 gc = s_gc[j];
      // det and alpha read with the rest of the record: deferred into the
      // branch, their reads each cost an address move from the scalar index
      asm volatile("" ::"v"(gb.det), "v"(gc.x));
      const float dx = px - ga.mx, dy = py - ga.my;
      const float tmpx = dx * gb.c3 - dy * gb.c2;
      const float tmpy = -dx * ga.c1 + dy * ga.c0;
      const float q = tmpx * dx + tmpy * dy;
      const float rq = q * gb.inv_det;
      if (rq > gc.y) continue;
      const float G = gauss_staged(q, gb.det, gb.inv_det, rq);
      const float al = gc.x;
      float coeff = al * T;
      coeff *= G;
      if (al * G < kMinRenderAlpha) continue;
      if (isnan(coeff)) coeff = 0.0f;
      const float* shj = s_sh + j * SHN;
      float y0, y1, y2;
      if constexpr (EXACT) {
        y0 = sigmoid_exact(dot_exact<CC>(shj, Y));
        y1 = sigmoid_exact(dot_exact<CC>(shj + CC, Y));
        y2 = sigmoid_exact(dot_exact<CC>(shj + 2 * CC, Y));
      } else {
        y0 = sigmoid_scaled(dot_fast<CC>(shj, Y));  // coefficients staged times -log2(e)
        y1 = sigmoid_scaled(dot_fast<CC>(shj + CC, Y));
        y2 = sigmoid_scaled(dot_fast<CC>(shj + 2 * CC, Y));
      }
      if constexpr (EXACT) {
        if (isnan(y0 * coeff)) y0 = 0.0f;
        if (isnan(y1 * coeff)) y1 = 0.0f;
        if (isnan(y2 * coeff)) y2 = 0.0f;
        o0 = o0 + coeff * y0;
        o1 = o1 + coeff * y1;
        o2 = o2 + coeff * y2;
      } else {
        // coeff is finite here and a sigmoid is NaN or in [0, 1], so
        // isnan(y * coeff) == isnan(y), and max(y, 0) maps NaN to 0.
        o0 = fmaf(coeff, fmaxf(y0, 0.0f), o0);
        o1 = fmaf(coeff, fmaxf(y1, 0.0f), o1);
        o2 = fmaf(coeff, fmaxf(y2, 0.0f), o2);
      }
      T = T * (1.0f - al * G);
      if (T < thresh) {
        done = true;
        ndone = nd1 + j;
      }
    }
    base += cnt;
    if (!__syncthreads_or(!done)) break;
  }
  if constexpr (TWO) {
    if (ndone < 0) ndone = base;  // list ran out first: every filtered entry examined
    if (a.fend && threadIdx.x == 0) a.fend[tile] = a.foff[tile] + base;
  }
  if (!inside) return;
  if (BG) {
    if (a.pre_bg) {
      a.pre_bg[3 * pix + 0] = o0;
      a.pre_bg[3 * pix + 1] = o1;
      a.pre_bg[3 * pix + 2] = o2;
    }
    if (T > thresh) {
      const float omt = 1.0f - T;
      o0 = o0 * T + a.bg[0] * omt;
      o1 = o1 * T + a.bg[1] * omt;
      o2 = o2 * T + a.bg[2] * omt;
    }
  }
  outb[3 * opix + 0] = o0;
  outb[3 * opix + 1] = o1;
  outb[3 * opix + 2] = o2;
  if (a.T_final) a.T_final[pix] = T;
  if (a.n_done) a.n_done[pix] = ndone;
}

// Next tile of a work-queue launch: the band of this workgroup's XCD first
// (consecutive workgroups land on consecutive XCDs, and a band is a run of
// neighbouring tiles, which share Gaussians in that XCD's L2), then the other
// bands in turn; -1 once every band is taken.  Block-uniform.
__device__ __forceinline__ int next_queued_tile(const RasterArgs& a, int& band, int* s_item) {
  const int per = (a.n_tiles_total + 7) >> 3;
  for (;;) {
    __syncthreads();  // everyone has read the previous item
    if (threadIdx.x == 0) {
      const int b = (blockIdx.x + band) & 7;
      const int k = atomicAdd(a.work + b, 1);
      *s_item = (k < per && b * per + k < a.n_tiles_total) ? b * per + k : -1;
    }
    __syncthreads();
    const int item = *s_item;
    if (item >= 0) return item;
    if (++band == 8) return -1;
  }
}

template <int C, bool BG, bool EXACT, bool TWO = false, bool PV = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_fwd_sh(RasterArgs a) {
  TileCtx tc;
  if (tile_setup(a, tc)) fwd_sh_tile<C, BG, EXACT, TWO, PV>(a, tc);
}

// The work-queue form (two-level lists, per-view arrays): a persistent
// launch whose workgroups composite tiles until the queues are empty.
template <int C, bool EXACT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_fwd_sh_queued(RasterArgs a) {
  __shared__ int s_item;
  TileCtx tc;
  int band = 0;
  for (;;) {
    const int tile = next_queued_tile(a, band, &s_item);
    if (tile < 0) return;
    if (tile_setup_at(a, tc, tile)) fwd_sh_tile<C, false, EXACT, true, true>(a, tc);
  }
}

__device__ __forceinline__ void stage_rgb(const RasterArgs& a, int s0, int base, int cnt, int* s_id, GeoA* s_ga,
                                          GeoB* s_gb, float2* s_gc, float* s_col);

// ---------------------------------------------------------------------------
// Forward state recomputed inside a backward launch, for the backward entry
// points that take only the reference's arguments (render.h:78-84, 95-101,
// 126-133: no saved transmittance).  Each thread walks its pixel's list front
// to back with the forward's own decisions (k_fwd_sh, k_fwd_rgb: the same
// rejection test, exact Gaussian value, 1/255 skip and transmittance break),
// so it ends with the bits the forward stores as T_final and n_done; with
// COLOUR (background variant) also the colour before the blend, in the
// forward's fast arithmetic.  Pixels outside the image get T = 1, nd = 0, as
// a backward reads them.  Batches of NB entries go through the calling
// kernel's staging arrays.  The reference recomputes the same front-to-back
// state in its backward (vol_render_sh.h:268-351).
// ---------------------------------------------------------------------------
template <int NB, int C, bool COLOUR, bool RGB, typename GC = float2>
__device__ __forceinline__ void forward_state(const RasterArgs& a, int s0, int n, const float* Y, float px,
                                              float py, bool inside, int* s_id, GeoA* s_ga, GeoB* s_gb,
                                              GC* s_gc, float* s_sh, float& T, int& nd, float& p0, float& p1,
                                              float& p2) {
  constexpr int CC = C * C;
  constexpr int SHN = 3 * CC;
  static_assert(NB <= 64, "one batch entry per lane in the relevance mask");
  const float thresh = a.thresh;
  T = 1.0f;
  p0 = p1 = p2 = 0.0f;
  bool done = !inside;
  nd = n;
  if (inside && T < thresh) {
    done = true;
    nd = 0;
  }
  const float4 wr = wave_pixel_rect(px, py);
  const int lane = threadIdx.x & 63;
  for (int base = 0; base < n; base += NB) {
    const int cnt = min(NB, n - base);
    __syncthreads();
    if constexpr (RGB)
      stage_rgb(a, s0, base, cnt, s_id, s_ga, s_gb, s_gc, s_sh);
    else
      stage_batch<SHN, CC, false, COLOUR>(a, s0, base, cnt, s_id, s_ga, s_gb, s_gc, s_sh);
    __syncthreads();
    const int le = lane < cnt ? lane : 0;
    uint64_t rel;
    if constexpr (RGB)
      rel = __ballot(lane < cnt);  // the RGB forward walks every entry
    else
      rel = __ballot(lane < cnt && may_reach(s_ga[le], s_gb[le], s_gc[le].y, wr.x, wr.y, wr.z, wr.w));
    while (rel) {
      const int j = __builtin_ctzll(rel);
      rel &= rel - 1;
      if (__all(done)) break;
      if (done) continue;
      const GeoA ga = s_ga[j];
      const GeoB gb = s_gb[j];
      const auto gc = s_gc[j];
      const float dx = px - ga.mx, dy = py - ga.my;
      const float tmpx = dx * gb.c3 - dy * gb.c2;
      const float tmpy = -dx * ga.c1 + dy * ga.c0;
      const float q = tmpx * dx + tmpy * dy;
      if (q * gb.inv_det > gc.y) continue;
      const float G = RGB ? gauss_exact_d(dx, dy, ga.c0, ga.c1, gb.c2, gb.c3) : gauss_exact(q, gb.det);
      const float al = gc.x;
      float coeff = al * T;
      coeff *= G;
      if (al * G < kMinRenderAlpha) continue;
      if constexpr (COLOUR && !RGB) {
        if (isnan(coeff)) coeff = 0.0f;
        const float* shj = s_sh + j * SHN;
        p0 = fmaf(coeff, fmaxf(sigmoid_fast(dot_fast<CC>(shj, Y)), 0.0f), p0);
        p1 = fmaf(coeff, fmaxf(sigmoid_fast(dot_fast<CC>(shj + CC, Y)), 0.0f), p1);
        p2 = fmaf(coeff, fmaxf(sigmoid_fast(dot_fast<CC>(shj + 2 * CC, Y)), 0.0f), p2);
      }
      T = T * (1.0f - al * G);
      if (T < thresh) {
        done = true;
        nd = base + j + 1;
      }
    }
    if (!__syncthreads_or(!done)) break;
  }
  if (!inside) {
    T = 1.0f;
    nd = 0;
  }
}

// ---------------------------------------------------------------------------
// Wave64 reduce-scatter: NV values per lane in, one column sum per lane out.
// After log2(NV) halving rounds (partner = lane ^ m, m = 32, 16, ...) each
// group of 64/NV lanes holds the partial sum of component lane >> (6 -
// log2 NV); the remaining rounds finish the sum.  63 shuffles for NV = 64.
// ---------------------------------------------------------------------------
template <int H, int M, int NV>
__device__ __forceinline__ void reduce_scatter_round(float (&v)[NV], int lane) {
  if constexpr (H >= 1) {
    const bool upper = (lane & M) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const float keep = upper ? v[i + H] : v[i];
      const float send = upper ? v[i] : v[i + H];
      v[i] = keep + __shfl_xor(send, M);
    }
    reduce_scatter_round<H / 2, M / 2, NV>(v, lane);
  }
}

template <int NV>
__device__ __forceinline__ float wave_reduce_scatter(float (&v)[NV]) {
  const int lane = lane_id();
  reduce_scatter_round<NV / 2, 32, NV>(v, lane);
  float s = v[0];
#pragma unroll
  for (int m = 32 / NV; m >= 1; m >>= 1) s += __shfl_xor(s, m);
  return s;
}

// Same result as wave_reduce_scatter<8> (lane L returns the wave sum of
// v[L >> 3]) without the LDS crossbar: the exchanges across 32 and 16 lanes
// are gfx950 permlane swaps (v_permlane32_swap: upper half of the first
// operand <-> lower half of the second; v_permlane16_swap: odd 16-lane rows of
// the first <-> even rows of the second), the one across 8 lanes a DPP row
// rotation by 8 (= lane ^ 8 inside a row), and the final sum over 8 lanes two
// DPP quad permutations and a row rotation by 12 (row_ror:n makes lane i read
// lane (i - n) mod 16 of its row).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xf, 0xf, false));
}

// The swaps are issued as inline VALU assembly: the compiler's lowering of
// __builtin_amdgcn_permlane{16,32}_swap inside this kernel added the swapped
// register to itself instead of to its partner (v_add v18, v55, v55 after
// v_permlane32_swap v55, v60), which corrupted the sums.  One wait state
// before each swap covers the VALU-write -> swap-read dependency.
__device__ __forceinline__ void permlane32_swap(float& x, float& y) {
  asm volatile("s_nop 1\n\tv_permlane32_swap_b32 %0, %1" : "+v"(x), "+v"(y));
}
__device__ __forceinline__ void permlane16_swap(float& x, float& y) {
  asm volatile("s_nop 1\n\tv_permlane16_swap_b32 %0, %1" : "+v"(x), "+v"(y));
}

__device__ __forceinline__ float wave_reduce_scatter8_swap(float (&v)[8], int lane) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {  // across lane ^ 32: component i (low half) / i + 4 (high half)
    float x = v[i], y = v[i + 4];
    permlane32_swap(x, y);
    v[i] = x + y;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {  // across lane ^ 16: + 2 on odd rows
    float x = v[i], y = v[i + 2];
    permlane16_swap(x, y);
    v[i] = x + y;
  }
  const bool upper = (lane & 8) != 0;  // across lane ^ 8: + 1 on the upper half row
  const float keep = upper ? v[1] : v[0];
  const float send = upper ? v[0] : v[1];
  float s = keep + dpp_f<0x128>(send);  // row_ror:8
  s = s + dpp_f<0xB1>(s);               // quad_perm [1,0,3,2]
  s = s + dpp_f<0x4E>(s);               // quad_perm [2,3,0,1]
  s = s + dpp_f<0x12C>(s);              // row_ror:12 (lane i reads (i + 4) mod 16): lanes 0, 8 hold the total
  return s;
}

constexpr int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Gradient component layout per Gaussian in the LDS accumulator:
// 0,1 mean; 2 cov00; 3 cov01 (= cov10); 4 cov11; 5 alpha; 6.. colour.
template <int NCOL>
__device__ __forceinline__ void emit_grad(const RasterArgs& a, int g, int c, float v) {
  if (c < 2) {
    unsafeAtomicAdd(a.grad_mean + 2 * (int64_t)g + c, v);
  } else if (c < 5) {
    if (c == 2) unsafeAtomicAdd(a.grad_cov + 4 * (int64_t)g + 0, v);
    if (c == 3) {
      unsafeAtomicAdd(a.grad_cov + 4 * (int64_t)g + 1, v);
      unsafeAtomicAdd(a.grad_cov + 4 * (int64_t)g + 2, v);
    }
    if (c == 4) unsafeAtomicAdd(a.grad_cov + 4 * (int64_t)g + 3, v);
  } else if (c == 5) {
    unsafeAtomicAdd(a.grad_alpha + g, v);
  } else {
    const int r = c - 6;
    if (a.sh_sc == 0) {  // RGB colour [V,3]
      unsafeAtomicAdd(a.grad_sh + 3 * (int64_t)g + r, v);
    } else {
      const int CCl = (int)(NCOL / 3);
      const int ch = r / CCl;
      const int k = r - ch * CCl;
      const int64_t row = a.sh_index ? (int64_t)a.sh_index[g] : (int64_t)g;
      unsafeAtomicAdd(a.grad_sh + row * a.sh_sg + ch * a.sh_sc + k, v);
    }
  }
}

// Adds a batch's per-entry gradient sums (s_acc[entry][6 + NCOL]) to the
// global gradients.
template <int NCOL>
__device__ __forceinline__ void flush_grads(const RasterArgs& a, const int* s_id, float* s_acc, int cnt) {
  constexpr int NCOMP = 6 + NCOL;
  const int t = threadIdx.x;
  for (int e = t; e < cnt * NCOMP; e += kTilePixels) {
    const float v = s_acc[e];
    if (v == 0.0f) continue;
    const int j = e / NCOMP;
    emit_grad<NCOL>(a, s_id[j], e - j * NCOMP, v);
  }
}

// Block-wide max of a per-thread int (256 threads).
__device__ __forceinline__ int block_max(int v, int* s_tmp) {
  const int t = threadIdx.x;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m));
  if ((t & 63) == 0) s_tmp[t >> 6] = v;
  __syncthreads();
  const int r = max(max(s_tmp[0], s_tmp[1]), max(s_tmp[2], s_tmp[3]));
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// Backward, SH colour: reverse-order adjoint.  Same gradients as the
// reference's forward-order recompute (vol_render_sh.h:268-455):
//   with T_j the transmittance in front of composited Gaussian j, w = a T_j G,
//   S_j = sum over later composited k of w_k y_k (for bg, plus F - pre_bg),
//   P = sum_ch g_ch (y_ch T_j - S_j,ch / (1 - a G)),
//   grad_sh += w y (1-y) g_ch Y_k;  grad_mean += P a G (s, u);
//   grad_cov += 0.5 P a G (s s, s u, u s, u u);  grad_alpha += P G,
// s = (dx c3 - dy c2)/det, u = (-dx c1 + dy c0)/det.  T_j is recovered from
// the stored final transmittance as T_j = T_{j+1} / (1 - a G).
// ---------------------------------------------------------------------------
template <int C, bool BG, bool RECOMP = false>
__global__ __launch_bounds__(256) void k_bwd_sh(RasterArgs a) {
  constexpr int CC = C * C;
  constexpr int SHN = 3 * CC;
  constexpr int NCOMP = 6 + SHN;
  constexpr int NV = next_pow2(NCOMP);
  static_assert(NV <= 64, "gradient vector wider than a wave");
  __shared__ GeoA s_ga[kBatch];
  __shared__ GeoB s_gb[kBatch];
  __shared__ float2 s_gc[kBatch];
  __shared__ int s_id[kBatch];
  __shared__ __align__(16) float s_sh[kBatch * SHN];
  __shared__ float s_acc[kBatch * NCOMP];
  __shared__ int s_tmp[4];

  TileCtx tc;
  if (!tile_setup(a, tc)) return;
  const int tile = tc.tile;
  const int t = threadIdx.x;
  const bool inside = tc.inside;
  const int64_t pix = tc.pix;
  const int s0 = a.start[tile];
  const int e0 = a.end[tile];
  if (s0 == -1 || e0 - s0 == 0) return;
  const int32_t* idsb = a.view_ids ? a.view_ids[tc.view] : a.ids;  // the tile's list entries

  const float px = tc.px;
  const float py = tc.py;
  float cw[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) cw[k] = tc.cw[k];
  float dir[3];
  pixel_direction(cw, px, py, dir);
  float Y[CC];
  sh_basis<C>(dir[0], dir[1], dir[2], Y);

  float T = 1.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, S0 = 0.0f, S1 = 0.0f, S2 = 0.0f;
  int nd = 0;
  float pre0 = 0.0f, pre1 = 0.0f, pre2 = 0.0f;
  if constexpr (RECOMP)
    forward_state<kBatch, C, BG, false>(a, s0, e0 - s0, Y, px, py, inside, s_id, s_ga, s_gb, s_gc, s_sh, T, nd,
                                        pre0, pre1, pre2);
  if (inside) {
    if constexpr (!RECOMP) {
      T = a.T_saved[pix];
      nd = a.n_done_saved[pix];
      if (BG) {
        pre0 = a.pre_bg_saved[3 * pix + 0];
        pre1 = a.pre_bg_saved[3 * pix + 1];
        pre2 = a.pre_bg_saved[3 * pix + 2];
      }
    }
    g0 = a.grad_out[3 * pix + 0];
    g1 = a.grad_out[3 * pix + 1];
    g2 = a.grad_out[3 * pix + 2];
    if (BG) {
      S0 = a.out_saved[3 * pix + 0] - pre0;
      S1 = a.out_saved[3 * pix + 1] - pre1;
      S2 = a.out_saved[3 * pix + 2] - pre2;
    }
  }
  const int nmax = block_max(nd, s_tmp);
  const int lane = lane_id();
  constexpr int kLaneShift = 6 - __builtin_ctz(NV);
  const int my_comp = lane >> kLaneShift;
  const bool writer = (lane & ((1 << kLaneShift) - 1)) == 0 && my_comp < NCOMP;

  for (int base = ((nmax - 1) / kBatch) * kBatch; nmax > 0 && base >= 0; base -= kBatch) {
    const int cnt = min(kBatch, nmax - base);
    __syncthreads();
    for (int e = t; e < cnt * NCOMP; e += kTilePixels) s_acc[e] = 0.0f;
    stage_batch_from<SHN, CC>(a, a.sh, idsb, s0, base, cnt, s_id, s_ga, s_gb, s_gc, s_sh);
    __syncthreads();
    for (int j = cnt - 1; j >= 0; --j) {
      const int idx = base + j;
      if (!wave_any(idx < nd)) continue;
      float cc0 = 0.0f, cc1 = 0.0f, cc2 = 0.0f, Ds = 0.0f, Du = 0.0f, Dss = 0.0f, Dsu = 0.0f, Duu = 0.0f,
            PG = 0.0f;
      bool contrib = false;
      if (idx < nd) {
        const GeoA ga = s_ga[j];
        const GeoB gb = s_gb[j];
        const float2 gc = s_gc[j];
        const float dx = px - ga.mx, dy = py - ga.my;
        const float tmpx = dx * gb.c3 - dy * gb.c2;
        const float tmpy = -dx * ga.c1 + dy * ga.c0;
        const float q = tmpx * dx + tmpy * dy;
        if (!(q * gb.inv_det > gc.y)) {
          const float G = gauss_exact(q, gb.det);
          const float al = gc.x;
          const float aG = al * G;
          if (!(aG < kMinRenderAlpha)) {
            contrib = true;
            const float om = 1.0f - aG;
            const float Tj = T / om;
            float w = al * Tj * G;
            if (isnan(w)) w = 0.0f;
            const float* shj = s_sh + j * SHN;
            // w is finite, a sigmoid NaN or in [0, 1]: isnan(y * w) == isnan(y)
            const float y0 = fmaxf(sigmoid_fast(dot_fast<CC>(shj, Y)), 0.0f);
            const float y1 = fmaxf(sigmoid_fast(dot_fast<CC>(shj + CC, Y)), 0.0f);
            const float y2 = fmaxf(sigmoid_fast(dot_fast<CC>(shj + 2 * CC, Y)), 0.0f);
            cc0 = w * (y0 * (1.0f - y0)) * g0;
            cc1 = w * (y1 * (1.0f - y1)) * g1;
            cc2 = w * (y2 * (1.0f - y2)) * g2;
            const float inv_om = 1.0f / om;
            float P = g0 * (y0 * Tj - S0 * inv_om);
            P = fmaf(g1, y1 * Tj - S1 * inv_om, P);
            P = fmaf(g2, y2 * Tj - S2 * inv_om, P);
            S0 = fmaf(w, y0, S0);
            S1 = fmaf(w, y1, S1);
            S2 = fmaf(w, y2, S2);
            const float D = P * aG;
            const float inv_det = 1.0f / gb.det;
            const float s = tmpx * inv_det;
            const float u = tmpy * inv_det;
            Ds = D * s;
            Du = D * u;
            Dss = 0.5f * Ds * s;
            Dsu = 0.5f * Ds * u;
            Duu = 0.5f * Du * u;
            PG = P * G;
            T = Tj;
          }
        }
      }
      if (!wave_any(contrib)) continue;
      float v[NV];
      v[0] = Ds;
      v[1] = Du;
      v[2] = Dss;
      v[3] = Dsu;
      v[4] = Duu;
      v[5] = PG;
#pragma unroll
      for (int k = 0; k < CC; ++k) {
        v[6 + k] = cc0 * Y[k];
        v[6 + CC + k] = cc1 * Y[k];
        v[6 + 2 * CC + k] = cc2 * Y[k];
      }
#pragma unroll
      for (int k = NCOMP; k < NV; ++k) v[k] = 0.0f;
      const float r = wave_reduce_scatter<NV>(v);
      if (writer && r != 0.0f) atomicAdd(s_acc + j * NCOMP + my_comp, r);
    }
    __syncthreads();
    flush_grads<SHN>(a, s_id, s_acc, cnt);
  }
}

// ---------------------------------------------------------------------------
// Backward for SH orders 3 and 4 with the SH-gradient reduction on the matrix
// cores.  Per wave, the per-pixel colour coefficients
// cc[ch](p) = w y (1-y) g_ch of the list entries that have a contributing
// pixel are staged in LDS, 5 entries at a time, as a 15 x 64 matrix
// (row = ch * 5 + entry slot, columns = the wave's 64 pixels), and
//   grad_sh[e][ch][k] += sum_p cc[ch,e](p) * Y_k(p)
// is one 16x64 by 64x16 product (row 15 unused) with fp32 accumulation.
//
// Default: fp32 products, as the reference forms them (vol_render_sh.h:28-36,
// :328-333): 16 v_mfma_f32_16x16x4_f32 per chunk.  Operand layout: A[i][k]
// from lane i + 16k, B[k][j] from lane j + 16k, D[i][j] in lane j + 16 (i / 4),
// accumulator register i % 4.  K-step t takes pixel 16 g + t from lane group
// g in both operands, so the product sums every pixel once.  A staged row
// keeps its 64 floats in 16-byte blocks, block b at position b ^ (row & 15),
// so a lane's 16 A values are four 16-byte loads and the rows of one load
// fall on distinct banks; the Y operand is constant per wave (16 VGPRs).
//
// SPLIT (opt-in; not fp32 arithmetic): the same product on bf16 matrix cores
// as a three-term split.  Each fp32 operand x is split into hi = x with the
// low 16 bits cleared (exactly a bf16) and lo = x - hi (exact in fp32, then
// truncated to bf16), and A B ~= Ah Bh + Ah Bl + Al Bh, dropping Al Bl and
// the truncation of lo: relative error up to 2^-15 per product, biased toward
// zero.  6 v_mfma_f32_16x16x32_bf16 (16 cycles each) per chunk instead of 16
// fp32 steps of 32 cycles.  Layout of the 16x16x32 bf16 MFMA: lane i + 16g
// holds 8 values of row i of A (and of column i of B) for the same 8 K
// indices; both operands give lane group g pixels 32t + 8g .. 32t + 8g + 7
// for K-step t.  A staged row holds the 64 pixels in 8 blocks of 8 (bf16),
// block b at position b ^ (row & 7).
//
// Entries without a contributing pixel take no slot.  The per-entry sums of
// the SH components go to per-wave LDS accumulators with plain stores: an
// entry takes at most one slot per wave and batch, so each (entry,
// component) is written at most once per wave, and the four waves' values are
// added when the batch is flushed to the global gradients.  (LDS float
// atomics from every lane of every flush were the kernel's largest cost.)
// The six scalar components (mean, cov, alpha) use an 8-wide wave
// reduce-scatter and LDS atomics (6 lanes per entry).  RECOMP: the forward
// state comes from forward_state instead of saved T_final / n_done.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// List entries staged per LDS batch of the MFMA backward: 16 keeps LDS under
// 32 KB with the per-wave accumulators (5 workgroups per CU).
#ifndef GSMI_BWD_BATCH
#define GSMI_BWD_BATCH 16
#endif
constexpr int kMfmaBatch = GSMI_BWD_BATCH;
constexpr int kChunk = 5;        // entry slots per MFMA reduction (3 * 5 = 15 of 16 rows)
constexpr int kCcRows = 3 * kChunk;
constexpr int kCcStride = 64;    // bf16 per staged row (128 B), SPLIT
constexpr int kCcPlane = kCcRows * kCcStride;  // bf16 per plane (hi or lo), SPLIT
constexpr int kCcWords = kCcRows * 64;          // 32-bit words per wave: 15 fp32 rows, or the two bf16 planes
constexpr int kSlotBits = 6;     // batch positions of the chunk's slots, packed in one word
static_assert(kMfmaBatch <= (1 << kSlotBits) && kChunk * kSlotBits <= 32, "slot packing");
static_assert(2 * kCcPlane * 2 == kCcWords * 4, "the two bf16 planes take the fp32 rows' bytes");

// Position (bf16 units) of pixel p in staged bf16 row `row`: 8-pixel blocks XOR-swizzled by row.
__device__ __forceinline__ int cc_pos(int row, int p) {
  return row * kCcStride + (p ^ ((row & 7) << 3));  // p < 64: block p / 8 -> (p / 8) ^ (row & 7)
}

// Position (floats) of pixel p in the staged fp32 rows: pixel-major with the
// odd stride kCcRows, so a row written by the 64 lanes (15 p + row) and one
// pixel of 16 rows x 4 pixel groups read by the matrix-core operand
// (15 (16 g + k) + i) both fall in distinct banks, and every row of a lane
// is its base plus a constant.
__device__ __forceinline__ int cc_pos_f32(int row, int p) { return p * kCcRows + row; }
static_assert(kCcRows % 2 == 1 && 64 * kCcRows <= kCcWords, "fp32 staging layout");

// Upper 16 bits of x (x truncated to bf16) and of the remainder x - hi.
__device__ __forceinline__ void split_bf16(float x, uint32_t& hi, uint32_t& lo) {
  const uint32_t b = __float_as_uint(x);
  hi = b >> 16;
  lo = __float_as_uint(x - __uint_as_float(b & 0xffff0000u)) >> 16;
}

// CMF: the colours (SH dot products) of up to kChunk relevant entries at a
// time are one 16x16 (rows ch * kChunk + slot) by 16x64 product on fp32
// matrix cores, written into the slots' staged rows, where each entry reads
// its three dot products just before it overwrites them with its
// coefficients; the SH-gradient product then runs on the same rows.  A chunk
// is then kChunk relevant entries (contributing or not) instead of kChunk
// contributing ones.  fp32 products, summed in another order than the VALU
// dot product.
template <int C, bool BG, bool SPLIT, bool RECOMP, bool CMF>
__device__ __forceinline__ void bwd_sh_mfma_body(const RasterArgs& a) {
  constexpr int CC = C * C;
  constexpr int SHN = 3 * CC;
  constexpr int NCOMP = 6 + SHN;
    constexpr int KS = (CC + 3) / 4;  // CMF: K-steps of 4 basis functions
  static_assert(CC <= 16, "SH basis wider than the MFMA tile");
  static_assert(!(CMF && SPLIT), "matrix-core colours with the fp32 reduction only");
  __shared__ GeoA s_ga[kMfmaBatch];
  __shared__ GeoB s_gb[kMfmaBatch];
  // {opacity, rejection radius, id, SH row} (the last two as int bits): the
  // stride of s_ga / s_gb, so one address serves an entry's three reads
  __shared__ float4 s_gc[kMfmaBatch];
  __shared__ __align__(16) float s_sh[kMfmaBatch * SHN];
  __shared__ float s_acc[kMfmaBatch * 6];                      // scalar components, shared
  __shared__ __align__(16) float s_shacc[4][kMfmaBatch * SHN];  // SH components, per wave
  __shared__ __align__(16) uint32_t s_cc[4][kCcWords];         // per wave: staged coefficient rows
  __shared__ int s_tmp[4];
  __shared__ uint32_t s_touch;          // batch entries with a contribution in some wave

  TileCtx tc;
  if (!tile_setup(a, tc)) return;
  const int tile = tc.tile;
  const int t = threadIdx.x;
  const int wv = t >> 6;
  const int lane = lane_id();
  const bool inside = tc.inside;
  const int64_t pix = tc.pix;
  const int s0 = a.start[tile];
  const int e0 = a.end[tile];
  if (s0 == -1 || e0 - s0 == 0) return;
  const int32_t* idsb = a.view_ids ? a.view_ids[tc.view] : a.ids;  // the tile's list entries

  const float px = tc.px;
  const float py = tc.py;
  float cw[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) cw[k] = tc.cw[k];
  float dir[3];
  pixel_direction(cw, px, py, dir);
  float Y[CC];
  sh_basis<C>(dir[0], dir[1], dir[2], Y);
  static_assert(kMfmaBatch <= 32, "one batch entry per lane in the relevance mask");
  const float4 wr = wave_pixel_rect(px, py);

  // gS = sum_c g_c S_c: the colour behind the entry (S_c, accumulated in
  // reverse order) enters the adjoint only through this grad_out-weighted sum.
  float T = 1.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, gS = 0.0f;
  int nd = 0;
  float pre0 = 0.0f, pre1 = 0.0f, pre2 = 0.0f;
  if constexpr (RECOMP)
    forward_state<kMfmaBatch, C, BG, false>(a, s0, e0 - s0, Y, px, py, inside, reinterpret_cast<int*>(s_acc), s_ga, s_gb, s_gc, s_sh, T,
                                            nd, pre0, pre1, pre2);
  if (inside) {
    if constexpr (!RECOMP) {
      T = a.T_saved[pix];
      nd = a.n_done_saved[pix];
      if (BG) {
        pre0 = a.pre_bg_saved[3 * pix + 0];
        pre1 = a.pre_bg_saved[3 * pix + 1];
        pre2 = a.pre_bg_saved[3 * pix + 2];
      }
    }
    g0 = a.grad_out[3 * pix + 0];
    g1 = a.grad_out[3 * pix + 1];
    g2 = a.grad_out[3 * pix + 2];
    if (BG) {
      const float S0 = a.out_saved[3 * pix + 0] - pre0;
      const float S1 = a.out_saved[3 * pix + 1] - pre1;
      const float S2 = a.out_saved[3 * pix + 2] - pre2;
      gS = fmaf(g2, S2, fmaf(g1, S1, g0 * S0));
    }
  }

  // B operand (the per-pixel SH basis), transposed through this wave's
  // staging area as fp32, 8 basis functions at a time.  FP32: bY[t] = Y_j at
  // pixel 16g + t for lane j + 16g.  SPLIT: lane j + 16g holds Y_j at pixels
  // 32t + 8g + e (e = 0..7) for K-step t, split into hi and lo bf16.
  static_assert(kCcWords >= 64 * 8, "transpose does not fit the staging area");
  uint32_t* cc_raw = s_cc[wv];
  float* tr = reinterpret_cast<float*>(cc_raw);
  u32x4 bh[2], bl[2];
  float bY[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) bY[k] = 0.0f;
  // CMF B operand: yB[4 s + b] = Y_{4 s + g}(pixel 16 b + j) for lane j + 16 g
  float yB[CMF ? 4 * KS : 1];
#pragma unroll
  for (int k = 0; k < (CMF ? 4 * KS : 1); ++k) yB[k] = 0.0f;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int k = 0; k < 8; ++k) tr[lane * 8 + k] = (8 * h + k) < CC ? Y[(8 * h + k) < CC ? 8 * h + k : 0] : 0.0f;
    __builtin_amdgcn_wave_barrier();
    const int col = (lane & 15) - 8 * h;
    if (col >= 0 && col < 8) {
      if constexpr (SPLIT) {
#pragma unroll
        for (int st = 0; st < 2; ++st) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int p = 32 * st + 8 * (lane >> 4) + 2 * q;
            uint32_t h0, l0, h1, l1;
            split_bf16(tr[p * 8 + col], h0, l0);
            split_bf16(tr[(p + 1) * 8 + col], h1, l1);
            bh[st][q] = h0 | (h1 << 16);
            bl[st][q] = l0 | (l1 << 16);
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) bY[k] = tr[(16 * (lane >> 4) + k) * 8 + col];
      }
    }
    if constexpr (CMF) {
#pragma unroll
      for (int st = 0; st < KS; ++st) {
        const int e = 4 * st + (lane >> 4) - 8 * h;  // basis function of this lane, within this half
        if (e >= 0 && e < 8 && 8 * h + e < CC) {
#pragma unroll
          for (int b = 0; b < 4; ++b) yB[4 * st + b] = tr[(16 * b + (lane & 15)) * 8 + e];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }

  const int nmax = block_max(nd, s_tmp);
  const int my_comp = lane >> 3;                      // scalar reduce-scatter (8 wide)
  const bool writer = (lane & 7) == 0 && my_comp < 6;
  // the covariance components (2..4) carry a factor 1/2, applied to the wave
  // sums (a power-of-two scale commutes with rounding)
  const float comp_scale = (my_comp >= 2 && my_comp <= 4) ? 0.5f : 1.0f;
  // MFMA A operand: lane i + 16g reads row i (row 15 repeats row 0; its
  // output row is discarded).  SPLIT: pixels 32t + 8g .. +7 of each plane.
  // FP32: pixels 16g .. 16g + 15 of the row.
  const int arow = (lane & 15) < kCcRows ? (lane & 15) : 0;
  uint16_t* cc_w = reinterpret_cast<uint16_t*>(cc_raw);
  float* cc_f = reinterpret_cast<float*>(cc_raw);
  const u32x4* a_src0 = reinterpret_cast<const u32x4*>(cc_w + cc_pos(arow, 8 * (lane >> 4)));
  const u32x4* a_src1 = reinterpret_cast<const u32x4*>(cc_w + cc_pos(arow, 32 + 8 * (lane >> 4)));
  const float* a_col = cc_f + cc_pos_f32(arow, 16 * (lane >> 4));  // FP32: pixel 16g + k at a_col[k kCcRows]
  float* shacc_w = s_shacc[wv];
  const int d_ch = lane >> 4;                          // channel of this lane's MFMA output rows
  const int d_k = lane & 15;                           // basis function of this lane's output column

  // The accumulators are zeroed once here; each batch's flush zeroes what it
  // reads, and only entries of the touched list are ever written.
  for (int e = t; e < kMfmaBatch * 6; e += kTilePixels) s_acc[e] = 0.0f;
  for (int e = t; e < 4 * kMfmaBatch * SHN / 4; e += kTilePixels)
    reinterpret_cast<float4*>(&s_shacc[0][0])[e] = make_float4(0.f, 0.f, 0.f, 0.f);
  // Flush of a batch: thread t takes elements e = t + i kTilePixels of the
  // (entry, component) grid, entry e / NCOMP, and reads, zeroes and emits
  // those of touched entries.  The ids and SH rows of its entries are kept
  // in registers (keep_rows, at the end of the batch's compositing), so the
  // flush of batch n - 1 can run while batch n is staged over them.
  // The thread index passes through an empty asm in both, so that their
  // index arithmetic, the same in every batch, is not hoisted out of the
  // batch loop into registers the compositing needs.
  constexpr int kFlushElems = (kMfmaBatch * NCOMP + kTilePixels - 1) / kTilePixels;
  int f_id[kFlushElems], f_row[kFlushElems];
  auto keep_rows = [&]() {
    int tv = t;
    asm volatile("" : "+v"(tv));
#pragma unroll
    for (int i = 0; i < kFlushElems; ++i) {
      const int k = min((tv + i * kTilePixels) / NCOMP, kMfmaBatch - 1);
      const int4 r = reinterpret_cast<const int4*>(s_gc)[k];
      f_id[i] = r.z;
      f_row[i] = r.w;
    }
  };
  auto flush = [&](uint32_t touch) {
    int tv = t;
    asm volatile("" : "+v"(tv));
#pragma unroll
    for (int i = 0; i < kFlushElems; ++i) {
      const int e = tv + i * kTilePixels;
      const int k = e / NCOMP;
      if (e >= kMfmaBatch * NCOMP || !((touch >> k) & 1u)) continue;
      const int c = e - k * NCOMP;
      if (c < 6) {
        const float v = s_acc[k * 6 + c];
        s_acc[k * 6 + c] = 0.0f;
        if (v != 0.0f) emit_grad<SHN>(a, f_id[i], c, v);
      } else {
        const int r = c - 6;
        const int o = k * SHN + r;
        const float v = ((s_shacc[0][o] + s_shacc[1][o]) + s_shacc[2][o]) + s_shacc[3][o];
        s_shacc[0][o] = s_shacc[1][o] = s_shacc[2][o] = s_shacc[3][o] = 0.0f;
        const int ch = r / CC;
        if (v != 0.0f) unsafeAtomicAdd(a.grad_sh + (int64_t)f_row[i] * a.sh_sg + ch * a.sh_sc + (r - ch * CC), v);
      }
    }
  };
  // Batch n, three barriers: stage its geometry | stage its SH rows, flush
  // batch n - 1 (zeroing the accumulators before batch n writes them) |
  // composite batch n, keep its ids and rows for its flush.
  int nb = 0;
  for (int base = ((nmax - 1) / kMfmaBatch) * kMfmaBatch; nmax > 0 && base >= 0; base -= kMfmaBatch, ++nb) {
    const int cnt = min(kMfmaBatch, nmax - base);
    const uint32_t ptouch = nb > 0 ? s_touch : 0u;  // batch n - 1's, complete since the last barrier
    if (t < cnt) {
      const int g = idsb[s0 + base + t];
      const int row = a.sh_index ? a.sh_index[g] : g;
      const float al = stage_geo(a, t, g, s_ga, s_gb);
      reinterpret_cast<int4*>(s_gc)[t] = make_int4(__float_as_int(al), __float_as_int(reject_radius(al)), g, row);
    }
    __syncthreads();
    if (t == 0) s_touch = 0u;  // every thread has read it before the barrier
    stage_sh_rows_of<SHN, CC, true>(
        a, a.sh, cnt, [&](int j) -> int64_t { return reinterpret_cast<const int4*>(s_gc)[j].w; }, s_sh);
    flush(ptouch);
    __syncthreads();
    int n_slots = 0;      // filled slots of the current chunk (wave-uniform)
    uint32_t slots = 0;   // batch position of slot m in bits [kSlotBits m, kSlotBits (m + 1))
    // entries of the batch that can reach a pixel of this wave, walked from
    // the last one down
    const int le = lane < cnt ? lane : 0;
    uint32_t rel =
        (uint32_t)__ballot(lane < cnt && may_reach(s_ga[le], s_gb[le], s_gc[le].y, wr.x, wr.y, wr.z, wr.w));
    uint32_t cmask = 0;  // CMF: chunk slots with a contributing lane
    uint32_t wtouch = 0;  // batch entries this wave contributed to (wave-uniform)
    while (rel) {
      int gj[kChunk];  // CMF: batch positions of the chunk's entries, last first (wave-uniform)
      int gn = 1;
      if constexpr (CMF) {
        gn = 0;
#pragma unroll
        for (int m = 0; m < kChunk; ++m) {
          gj[m] = 0;
          if (rel) {
            gj[m] = 31 - __builtin_clz(rel);
            rel &= ~(1u << gj[m]);
            gn = m + 1;
          }
        }
        // A operand: row ch * kChunk + m (this lane's row i = lane % 16), basis 4 s + lane / 16
        const int ai = lane & 15, am = ai % kChunk, ach = ai / kChunk;
        int aj = gj[0];
#pragma unroll
        for (int m = 1; m < kChunk; ++m) aj = am == m ? gj[m] : aj;
        const bool arow_ok = ai < kCcRows && am < gn;
        float av[KS];
#pragma unroll
        for (int st = 0; st < KS; ++st) {
          const int k = 4 * st + (lane >> 4);
          av[st] = (arow_ok && k < CC) ? s_sh[aj * SHN + ach * CC + k] : 0.0f;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int st = 0; st < KS; ++st)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[st], yB[4 * st + b], acc, 0, 0, 0);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = 4 * (lane >> 4) + r;
            if (row < kCcRows) cc_f[cc_pos_f32(row, 16 * b + (lane & 15))] = acc[r];
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
      for (int gi = 0; gi < gn; ++gi) {
      int j;
      if constexpr (CMF) {
        j = gj[gi];
      } else {
        j = 31 - __builtin_clz(rel);
        rel &= ~(1u << j);
      }
      const int idx = base + j;
      // Per-lane decisions are selects, not branches: the wave runs the
      // entry when any lane needs it, and lanes that do not contribute keep
      // their state (a divergent branch here cost more in exec-mask handling
      // and re-zeroing than the masked arithmetic it skipped).
      const bool live = idx < nd;
      if (wave_any(live)) {
        const GeoA ga = s_ga[j];
        const GeoB gb = s_gb[j];
        const float4 gc = s_gc[j];
        const float al = gc.x;
        const float dx = px - ga.mx, dy = py - ga.my;
        const float tmpx = dx * gb.c3 - dy * gb.c2;
        const float tmpy = -dx * ga.c1 + dy * ga.c0;
        const float q = tmpx * dx + tmpy * dy;
        const float rq = q * gb.inv_det;
        // bitwise &, not &&: the conditions stay lane masks (no branch); the
        // wave tests combine the ballots of single comparisons.  `live` is
        // compared again here (one instruction) rather than carried into the
        // block, which the compiler does through a vector register (two).
        int idx_s = __builtin_amdgcn_readfirstlane(idx);  // wave-uniform
        asm volatile("" : "+s"(idx_s));  // a scalar copy: the compare is not merged with the first
        const bool live_in = idx_s < nd;
        const bool near = live_in & !(rq > gc.y);
        const uint64_t b_near = __builtin_amdgcn_ballot_w64(live_in) & __builtin_amdgcn_ballot_w64(!(rq > gc.y));
        float G = gauss_fast(rq);
        const bool exact = near & gauss_needs_exact(al * G, q);
        if (b_near & ballot_needs_exact(al * G, q)) {
          if (exact) G = gauss_exact(q, gb.det);
        }
        const float aG = al * G;
        const bool contrib = near & !(aG < kMinRenderAlpha);
        if (b_near & __builtin_amdgcn_ballot_w64(!(aG < kMinRenderAlpha))) {
          const float om = 1.0f - aG;
          const float inv_om = __builtin_amdgcn_rcpf(om);
          const float Tj = T * inv_om;
          // the weight, zero where the entry does not contribute (or is NaN)
          float w = al * Tj * G;
          w = (contrib && !isnan(w)) ? w : 0.0f;
          const int m = CMF ? gi : n_slots;
          float d0, d1, d2;
          if constexpr (CMF) {  // this entry's dot products, from the chunk's matrix-core product
            d0 = cc_f[cc_pos_f32(0 * kChunk + m, lane)];
            d1 = cc_f[cc_pos_f32(1 * kChunk + m, lane)];
            d2 = cc_f[cc_pos_f32(2 * kChunk + m, lane)];
          } else {
            const float* shj = s_sh + j * SHN;
            d0 = dot_fast<CC>(shj, Y);
            d1 = dot_fast<CC>(shj + CC, Y);
            d2 = dot_fast<CC>(shj + 2 * CC, Y);
          }
          // w is finite, a sigmoid NaN or in [0, 1]: isnan(y * w) == isnan(y)
          // the staged coefficients carry -log2(e): d is the exponent's argument
          const float y0 = fmaxf(sigmoid_scaled(d0), 0.0f);
          const float y1 = fmaxf(sigmoid_scaled(d1), 0.0f);
          const float y2 = fmaxf(sigmoid_scaled(d2), 0.0f);
          const float cc0 = contrib ? w * fmaf(-y0, y0, y0) * g0 : 0.0f;  // y (1 - y): the sigmoid's slope
          const float cc1 = contrib ? w * fmaf(-y1, y1, y1) * g1 : 0.0f;
          const float cc2 = contrib ? w * fmaf(-y2, y2, y2) * g2 : 0.0f;
          const float gy = fmaf(g2, y2, fmaf(g1, y1, g0 * y0));
          // sum_c g_c (y_c T_j - S_c / (1 - aG))
          const float P = fmaf(gy, Tj, -(gS * inv_om));
          // w = 0 leaves gS unchanged where gy is finite; where it is not, so
          // is P of every later contributing entry of the pixel
          gS = fmaf(w, gy, gS);
          T = contrib ? Tj : T;
          // zero factors, selected before the products, make the mean and
          // covariance terms exact zeros where the entry does not contribute
          const float D = contrib ? P * aG : 0.0f;
          const float sx = contrib ? tmpx * gb.inv_det : 0.0f;
          const float uy = contrib ? tmpy * gb.inv_det : 0.0f;
          const float Ds = D * sx, Du = D * uy;
          float v[8];
          v[0] = Ds;
          v[1] = Du;
          v[2] = Ds * sx;  // x 1/2 after the reduction (comp_scale)
          v[3] = Ds * uy;
          v[4] = Du * uy;
          v[5] = contrib ? P * G : 0.0f;
          // components 6 and 7 are unused and their lanes never write, so
          // any register serves as their partner in the exchanges (its
          // value stays in those lanes' sums)
          float spare6, spare7;  // two: each exchange rewrites its partner
          asm volatile("" : "=v"(spare6), "=v"(spare7));
          v[6] = spare6;
          v[7] = spare7;
          if constexpr (SPLIT) {
            uint32_t h, l;
            split_bf16(cc0, h, l);
            cc_w[cc_pos(0 * kChunk + m, lane)] = h;
            cc_w[kCcPlane + cc_pos(0 * kChunk + m, lane)] = l;
            split_bf16(cc1, h, l);
            cc_w[cc_pos(1 * kChunk + m, lane)] = h;
            cc_w[kCcPlane + cc_pos(1 * kChunk + m, lane)] = l;
            split_bf16(cc2, h, l);
            cc_w[cc_pos(2 * kChunk + m, lane)] = h;
            cc_w[kCcPlane + cc_pos(2 * kChunk + m, lane)] = l;
          } else {
            cc_f[cc_pos_f32(0 * kChunk + m, lane)] = cc0;
            cc_f[cc_pos_f32(1 * kChunk + m, lane)] = cc1;
            cc_f[cc_pos_f32(2 * kChunk + m, lane)] = cc2;
          }
          if constexpr (CMF) {
            cmask |= 1u << m;
          } else {
            slots |= static_cast<uint32_t>(j) << (kSlotBits * m);
            n_slots = m + 1;
          }
          wtouch |= 1u << j;
          const float r = wave_reduce_scatter8_swap(v, lane);
          if (writer && r != 0.0f) {
            // LDS float add at s_acc[6 j + my_comp]: the entry's byte offset
            // formed in a scalar register (the empty asm keeps it there) and
            // added to the lane's address; otherwise the compiler fuses both
            // into a 64-bit multiply-add per entry
            typedef __attribute__((address_space(3))) float lds_float;
            typedef __attribute__((address_space(3))) char lds_char;
            uint32_t jb = __builtin_amdgcn_readfirstlane(j * 24);  // j is wave-uniform
            asm volatile("" : "+s"(jb));
            lds_float* dst = (lds_float*)((lds_char*)((lds_float*)s_acc + my_comp) + jb);
            __hip_atomic_fetch_add(dst, r * comp_scale, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
      }
      }  // entries of the chunk
      if (CMF ? cmask != 0u : (n_slots == kChunk || (rel == 0 && n_slots > 0))) {
        // LDS operations of one wave complete in order; the barrier keeps the
        // compiler from moving the loads above the stores.
        __builtin_amdgcn_wave_barrier();
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        if constexpr (SPLIT) {
          const u32x4 ah0 = a_src0[0], ah1 = a_src1[0];  // pixels 8g.., 32 + 8g..
          const u32x4 al0 = a_src0[kCcPlane / 8], al1 = a_src1[kCcPlane / 8];
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ah0),
                                                        __builtin_bit_cast(bf16x8, bh[0]), acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ah0),
                                                        __builtin_bit_cast(bf16x8, bl[0]), acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, al0),
                                                        __builtin_bit_cast(bf16x8, bh[0]), acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ah1),
                                                        __builtin_bit_cast(bf16x8, bh[1]), acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ah1),
                                                        __builtin_bit_cast(bf16x8, bl[1]), acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, al1),
                                                        __builtin_bit_cast(bf16x8, bh[1]), acc, 0, 0, 0);
        } else {
          float av[16];
#pragma unroll
          for (int k = 0; k < 16; ++k) av[k] = a_col[k * kCcRows];
#pragma unroll
          for (int k = 0; k < 16; ++k) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[k], bY[k], acc, 0, 0, 0);
          // all operand reads in flight before the first product (one LDS
          // round trip per chunk instead of one per pair of products)
          __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);
        }
        if (d_k < CC) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = 4 * d_ch + r;  // = ch * kChunk + slot
            const int ch = row / kChunk, m = row - ch * kChunk;
            if constexpr (CMF) {
              if (row < kCcRows && ((cmask >> m) & 1u)) {
                int jj = gj[0];
#pragma unroll
                for (int mm = 1; mm < kChunk; ++mm) jj = m == mm ? gj[mm] : jj;
                shacc_w[jj * SHN + ch * CC + d_k] = acc[r];
              }
            } else if (row < kCcRows && m < n_slots) {
              const int jj = (slots >> (kSlotBits * m)) & ((1u << kSlotBits) - 1);
              shacc_w[jj * SHN + ch * CC + d_k] = acc[r];
            }
          }
        }
        __builtin_amdgcn_wave_barrier();
        n_slots = 0;
        slots = 0;
        cmask = 0u;
      }
    }
    if (lane == 0 && wtouch) atomicOr(&s_touch, wtouch);
    keep_rows();
    __syncthreads();
  }
  if (nb > 0) flush(s_touch);  // the last batch
}

template <int C, bool BG, bool SPLIT, bool RECOMP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_bwd_sh_mfma(RasterArgs a) {
  bwd_sh_mfma_body<C, BG, SPLIT, RECOMP, false>(a);
}

// The matrix-core colours need the 16 basis values of the product's B
// operand in registers as well: four waves per SIMD instead of five.
template <int C, bool BG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_bwd_sh_mfma_cmf(RasterArgs a) {
  bwd_sh_mfma_body<C, BG, false, false, true>(a);
}

// ---------------------------------------------------------------------------
// RGB compositing (render_start_end): vol_render.h:716-798 forward,
// vol_render.h:800-905 backward; the Gaussian is evaluated in double
// (kernels.h:195-224).  Colour [V,3] is staged through the SH slots with
// sh_sg = 3, sh_sc = 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stage_rgb(const RasterArgs& a, int s0, int base, int cnt, int* s_id, GeoA* s_ga,
                                          GeoB* s_gb, float2* s_gc, float* s_col) {
  const int t = threadIdx.x;
  if (t < cnt) {
    const int g = a.ids[s0 + base + t];
    s_id[t] = g;
    const float2 m = reinterpret_cast<const float2*>(a.mean)[g];
    const float4 c = reinterpret_cast<const float4*>(a.cov)[g];
    const float det = c.x * c.w - c.y * c.z;
    const float al = fminf(a.alpha[g], 0.99f);
    s_ga[t] = GeoA{m.x, m.y, c.x, c.y};
    s_gb[t] = GeoB{c.z, c.w, det, 1.0f / det};
    s_gc[t] = make_float2(al, reject_radius(al));
    s_col[3 * t + 0] = a.sh[3 * (int64_t)g + 0];
    s_col[3 * t + 1] = a.sh[3 * (int64_t)g + 1];
    s_col[3 * t + 2] = a.sh[3 * (int64_t)g + 2];
  }
}

__global__ __launch_bounds__(256) void k_fwd_rgb(RasterArgs a) {
  __shared__ GeoA s_ga[kBatch];
  __shared__ GeoB s_gb[kBatch];
  __shared__ float2 s_gc[kBatch];
  __shared__ int s_id[kBatch];
  __shared__ float s_col[kBatch * 3];
  TileCtx tc;
  if (!tile_setup(a, tc)) return;
  const int tile = tc.tile;
  const bool inside = tc.inside;
  const int64_t pix = tc.pix;
  const int s0 = a.start[tile];
  const int e0 = a.end[tile];
  if (s0 == -1 || e0 - s0 == 0) {
    if (inside) {
      if (a.write_empty) a.out[3 * pix] = a.out[3 * pix + 1] = a.out[3 * pix + 2] = 0.0f;
      if (a.T_final) a.T_final[pix] = 1.0f;
      if (a.n_done) a.n_done[pix] = 0;
    }
    return;
  }
  const float px = tc.px;
  const float py = tc.py;
  const float thresh = a.thresh;
  float T = 1.0f, o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
  bool done = !inside;
  const int n = e0 - s0;
  int ndone = n;
  if (inside && T < thresh) {
    done = true;
    ndone = 0;
  }
  for (int base = 0; base < n; base += kBatch) {
    const int cnt = min(kBatch, n - base);
    __syncthreads();
    stage_rgb(a, s0, base, cnt, s_id, s_ga, s_gb, s_gc, s_col);
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      if (__all(done)) break;
      if (done) continue;
      const GeoA ga = s_ga[j];
      const GeoB gb = s_gb[j];
      const float2 gc = s_gc[j];
      const float dx = px - ga.mx, dy = py - ga.my;
      const float tmpx = dx * gb.c3 - dy * gb.c2;
      const float tmpy = -dx * ga.c1 + dy * ga.c0;
      const float q = tmpx * dx + tmpy * dy;
      if (q * gb.inv_det > gc.y) continue;
      const float G = gauss_exact_d(dx, dy, ga.c0, ga.c1, gb.c2, gb.c3);
      const float al = gc.x;
      float coeff = al * T;
      coeff *= G;
      if (al * G < kMinRenderAlpha) continue;
      o0 = o0 + s_col[3 * j + 0] * coeff;
      o1 = o1 + s_col[3 * j + 1] * coeff;
      o2 = o2 + s_col[3 * j + 2] * coeff;
      T = T * (1.0f - al * G);
      if (T < thresh) {
        done = true;
        ndone = base + j + 1;
      }
    }
    if (!__syncthreads_or(!done)) break;
  }
  if (!inside) return;
  a.out[3 * pix + 0] = o0;
  a.out[3 * pix + 1] = o1;
  a.out[3 * pix + 2] = o2;
  if (a.T_final) a.T_final[pix] = T;
  if (a.n_done) a.n_done[pix] = ndone;
}

template <bool RECOMP>
__global__ __launch_bounds__(256) void k_bwd_rgb(RasterArgs a) {
  constexpr int NCOMP = 9;
  constexpr int NV = 16;
  __shared__ GeoA s_ga[kBatch];
  __shared__ GeoB s_gb[kBatch];
  __shared__ float2 s_gc[kBatch];
  __shared__ int s_id[kBatch];
  __shared__ float s_col[kBatch * 3];
  __shared__ float s_acc[kBatch * NCOMP];
  __shared__ int s_tmp[4];
  TileCtx tc;
  if (!tile_setup(a, tc)) return;
  const int tile = tc.tile;
  const int t = threadIdx.x;
  const bool inside = tc.inside;
  const int64_t pix = tc.pix;
  const int s0 = a.start[tile];
  const int e0 = a.end[tile];
  if (s0 == -1 || e0 - s0 == 0) return;
  const float px = tc.px;
  const float py = tc.py;
  float T = 1.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, S0 = 0.0f, S1 = 0.0f, S2 = 0.0f;
  int nd = 0;
  if constexpr (RECOMP) {
    float p0, p1, p2;
    forward_state<kBatch, 1, false, true>(a, s0, e0 - s0, nullptr, px, py, inside, s_id, s_ga, s_gb, s_gc, s_col,
                                          T, nd, p0, p1, p2);
  }
  if (inside) {
    if constexpr (!RECOMP) {
      T = a.T_saved[pix];
      nd = a.n_done_saved[pix];
    }
    g0 = a.grad_out[3 * pix + 0];
    g1 = a.grad_out[3 * pix + 1];
    g2 = a.grad_out[3 * pix + 2];
  }
  const int nmax = block_max(nd, s_tmp);
  const int lane = lane_id();
  constexpr int kLaneShift = 6 - __builtin_ctz(NV);
  const int my_comp = lane >> kLaneShift;
  const bool writer = (lane & ((1 << kLaneShift) - 1)) == 0 && my_comp < NCOMP;
  for (int base = ((nmax - 1) / kBatch) * kBatch; nmax > 0 && base >= 0; base -= kBatch) {
    const int cnt = min(kBatch, nmax - base);
    __syncthreads();
    for (int e = t; e < cnt * NCOMP; e += kTilePixels) s_acc[e] = 0.0f;
    stage_rgb(a, s0, base, cnt, s_id, s_ga, s_gb, s_gc, s_col);
    __syncthreads();
    for (int j = cnt - 1; j >= 0; --j) {
      const int idx = base + j;
      if (!wave_any(idx < nd)) continue;
      float c0v = 0.0f, c1v = 0.0f, c2v = 0.0f, Ds = 0.0f, Du = 0.0f, Dss = 0.0f, Dsu = 0.0f, Duu = 0.0f,
            PG = 0.0f;
      bool contrib = false;
      if (idx < nd) {
        const GeoA ga = s_ga[j];
        const GeoB gb = s_gb[j];
        const float2 gc = s_gc[j];
        const float dx = px - ga.mx, dy = py - ga.my;
        const float tmpx = dx * gb.c3 - dy * gb.c2;
        const float tmpy = -dx * ga.c1 + dy * ga.c0;
        const float q = tmpx * dx + tmpy * dy;
        if (!(q * gb.inv_det > gc.y)) {
          const float G = gauss_exact_d(dx, dy, ga.c0, ga.c1, gb.c2, gb.c3);
          const float al = gc.x;
          const float aG = al * G;
          if (!(aG < kMinRenderAlpha)) {
            contrib = true;
            const float om = 1.0f - aG;
            const float Tj = T / om;
            const float w = al * Tj * G;
            const float y0 = s_col[3 * j], y1 = s_col[3 * j + 1], y2 = s_col[3 * j + 2];
            c0v = w * g0;
            c1v = w * g1;
            c2v = w * g2;
            const float inv_om = 1.0f / om;
            float P = g0 * (y0 * Tj - S0 * inv_om);
            P = fmaf(g1, y1 * Tj - S1 * inv_om, P);
            P = fmaf(g2, y2 * Tj - S2 * inv_om, P);
            S0 = fmaf(w, y0, S0);
            S1 = fmaf(w, y1, S1);
            S2 = fmaf(w, y2, S2);
            const float D = P * aG;
            const float inv_det = 1.0f / gb.det;
            const float s = tmpx * inv_det;
            const float u = tmpy * inv_det;
            Ds = D * s;
            Du = D * u;
            Dss = 0.5f * Ds * s;
            Dsu = 0.5f * Ds * u;
            Duu = 0.5f * Du * u;
            PG = P * G;
            T = Tj;
          }
        }
      }
      if (!wave_any(contrib)) continue;
      float v[NV] = {Ds, Du, Dss, Dsu, Duu, PG, c0v, c1v, c2v, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      const float r = wave_reduce_scatter<NV>(v);
      if (writer && r != 0.0f) atomicAdd(s_acc + j * NCOMP + my_comp, r);
    }
    __syncthreads();
    flush_grads<3>(a, s_id, s_acc, cnt);
  }
}

// ---------------------------------------------------------------------------
// Backward launch order.  A tile's backward cost is set by how far its
// workgroup walks the list: its four waves meet at a barrier every batch, so
// the largest n_done over the whole tile sets the number of batches and the
// slowest wave each batch's length.  Costs vary ~3x between tiles, and tiles
// in image order put expensive ones late; launching the tiles most expensive
// first (longest processing time first) shortens the tail of the launch.
// The order is a stable counting sort on the cost quantised to 64 levels, in
// one workgroup: within a level the tiles keep image order, so neighbours
// (which share Gaussians) still run close together.
// ---------------------------------------------------------------------------
// Cost of each tile: four times the most entries a pixel of the tile examined
// (its largest n_done), plus a fixed 32 (0 for an empty list).  One thread per
// 8x8 quadrant; the four quadrants of a tile are consecutive lanes.  (The sum
// over the four waves of their own largest n_done predicts the time less
// well: 1.60 against 1.54 ms at 3M Gaussians, 1080p.)
#ifndef GSMI_ORDER_CONST
#define GSMI_ORDER_CONST 32
#endif
__global__ __launch_bounds__(256) void k_tile_cost(RasterArgs a, uint32_t* __restrict__ cost) {
  const int64_t qd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int tile = (int)(qd >> 2);
  const bool valid = tile < a.n_tiles_total;
  int mx = 0;
  if (valid) {
    int local = tile, ntw = a.n_tiles_w, W = a.W, H = a.H;
    int64_t pix_base = 0;
    if (a.views) {
      const gsmi_view& v = a.views[view_of_tile(a.views, a.n_views, tile)];
      local = tile - v.tile_base;
      ntw = v.n_tiles_w;
      W = v.W;
      H = v.H;
      pix_base = v.pix_base;
    }
    const int w = (int)(qd & 3);
    const int gx0 = (local % ntw) * kTile + ((w & 1) << 3), gy0 = (local / ntw) * kTile + ((w >> 1) << 3);
    for (int r = 0; r < 8 && gy0 + r < H; ++r) {
      const int32_t* row = a.n_done_saved + pix_base + (int64_t)(gy0 + r) * W;
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (gx0 + c < W) mx = max(mx, row[gx0 + c]);
    }
  }
  int m4 = max(mx, __shfl_xor(mx, 1));
  m4 = max(m4, __shfl_xor(m4, 2));
  if (valid && (qd & 3) == 0) {
    const int s0 = a.start[tile], e0 = a.end[tile];
    const bool empty = s0 == -1 || e0 <= s0;
    cost[tile] = empty ? 0u : 4u * (uint32_t)m4 + (uint32_t)GSMI_ORDER_CONST;
  }
}

constexpr int kOrderThreads = 1024;
#ifndef GSMI_ORDER_LEVELS
#define GSMI_ORDER_LEVELS 64
#endif
constexpr int kOrderLevels = GSMI_ORDER_LEVELS;  // cost levels; within a level tiles keep their index order
// bits of a level or of the marker kOrderLevels (no level)
constexpr int kLevelBits = 32 - __builtin_clz((unsigned)kOrderLevels);
static_assert(kOrderLevels >= 2 && kOrderLevels <= 64, "one lane per level in the placement counts");
constexpr int kOrderSingleMax = 1 << 15;  // grids up to this many tiles are ordered by one workgroup

// Cost level of a tile: 0 holds the most expensive.  scale = (levels - 1) /
// the largest cost (>= 1), from order_scale; the level is a launch-order
// heuristic, so float rounding at a level's edge is of no consequence.
__device__ __forceinline__ float order_scale(uint32_t mx) { return (float)(kOrderLevels - 1) / (float)max(mx, 1u); }
__device__ __forceinline__ uint32_t order_level(uint32_t c, float scale) {
  return (uint32_t)((kOrderLevels - 1) - min((int)((float)c * scale), kOrderLevels - 1));
}

__global__ __launch_bounds__(kOrderThreads) void k_tile_order(const uint32_t* __restrict__ cost, int n,
                                                              int32_t* __restrict__ order) {
  constexpr int NW = kOrderThreads / 64;
  __shared__ uint32_t s_cursor[kOrderLevels];
  __shared__ uint32_t s_wc[NW][kOrderLevels];
  __shared__ uint32_t s_max;
  __shared__ uint8_t s_lv[kOrderSingleMax];  // each tile's level, from the counting pass
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  if (t == 0) s_max = 1u;
  if (t < kOrderLevels) s_cursor[t] = 0u;
  __syncthreads();
  uint32_t m = 0u;
  for (int i = t; i < n; i += kOrderThreads) m = max(m, cost[i]);
  atomicMax(&s_max, m);
  __syncthreads();
  const float scale = order_scale(s_max);
  for (int i = t; i < n; i += kOrderThreads) {
    const uint32_t lv = order_level(cost[i], scale);
    s_lv[i] = (uint8_t)lv;
    atomicAdd(&s_cursor[lv], 1u);
  }
  __syncthreads();
  if (w == 0) {  // exclusive prefix over the levels, one lane per level
    const uint32_t c = lane < kOrderLevels ? s_cursor[lane] : 0u;
    uint32_t inc = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = (uint32_t)__shfl_up((int)inc, d);
      if (lane >= d) inc += o;
    }
    if (lane < kOrderLevels) s_cursor[lane] = inc - c;
  }
  __syncthreads();
  // Stable placement, kOrderThreads tiles per round: rank among the lanes of
  // the wave with the same level (ballot match on the level bits), then the
  // counts of the waves before it, then the level's cursor.
  for (int c0 = 0; c0 < n; c0 += kOrderThreads) {
    const int i = c0 + t;
    const bool valid = i < n;
    const uint32_t lv = valid ? (uint32_t)s_lv[i] : (uint32_t)kOrderLevels;  // kOrderLevels: no level
    uint64_t peers = ~0ull;
#pragma unroll
    for (int bit = 0; bit < kLevelBits; ++bit) {
      const bool on = (lv >> bit) & 1u;
      const uint64_t bal = __ballot(on);
      peers &= on ? bal : ~bal;
    }
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
    if (lane < kOrderLevels) s_wc[w][lane] = 0u;
    __syncthreads();
    if (valid && below == 0) s_wc[w][lv] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (t < kOrderLevels) {  // per level: the waves' counts -> their start positions
      uint32_t run = s_cursor[t];
      for (int v = 0; v < NW; ++v) {
        const uint32_t c = s_wc[v][t];
        s_wc[v][t] = run;
        run += c;
      }
      s_cursor[t] = run;
    }
    __syncthreads();
    if (valid) order[s_wc[w][lv] + below] = i;
    __syncthreads();
  }
}

// Larger grids: the same levels, placed by many workgroups.  Counter words
// (cw): [0] the largest cost, [1 + l] the tiles of level l, [1 + kOrderLevels
// + l] level l's next free position.  Each workgroup places kOrderThreads
// tiles, reserving its run of each level with one atomic, so within a level
// tiles keep their index order inside a workgroup and workgroups follow in
// the order they reserve.
constexpr int kOrderWords = 1 + 2 * kOrderLevels;

__global__ void k_order_zero(uint32_t* __restrict__ cw) {
  for (int i = threadIdx.x; i < kOrderWords; i += blockDim.x) cw[i] = 0u;
}

__global__ __launch_bounds__(256) void k_order_max(const uint32_t* __restrict__ cost, int n, uint32_t* __restrict__ cw) {
  uint32_t m = 0u;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = max(m, cost[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
  if ((threadIdx.x & 63) == 0 && m) atomicMax(cw, m);
}

__global__ __launch_bounds__(256) void k_order_count(const uint32_t* __restrict__ cost, int n,
                                                     uint32_t* __restrict__ cw) {
  __shared__ uint32_t s_c[kOrderLevels];
  if (threadIdx.x < kOrderLevels) s_c[threadIdx.x] = 0u;
  __syncthreads();
  const float scale = order_scale(cw[0]);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) atomicAdd(&s_c[order_level(cost[i], scale)], 1u);
  __syncthreads();
  if (threadIdx.x < kOrderLevels && s_c[threadIdx.x]) atomicAdd(&cw[1 + threadIdx.x], s_c[threadIdx.x]);
}

__global__ void k_order_scan(uint32_t* __restrict__ cw) {
  if (threadIdx.x == 0) {
    uint32_t run = 0u;
    for (int l = 0; l < kOrderLevels; ++l) {
      cw[1 + kOrderLevels + l] = run;
      run += cw[1 + l];
    }
  }
}

__global__ __launch_bounds__(kOrderThreads) void k_order_place(const uint32_t* __restrict__ cost, int n,
                                                               uint32_t* __restrict__ cw, int32_t* __restrict__ order) {
  constexpr int NW = kOrderThreads / 64;
  __shared__ uint32_t s_wc[NW][kOrderLevels];
  __shared__ uint32_t s_base[kOrderLevels];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const float scale = order_scale(cw[0]);
  const int i = blockIdx.x * kOrderThreads + t;
  const bool valid = i < n;
  const uint32_t lv = valid ? order_level(cost[i], scale) : (uint32_t)kOrderLevels;  // kOrderLevels: no level
  uint64_t peers = ~0ull;
#pragma unroll
  for (int bit = 0; bit < kLevelBits; ++bit) {
    const bool on = (lv >> bit) & 1u;
    const uint64_t bal = __ballot(on);
    peers &= on ? bal : ~bal;
  }
  const uint32_t below =
      __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
  if (lane < kOrderLevels) s_wc[w][lane] = 0u;
  __syncthreads();
  if (valid && below == 0) s_wc[w][lv] = (uint32_t)__popcll(peers);
  __syncthreads();
  if (t < kOrderLevels) {
    uint32_t tot = 0u;
    for (int v = 0; v < NW; ++v) tot += s_wc[v][t];
    s_base[t] = tot ? atomicAdd(&cw[1 + kOrderLevels + t], tot) : 0u;
  }
  __syncthreads();
  if (valid) {
    uint32_t pos = s_base[lv] + below;
    for (int v = 0; v < w; ++v) pos += s_wc[v][lv];
    order[pos] = i;
  }
}

static size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }
// costs, order, then the counter words of large grids
static size_t order_ws_bytes(int n_tiles) {
  return 2 * align_up((size_t)(n_tiles > 0 ? n_tiles : 0) * 4) + align_up(kOrderWords * 4);
}

// Order of the tiles by cost[] (most expensive level first, stable within a
// level); cw: the counter words of grids beyond kOrderSingleMax tiles.
static void place_tile_order(const uint32_t* cost, int n, int32_t* order, uint32_t* cw, hipStream_t st) {
  if (n <= kOrderSingleMax) {
    hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(kOrderThreads), 0, st, cost, n, order);
  } else {
    const int g = std::min((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_order_zero, dim3(1), dim3(64), 0, st, cw);
    hipLaunchKernelGGL(k_order_max, dim3(g), dim3(256), 0, st, cost, n, cw);
    hipLaunchKernelGGL(k_order_count, dim3(g), dim3(256), 0, st, cost, n, cw);
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(64), 0, st, cw);
    hipLaunchKernelGGL(k_order_place, dim3((n + kOrderThreads - 1) / kOrderThreads), dim3(kOrderThreads), 0, st,
                       cost, n, cw, order);
  }
}

// Fills a.tile_order from the saved n_done (workspace: costs, then order).
static int make_tile_order(RasterArgs& a, void* ws, hipStream_t st) {
  const int n = a.n_tiles_total;
  uint32_t* cost = reinterpret_cast<uint32_t*>(ws);
  int32_t* order = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(ws) + align_up((size_t)n * 4));
  uint32_t* cw = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ws) + 2 * align_up((size_t)n * 4));
  a.tile_order = nullptr;
  hipLaunchKernelGGL(k_tile_cost, dim3((unsigned)((4 * (int64_t)n + 255) / 256)), dim3(256), 0, st, a, cost);
  place_tile_order(cost, n, order, cw, st);
  GSMI_CHECK_LAUNCH();
  a.tile_order = order;
  return GSMI_OK;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
template <int C, bool BG>
static int launch_fwd_sh(const RasterArgs& a, int exact, hipStream_t st) {
  const int grid = xcd_grid(a.n_tiles_total);
  if (exact)
    hipLaunchKernelGGL((k_fwd_sh<C, BG, true>), dim3(grid), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((k_fwd_sh<C, BG, false>), dim3(grid), dim3(256), 0, st, a);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

// recomp: no saved forward state (a.T_saved == null); split: bf16 split of
// the SH-gradient reduction (orders 3 and 4, saved state only).
template <int C, bool BG>
static int launch_bwd_sh(const RasterArgs& a, bool split, bool recomp, hipStream_t st, bool cmf = false) {
  const int grid = xcd_grid(a.n_tiles_total);
  if constexpr (C >= 3) {
    if (recomp)
      hipLaunchKernelGGL((k_bwd_sh_mfma<C, BG, false, true>), dim3(grid), dim3(256), 0, st, a);
    else if (cmf)
      hipLaunchKernelGGL((k_bwd_sh_mfma_cmf<C, BG>), dim3(grid), dim3(256), 0, st, a);
    else if (split)
      hipLaunchKernelGGL((k_bwd_sh_mfma<C, BG, true, false>), dim3(grid), dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL((k_bwd_sh_mfma<C, BG, false, false>), dim3(grid), dim3(256), 0, st, a);
  } else {
    if (recomp)
      hipLaunchKernelGGL((k_bwd_sh<C, BG, true>), dim3(grid), dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL((k_bwd_sh<C, BG, false>), dim3(grid), dim3(256), 0, st, a);
  }
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

static int dispatch_bwd_sh(const RasterArgs& a, int C, int flags, bool recomp, hipStream_t st) {
  const bool B = a.bg != nullptr;
  const bool split = (flags & GSMI_BWD_SPLIT_BF16) != 0, cmf = (flags & GSMI_BWD_COLOUR_MFMA) != 0;
  switch (C) {
    case 1: return B ? launch_bwd_sh<1, true>(a, split, recomp, st) : launch_bwd_sh<1, false>(a, split, recomp, st);
    case 2: return B ? launch_bwd_sh<2, true>(a, split, recomp, st) : launch_bwd_sh<2, false>(a, split, recomp, st);
    case 3:
      return B ? launch_bwd_sh<3, true>(a, split, recomp, st, cmf) : launch_bwd_sh<3, false>(a, split, recomp, st, cmf);
    default:
      return B ? launch_bwd_sh<4, true>(a, split, recomp, st, cmf) : launch_bwd_sh<4, false>(a, split, recomp, st, cmf);
  }
}

static int check_common(int64_t V, int64_t P, int tile_size, int n_tiles_h, int n_tiles_w, int H, int W) {
  if (V < 0 || P < 0 || n_tiles_h <= 0 || n_tiles_w <= 0 || H <= 0 || W <= 0) return GSMI_ERR_ARG;
  if (tile_size != kTile) return GSMI_ERR_UNSUPPORTED;
  if ((int64_t)n_tiles_h * kTile < H || (int64_t)n_tiles_w * kTile < W) return GSMI_ERR_ARG;
  if ((int64_t)n_tiles_h * n_tiles_w > (1ll << 31) / 2) return GSMI_ERR_ARG;
  return GSMI_OK;
}

// Checks shared by the SH backward entry points; sets *recomp when the
// forward state is to be recomputed (T_final and n_done both null).
static int check_bwd_state(const float* out, const float* T_final, const int32_t* n_done, const float* pre_bg,
                           const float* bg, int flags, const void* workspace, bool* recomp) {
  if ((T_final == nullptr) != (n_done == nullptr)) return GSMI_ERR_ARG;
  *recomp = T_final == nullptr;
  if (flags & ~(GSMI_BWD_SPLIT_BF16 | GSMI_BWD_COLOUR_MFMA)) return GSMI_ERR_ARG;
  if (*recomp && flags) return GSMI_ERR_UNSUPPORTED;
  if ((flags & GSMI_BWD_SPLIT_BF16) && (flags & GSMI_BWD_COLOUR_MFMA)) return GSMI_ERR_UNSUPPORTED;
  if (*recomp && workspace) return GSMI_ERR_ARG;  // the launch order is derived from the saved n_done
  if (bg && (!out || (!*recomp && !pre_bg))) return GSMI_ERR_ARG;
  return GSMI_OK;
}

}  // namespace gsmi

using namespace gsmi;

extern "C" int gsmi_render_sh_forward(int64_t V, int64_t P, const float* mean, const float* cov, const float* sh,
                                      int64_t sh_stride_g, int64_t sh_stride_c, const int32_t* sh_index,
                                      const float* alpha, const int32_t* start, const int32_t* end,
                                      const int32_t* gaussian_ids, float* out, const float* topleft,
                                      const float* c2w, int tile_size, int n_tiles_h, int n_tiles_w,
                                      float pixel_size_x, float pixel_size_y, int H, int W, int C, float thresh,
                                      const float* bg, float* T_final, int32_t* n_done, float* pre_bg, int flags,
                                      int exact, void* stream) {
  int rc = check_common(V, P, tile_size, n_tiles_h, n_tiles_w, H, W);
  if (rc) return rc;
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (!start || !end || !out || !topleft || !c2w) return GSMI_ERR_ARG;
  if (P > 0 && (!mean || !cov || !sh || !alpha || !gaussian_ids)) return GSMI_ERR_ARG;
  if (sh_stride_c < C * C || sh_stride_g < 3 * sh_stride_c) return GSMI_ERR_ARG;
  if (((uintptr_t)mean & 7) || ((uintptr_t)cov & 15)) return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh = sh;
  a.sh_index = sh_index;
  a.alpha = alpha;
  a.start = start;
  a.end = end;
  a.ids = gaussian_ids;
  a.topleft = topleft;
  a.c2w = c2w;
  a.bg = bg;
  a.out = out;
  a.T_final = T_final;
  a.n_done = n_done;
  a.pre_bg = pre_bg;
  a.sh_sg = sh_stride_g;
  a.sh_sc = sh_stride_c;
  a.psx = pixel_size_x;
  a.psy = pixel_size_y;
  a.thresh = thresh;
  a.n_tiles_h = n_tiles_h;
  a.n_tiles_w = n_tiles_w;
  a.n_tiles_total = n_tiles_h * n_tiles_w;
  a.H = H;
  a.W = W;
  a.write_empty = flags & 1;
  hipStream_t st = (hipStream_t)stream;
  const bool B = bg != nullptr;
  switch (C) {
    case 1: return B ? launch_fwd_sh<1, true>(a, exact, st) : launch_fwd_sh<1, false>(a, exact, st);
    case 2: return B ? launch_fwd_sh<2, true>(a, exact, st) : launch_fwd_sh<2, false>(a, exact, st);
    case 3: return B ? launch_fwd_sh<3, true>(a, exact, st) : launch_fwd_sh<3, false>(a, exact, st);
    default: return B ? launch_fwd_sh<4, true>(a, exact, st) : launch_fwd_sh<4, false>(a, exact, st);
  }
}

extern "C" size_t gsmi_backward_order_workspace_size(int n_tiles_total) { return order_ws_bytes(n_tiles_total); }

extern "C" int gsmi_render_sh_backward(int64_t V, int64_t P, const float* mean, const float* cov, const float* sh,
                                       int64_t sh_stride_g, int64_t sh_stride_c, const int32_t* sh_index,
                                       const float* alpha, const int32_t* start, const int32_t* end,
                                       const int32_t* gaussian_ids, const float* out, const float* T_final,
                                       const int32_t* n_done, const float* pre_bg, const float* grad_out,
                                       float* grad_mean, float* grad_cov, float* grad_sh, float* grad_alpha,
                                       const float* topleft, const float* c2w, int tile_size, int n_tiles_h,
                                       int n_tiles_w, float pixel_size_x, float pixel_size_y, int H, int W, int C,
                                       float thresh, const float* bg, int flags, void* workspace,
                                       size_t workspace_bytes, void* stream) {
  int rc = check_common(V, P, tile_size, n_tiles_h, n_tiles_w, H, W);
  if (rc) return rc;
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (!start || !end || !topleft || !c2w || !grad_out) return GSMI_ERR_ARG;
  if (P > 0 && (!mean || !cov || !sh || !alpha || !gaussian_ids || !grad_mean || !grad_cov || !grad_sh ||
                !grad_alpha))
    return GSMI_ERR_ARG;
  bool recomp;
  rc = check_bwd_state(out, T_final, n_done, pre_bg, bg, flags, workspace, &recomp);
  if (rc) return rc;
  if (sh_stride_c < C * C || sh_stride_g < 3 * sh_stride_c) return GSMI_ERR_ARG;
  if (((uintptr_t)mean & 7) || ((uintptr_t)cov & 15)) return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh = sh;
  a.sh_index = sh_index;
  a.alpha = alpha;
  a.start = start;
  a.end = end;
  a.ids = gaussian_ids;
  a.topleft = topleft;
  a.c2w = c2w;
  a.bg = bg;
  a.out_saved = out;
  a.T_saved = T_final;
  a.n_done_saved = n_done;
  a.pre_bg_saved = pre_bg;
  a.grad_out = grad_out;
  a.grad_mean = grad_mean;
  a.grad_cov = grad_cov;
  a.grad_sh = grad_sh;
  a.grad_alpha = grad_alpha;
  a.sh_sg = sh_stride_g;
  a.sh_sc = sh_stride_c;
  a.psx = pixel_size_x;
  a.psy = pixel_size_y;
  a.thresh = thresh;
  a.n_tiles_h = n_tiles_h;
  a.n_tiles_w = n_tiles_w;
  a.n_tiles_total = n_tiles_h * n_tiles_w;
  a.H = H;
  a.W = W;
  hipStream_t st = (hipStream_t)stream;
  if (workspace) {
    if (workspace_bytes < order_ws_bytes(a.n_tiles_total)) return GSMI_ERR_WORKSPACE;
    rc = make_tile_order(a, workspace, st);
    if (rc) return rc;
  }
  return dispatch_bwd_sh(a, C, flags, recomp, st);
}

// ---------------------------------------------------------------------------
// Backward without allocation: the tile order (costs, then the ordered tile
// indices) goes to a file-scope device array instead of a caller workspace.
// Each GPU holds its own copy of `scratch` at its own address; the address is
// looked up once per device (with that device current) and only that address
// is handed to kernels.  Every call on a GPU shares the array and the order is
// rebuilt in full by each call (every tile's cost and position is written,
// and the counter words of grids beyond 32768 tiles are cleared by a kernel
// first); calls that overlap on one GPU (other streams or threads) race on
// it.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxTiles = 1 << 18;  // 262,144 tiles: 8 views of 3840x2160
static __device__ __align__(256) uint8_t scratch[2 * kNoAllocMaxTiles * 4 + ((kOrderWords * 4 + 255) & ~255)];  // = order_ws_bytes(max)
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

// Device address of this device's `scratch`; null on failure.
static void* scratch_on_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

extern "C" int gsmi_render_sh_backward_noalloc(
    int64_t V, int64_t P, const float* mean, const float* cov, const float* sh, int64_t sh_stride_g,
    int64_t sh_stride_c, const int32_t* sh_index, const float* alpha, const int32_t* start, const int32_t* end,
    const int32_t* gaussian_ids, const float* out, const float* T_final, const int32_t* n_done, const float* pre_bg,
    const float* grad_out, float* grad_mean, float* grad_cov, float* grad_sh, float* grad_alpha,
    const float* topleft, const float* c2w, int tile_size, int n_tiles_h, int n_tiles_w, float pixel_size_x,
    float pixel_size_y, int H, int W, int C, float thresh, const float* bg, int flags, void* stream) {
  if (!T_final || !n_done) return GSMI_ERR_ARG;  // the order comes from the saved n_done
  if (n_tiles_h <= 0 || n_tiles_w <= 0 || (int64_t)n_tiles_h * n_tiles_w > kNoAllocMaxTiles)
    return n_tiles_h <= 0 || n_tiles_w <= 0 ? GSMI_ERR_ARG : GSMI_ERR_WORKSPACE;
  void* ws = scratch_on_current_device();
  if (!ws) return GSMI_ERR_LAUNCH;
  return gsmi_render_sh_backward(V, P, mean, cov, sh, sh_stride_g, sh_stride_c, sh_index, alpha, start, end,
                                 gaussian_ids, out, T_final, n_done, pre_bg, grad_out, grad_mean, grad_cov, grad_sh,
                                 grad_alpha, topleft, c2w, tile_size, n_tiles_h, n_tiles_w, pixel_size_x,
                                 pixel_size_y, H, W, C, thresh, bg, flags, ws, sizeof(scratch), stream);
}

// The reference's own signatures (render.h:86-101, 119-133): sh_coeffs is the
// contiguous [V, 3, C*C] slice, out the flat [H*W*3] image the caller zeroed,
// and the backward recomputes the forward state itself.
extern "C" int gsmi_tile_based_vol_rendering_sh(int64_t V, int64_t P, const float* mean, const float* cov,
                                                const float* sh_coeffs, const float* alpha, const int32_t* start,
                                                const int32_t* end, const int32_t* gaussian_ids, float* out,
                                                const float* topleft, const float* c2w, int tile_size,
                                                int n_tiles_h, int n_tiles_w, float pixel_size_x,
                                                float pixel_size_y, int H, int W, int C, float thresh,
                                                void* stream) {
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  return gsmi_render_sh_forward(V, P, mean, cov, sh_coeffs, 3 * C * C, C * C, nullptr, alpha, start, end,
                                gaussian_ids, out, topleft, c2w, tile_size, n_tiles_h, n_tiles_w, pixel_size_x,
                                pixel_size_y, H, W, C, thresh, nullptr, nullptr, nullptr, nullptr, 0, 0, stream);
}

extern "C" int gsmi_tile_based_vol_rendering_sh_with_bg(int64_t V, int64_t P, const float* mean, const float* cov,
                                                        const float* sh_coeffs, const float* alpha,
                                                        const int32_t* start, const int32_t* end,
                                                        const int32_t* gaussian_ids, float* out,
                                                        const float* topleft, const float* c2w, int tile_size,
                                                        int n_tiles_h, int n_tiles_w, float pixel_size_x,
                                                        float pixel_size_y, int H, int W, int C, float thresh,
                                                        const float* bg_rgb, void* stream) {
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (!bg_rgb) return GSMI_ERR_ARG;
  return gsmi_render_sh_forward(V, P, mean, cov, sh_coeffs, 3 * C * C, C * C, nullptr, alpha, start, end,
                                gaussian_ids, out, topleft, c2w, tile_size, n_tiles_h, n_tiles_w, pixel_size_x,
                                pixel_size_y, H, W, C, thresh, bg_rgb, nullptr, nullptr, nullptr, 0, 0, stream);
}

extern "C" int gsmi_tile_based_vol_rendering_backward_sh(
    int64_t V, int64_t P, const float* mean, const float* cov, const float* sh_coeffs, const float* alpha,
    const int32_t* start, const int32_t* end, const int32_t* gaussian_ids, const float* out, float* grad_mean,
    float* grad_cov, float* grad_sh_coeffs, float* grad_alpha, const float* grad_out, const float* topleft,
    const float* c2w, int tile_size, int n_tiles_h, int n_tiles_w, float pixel_size_x, float pixel_size_y, int H,
    int W, int C, float thresh, void* stream) {
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  return gsmi_render_sh_backward(V, P, mean, cov, sh_coeffs, 3 * C * C, C * C, nullptr, alpha, start, end,
                                 gaussian_ids, out, nullptr, nullptr, nullptr, grad_out, grad_mean, grad_cov,
                                 grad_sh_coeffs, grad_alpha, topleft, c2w, tile_size, n_tiles_h, n_tiles_w,
                                 pixel_size_x, pixel_size_y, H, W, C, thresh, nullptr, 0, nullptr, 0, stream);
}

extern "C" int gsmi_tile_based_vol_rendering_backward_sh_with_bg(
    int64_t V, int64_t P, const float* mean, const float* cov, const float* sh_coeffs, const float* alpha,
    const int32_t* start, const int32_t* end, const int32_t* gaussian_ids, const float* out, float* grad_mean,
    float* grad_cov, float* grad_sh_coeffs, float* grad_alpha, const float* grad_out, const float* topleft,
    const float* c2w, int tile_size, int n_tiles_h, int n_tiles_w, float pixel_size_x, float pixel_size_y, int H,
    int W, int C, float thresh, const float* bg_rgb, void* stream) {
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (!bg_rgb) return GSMI_ERR_ARG;
  return gsmi_render_sh_backward(V, P, mean, cov, sh_coeffs, 3 * C * C, C * C, nullptr, alpha, start, end,
                                 gaussian_ids, out, nullptr, nullptr, nullptr, grad_out, grad_mean, grad_cov,
                                 grad_sh_coeffs, grad_alpha, topleft, c2w, tile_size, n_tiles_h, n_tiles_w,
                                 pixel_size_x, pixel_size_y, H, W, C, thresh, bg_rgb, 0, nullptr, 0, stream);
}

extern "C" int gsmi_render_rgb_forward(int64_t V, int64_t P, const float* mean, const float* cov,
                                       const float* color, const float* alpha, const int32_t* start,
                                       const int32_t* end, const int32_t* gaussian_ids, float* out,
                                       const float* topleft, int tile_size, int n_tiles_h, int n_tiles_w,
                                       float pixel_size_x, float pixel_size_y, int H, int W, float thresh,
                                       float* T_final, int32_t* n_done, int flags, void* stream) {
  int rc = check_common(V, P, tile_size, n_tiles_h, n_tiles_w, H, W);
  if (rc) return rc;
  if (!start || !end || !out || !topleft) return GSMI_ERR_ARG;
  if (P > 0 && (!mean || !cov || !color || !alpha || !gaussian_ids)) return GSMI_ERR_ARG;
  if (((uintptr_t)mean & 7) || ((uintptr_t)cov & 15)) return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh = color;
  a.alpha = alpha;
  a.start = start;
  a.end = end;
  a.ids = gaussian_ids;
  a.topleft = topleft;
  a.out = out;
  a.T_final = T_final;
  a.n_done = n_done;
  a.sh_sg = 3;
  a.sh_sc = 0;
  a.psx = pixel_size_x;
  a.psy = pixel_size_y;
  a.thresh = thresh;
  a.n_tiles_h = n_tiles_h;
  a.n_tiles_w = n_tiles_w;
  a.n_tiles_total = n_tiles_h * n_tiles_w;
  a.H = H;
  a.W = W;
  a.write_empty = flags & 1;
  hipLaunchKernelGGL(k_fwd_rgb, dim3(xcd_grid(n_tiles_h * n_tiles_w)), dim3(256), 0, (hipStream_t)stream, a);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_render_rgb_backward(int64_t V, int64_t P, const float* mean, const float* cov,
                                        const float* color, const float* alpha, const int32_t* start,
                                        const int32_t* end, const int32_t* gaussian_ids, const float* out,
                                        const float* T_final, const int32_t* n_done, const float* grad_out,
                                        float* grad_mean, float* grad_cov, float* grad_color, float* grad_alpha,
                                        const float* topleft, int tile_size, int n_tiles_h, int n_tiles_w,
                                        float pixel_size_x, float pixel_size_y, int H, int W, float thresh,
                                        void* stream) {
  int rc = check_common(V, P, tile_size, n_tiles_h, n_tiles_w, H, W);
  if (rc) return rc;
  if (!start || !end || !topleft || !grad_out) return GSMI_ERR_ARG;
  if ((T_final == nullptr) != (n_done == nullptr)) return GSMI_ERR_ARG;
  if (P > 0 && (!mean || !cov || !color || !alpha || !gaussian_ids || !grad_mean || !grad_cov || !grad_color ||
                !grad_alpha))
    return GSMI_ERR_ARG;
  (void)out;  // the adjoint needs no output colour without a background
  if (((uintptr_t)mean & 7) || ((uintptr_t)cov & 15)) return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh = color;
  a.alpha = alpha;
  a.start = start;
  a.end = end;
  a.ids = gaussian_ids;
  a.topleft = topleft;
  a.T_saved = T_final;
  a.n_done_saved = n_done;
  a.grad_out = grad_out;
  a.grad_mean = grad_mean;
  a.grad_cov = grad_cov;
  a.grad_sh = grad_color;
  a.grad_alpha = grad_alpha;
  a.sh_sg = 3;
  a.sh_sc = 0;
  a.psx = pixel_size_x;
  a.psy = pixel_size_y;
  a.thresh = thresh;
  a.n_tiles_h = n_tiles_h;
  a.n_tiles_w = n_tiles_w;
  a.n_tiles_total = n_tiles_h * n_tiles_w;
  a.H = H;
  a.W = W;
  const dim3 grid(xcd_grid(n_tiles_h * n_tiles_w));
  if (T_final)
    hipLaunchKernelGGL(k_bwd_rgb<false>, grid, dim3(256), 0, (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL(k_bwd_rgb<true>, grid, dim3(256), 0, (hipStream_t)stream, a);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

// The reference's own signatures (render.h:70-84).
extern "C" int gsmi_tile_based_vol_rendering_start_end(int64_t V, int64_t P, const float* mean, const float* cov,
                                                       const float* color, const float* alpha,
                                                       const int32_t* start, const int32_t* end,
                                                       const int32_t* gaussian_ids, float* out,
                                                       const float* topleft, int tile_size, int n_tiles_h,
                                                       int n_tiles_w, float pixel_size_x, float pixel_size_y,
                                                       int H, int W, float thresh, void* stream) {
  return gsmi_render_rgb_forward(V, P, mean, cov, color, alpha, start, end, gaussian_ids, out, topleft, tile_size,
                                 n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, thresh, nullptr, nullptr,
                                 0, stream);
}

extern "C" int gsmi_tile_based_vol_rendering_backward_start_end(
    int64_t V, int64_t P, const float* mean, const float* cov, const float* color, const float* alpha,
    const int32_t* start, const int32_t* end, const int32_t* gaussian_ids, const float* out, float* grad_mean,
    float* grad_cov, float* grad_color, float* grad_alpha, const float* grad_out, const float* topleft,
    int tile_size, int n_tiles_h, int n_tiles_w, float pixel_size_x, float pixel_size_y, int H, int W,
    float thresh, void* stream) {
  return gsmi_render_rgb_backward(V, P, mean, cov, color, alpha, start, end, gaussian_ids, out, nullptr, nullptr,
                                  grad_out, grad_mean, grad_cov, grad_color, grad_alpha, topleft, tile_size,
                                  n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, thresh, stream);
}

// ---------------------------------------------------------------------------
// Multi-view batches: one launch over the tiles of every view (tile ids
// numbered with each view's tile_base, pixels with its pix_base).  The
// per-pixel arithmetic is the single-view kernels', so every view's image is
// bit-identical to rendering it alone.
// ---------------------------------------------------------------------------
static int check_views(int64_t V, int64_t P, int B, const gsmi_view* views, int n_tiles_total) {
  if (V < 0 || P < 0 || B <= 0 || !views || n_tiles_total <= 0) return GSMI_ERR_ARG;
  if ((uintptr_t)views & 7) return GSMI_ERR_ARG;
  if (n_tiles_total > (1 << 30)) return GSMI_ERR_UNSUPPORTED;
  return GSMI_OK;
}

extern "C" int gsmi_render_views_sh(int64_t V, int64_t P, int B, const gsmi_view* views, int n_tiles_total,
                                    const float* mean, const float* cov, const float* sh, int64_t sh_stride_g,
                                    int64_t sh_stride_c, const int32_t* sh_index, const float* alpha,
                                    const int32_t* start, const int32_t* end, const int32_t* gaussian_ids,
                                    float* out, int C, float thresh, const float* bg, float* T_final,
                                    int32_t* n_done, float* pre_bg, int flags, int exact, void* stream) {
  int rc = check_views(V, P, B, views, n_tiles_total);
  if (rc) return rc;
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (!start || !end || !out) return GSMI_ERR_ARG;
  if (P > 0 && (!mean || !cov || !sh || !alpha || !gaussian_ids)) return GSMI_ERR_ARG;
  if (sh_stride_c < C * C || sh_stride_g < 3 * sh_stride_c) return GSMI_ERR_ARG;
  if (((uintptr_t)mean & 7) || ((uintptr_t)cov & 15)) return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh = sh;
  a.sh_index = sh_index;
  a.alpha = alpha;
  a.start = start;
  a.end = end;
  a.ids = gaussian_ids;
  a.bg = bg;
  a.out = out;
  a.T_final = T_final;
  a.n_done = n_done;
  a.pre_bg = pre_bg;
  a.sh_sg = sh_stride_g;
  a.sh_sc = sh_stride_c;
  a.thresh = thresh;
  a.write_empty = flags & 1;
  a.views = views;
  a.n_views = B;
  a.n_tiles_total = n_tiles_total;
  hipStream_t st = (hipStream_t)stream;
  const bool G = bg != nullptr;
  switch (C) {
    case 1: return G ? launch_fwd_sh<1, true>(a, exact, st) : launch_fwd_sh<1, false>(a, exact, st);
    case 2: return G ? launch_fwd_sh<2, true>(a, exact, st) : launch_fwd_sh<2, false>(a, exact, st);
    case 3: return G ? launch_fwd_sh<3, true>(a, exact, st) : launch_fwd_sh<3, false>(a, exact, st);
    default: return G ? launch_fwd_sh<4, true>(a, exact, st) : launch_fwd_sh<4, false>(a, exact, st);
  }
}

// The multi-view adjoint over shared lists (gaussian_ids) or per-view lists
// (view_ids: device array of B pointers; start/end relative to each).
static int views_backward_impl(int64_t V, int64_t P, int B, const gsmi_view* views, int n_tiles_total,
                               const float* mean, const float* cov, const float* sh, int64_t sh_stride_g,
                               int64_t sh_stride_c, const int32_t* sh_index, const float* alpha, const int32_t* start,
                               const int32_t* end, const int32_t* gaussian_ids, const int32_t* const* view_ids,
                               const float* out, const float* T_final, const int32_t* n_done, const float* pre_bg,
                               const float* grad_out, float* grad_mean, float* grad_cov, float* grad_sh,
                               float* grad_alpha, int C, float thresh, const float* bg, int flags, void* workspace,
                               size_t workspace_bytes, hipStream_t st) {
  int rc = check_views(V, P, B, views, n_tiles_total);
  if (rc) return rc;
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (!start || !end || !grad_out) return GSMI_ERR_ARG;
  if (P > 0 && (!mean || !cov || !sh || !alpha || !(gaussian_ids || view_ids) || !grad_mean || !grad_cov ||
                !grad_sh || !grad_alpha))
    return GSMI_ERR_ARG;
  bool recomp;
  rc = check_bwd_state(out, T_final, n_done, pre_bg, bg, flags, workspace, &recomp);
  if (rc) return rc;
  if (sh_stride_c < C * C || sh_stride_g < 3 * sh_stride_c) return GSMI_ERR_ARG;
  if (((uintptr_t)mean & 7) || ((uintptr_t)cov & 15)) return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh = sh;
  a.sh_index = sh_index;
  a.alpha = alpha;
  a.start = start;
  a.end = end;
  a.ids = gaussian_ids;
  a.view_ids = view_ids;
  a.bg = bg;
  a.out_saved = out;
  a.T_saved = T_final;
  a.n_done_saved = n_done;
  a.pre_bg_saved = pre_bg;
  a.grad_out = grad_out;
  a.grad_mean = grad_mean;
  a.grad_cov = grad_cov;
  a.grad_sh = grad_sh;
  a.grad_alpha = grad_alpha;
  a.sh_sg = sh_stride_g;
  a.sh_sc = sh_stride_c;
  a.thresh = thresh;
  a.views = views;
  a.n_views = B;
  a.n_tiles_total = n_tiles_total;
  if (workspace) {
    if (workspace_bytes < order_ws_bytes(n_tiles_total)) return GSMI_ERR_WORKSPACE;
    rc = make_tile_order(a, workspace, st);
    if (rc) return rc;
  }
  return dispatch_bwd_sh(a, C, flags, recomp, st);
}

extern "C" int gsmi_render_views_backward_sh(int64_t V, int64_t P, int B, const gsmi_view* views,
                                             int n_tiles_total, const float* mean, const float* cov,
                                             const float* sh, int64_t sh_stride_g, int64_t sh_stride_c,
                                             const int32_t* sh_index, const float* alpha, const int32_t* start,
                                             const int32_t* end, const int32_t* gaussian_ids, const float* out,
                                             const float* T_final, const int32_t* n_done, const float* pre_bg,
                                             const float* grad_out, float* grad_mean, float* grad_cov,
                                             float* grad_sh, float* grad_alpha, int C, float thresh,
                                             const float* bg, int flags, void* workspace, size_t workspace_bytes,
                                             void* stream) {
  return views_backward_impl(V, P, B, views, n_tiles_total, mean, cov, sh, sh_stride_g, sh_stride_c, sh_index, alpha,
                             start, end, gaussian_ids, nullptr, out, T_final, n_done, pre_bg, grad_out, grad_mean,
                             grad_cov, grad_sh, grad_alpha, C, thresh, bg, flags, workspace, workspace_bytes,
                             (hipStream_t)stream);
}

extern "C" int gsmi_render_views_backward_sh_lists(int64_t V, int64_t P, int B, const gsmi_view* views,
                                                   int n_tiles_total, const float* mean, const float* cov,
                                                   const float* sh, int64_t sh_stride_g, int64_t sh_stride_c,
                                                   const int32_t* sh_index, const float* alpha, const int32_t* start,
                                                   const int32_t* end, const int32_t* const* view_ids,
                                                   const float* T_final, const int32_t* n_done,
                                                   const float* grad_out, float* grad_mean, float* grad_cov,
                                                   float* grad_sh, float* grad_alpha, int C, float thresh, int flags,
                                                   void* workspace, size_t workspace_bytes, void* stream) {
  if (!view_ids || !T_final || !n_done) return GSMI_ERR_ARG;
  return views_backward_impl(V, P, B, views, n_tiles_total, mean, cov, sh, sh_stride_g, sh_stride_c, sh_index, alpha,
                             start, end, nullptr, view_ids, nullptr, T_final, n_done, nullptr, grad_out, grad_mean,
                             grad_cov, grad_sh, grad_alpha, C, thresh, nullptr, flags, workspace, workspace_bytes,
                             (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Two-level lists (single view, no background): per-tile offsets of the
// recorded filtered lists, and the forward that fills them.
// ---------------------------------------------------------------------------
namespace gsmi {
// Capacity of tile t's recorded list: the length of its block's list.
__global__ void k_tile_caps(int n_tiles_h, int n_tiles_w, int shift, int n_super_w, const int32_t* __restrict__ sstart,
                            const int32_t* __restrict__ send, int32_t* __restrict__ cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles_h * n_tiles_w) return;
  const int tx = t % n_tiles_w, ty = t / n_tiles_w;
  const int sb = (ty >> shift) * n_super_w + (tx >> shift);
  const int s0 = sstart[sb], e0 = send[sb];
  cap[t] = (s0 == -1 || e0 <= s0) ? 0 : e0 - s0;
}

// Per tile of a multi-view batch: the length of its block's list.
__global__ void k_tile_caps_views(const gsmi_view* __restrict__ views, const gsmi_view* __restrict__ bviews, int B,
                                  int n_tiles_total, int shift, const int32_t* __restrict__ bstart,
                                  const int32_t* __restrict__ bend, int32_t* __restrict__ cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles_total) return;
  const int v = view_of_tile(views, B, t);
  const int local = t - views[v].tile_base;
  const int ntw = views[v].n_tiles_w;
  const int tx = local % ntw, ty = local / ntw;
  const int sb = bviews[v].tile_base + (ty >> shift) * bviews[v].n_tiles_w + (tx >> shift);
  const int s0 = bstart[sb], e0 = bend[sb];
  cap[t] = (s0 == -1 || e0 <= s0) ? 0 : e0 - s0;
}
}  // namespace gsmi

__global__ void k_zero_work(int* w) {
  if (threadIdx.x < 8) w[threadIdx.x] = 0;
}

// grid_cap > 0 with a.work: a work-queue launch of at most grid_cap workgroups.
static int launch_fwd_two(const RasterArgs& a, int C, int exact, hipStream_t st, int grid_cap = 0) {
  int grid = xcd_grid(a.n_tiles_total);
  const bool queued = a.work != nullptr;
  const bool pv = a.view_sh || a.view_out || a.view_ids || a.view_fids;
  if (queued) {
    hipLaunchKernelGGL(k_zero_work, dim3(1), dim3(64), 0, st, a.work);  // a fill kernel, not a memset node
    grid = std::min(grid, std::max(grid_cap, 8));
  }
#define GSMI_FWD2(CC_)                                                                                     \
  do {                                                                                                      \
    if (queued) {                                                                                           \
      if (exact)                                                                                            \
        hipLaunchKernelGGL((k_fwd_sh_queued<CC_, true>), dim3(grid), dim3(256), 0, st, a);                  \
      else                                                                                                  \
        hipLaunchKernelGGL((k_fwd_sh_queued<CC_, false>), dim3(grid), dim3(256), 0, st, a);                 \
    } else if (pv) {                                                                                        \
      if (exact)                                                                                            \
        hipLaunchKernelGGL((k_fwd_sh<CC_, false, true, true, true>), dim3(grid), dim3(256), 0, st, a);      \
      else                                                                                                  \
        hipLaunchKernelGGL((k_fwd_sh<CC_, false, false, true, true>), dim3(grid), dim3(256), 0, st, a);     \
    } else {                                                                                                \
      if (exact)                                                                                            \
        hipLaunchKernelGGL((k_fwd_sh<CC_, false, true, true>), dim3(grid), dim3(256), 0, st, a);            \
      else                                                                                                  \
        hipLaunchKernelGGL((k_fwd_sh<CC_, false, false, true>), dim3(grid), dim3(256), 0, st, a);           \
    }                                                                                                       \
  } while (0)
  switch (C) {
    case 1: GSMI_FWD2(1); break;
    case 2: GSMI_FWD2(2); break;
    case 3: GSMI_FWD2(3); break;
    default: GSMI_FWD2(4); break;
  }
#undef GSMI_FWD2
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_two_level_offsets_views(int B, const gsmi_view* views, const gsmi_view* block_views,
                                            const int32_t* view_tiles, int shift, const int32_t* bstart,
                                            const int32_t* bend, int32_t* foff, void* workspace,
                                            size_t workspace_bytes, void* stream) {
  if (B <= 0 || !views || !block_views || !view_tiles || shift < 0 || shift > 4 || !bstart || !bend || !foff)
    return GSMI_ERR_ARG;
  int64_t n = 0;
  int n_max = 0;
  for (int b = 0; b < B; ++b) {
    if (view_tiles[b] <= 0) return GSMI_ERR_ARG;
    n += view_tiles[b];
    n_max = std::max(n_max, (int)view_tiles[b]);
  }
  if (n > (1 << 30)) return GSMI_ERR_UNSUPPORTED;
  if (!workspace || workspace_bytes < gsmi_two_level_offsets_workspace_size(n_max)) return GSMI_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_tile_caps_views, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, views, block_views, B,
                     (int)n, shift, bstart, bend, foff);
  GSMI_CHECK_LAUNCH();
  int64_t base = 0;  // each view's offsets start at 0 in its own list array
  for (int b = 0; b < B; ++b) {
    const int rc = exclusive_scan<int32_t>(foff + base, foff + base, view_tiles[b], nullptr, workspace, st);
    if (rc) return rc;
    base += view_tiles[b];
  }
  return GSMI_OK;
}

extern "C" int gsmi_render_views_sh_two_level_record(
    int64_t V, int B, const gsmi_view* views, const gsmi_view* block_views, int n_tiles_total, const float* mean,
    const float* cov, const float* sh, int64_t sh_stride_g, int64_t sh_stride_c, const int32_t* sh_index,
    const float* alpha, const int32_t* tl, const int32_t* br, const int32_t* bstart, const int32_t* bend,
    const int32_t* bids, int shift, const int32_t* foff, int32_t* const* view_fids, int32_t* fend, float* out, int C,
    float thresh, float* T_final, int32_t* n_done, int exact, void* stream) {
  int rc = check_views(V, 0, B, views, n_tiles_total);
  if (rc) return rc;
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (shift < 0 || shift > 4) return GSMI_ERR_ARG;
  if (!block_views || ((uintptr_t)block_views & 7) || !bstart || !bend || !out || !foff || !view_fids || !fend)
    return GSMI_ERR_ARG;
  if (V > 0 && (!mean || !cov || !sh || !alpha || !tl || !br || !bids)) return GSMI_ERR_ARG;
  if (sh_stride_c < C * C || sh_stride_g < 3 * sh_stride_c) return GSMI_ERR_ARG;
  if (((uintptr_t)mean & 7) || ((uintptr_t)cov & 15) || ((uintptr_t)tl & 7) || ((uintptr_t)br & 7))
    return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh = sh;
  a.sh_index = sh_index;
  a.alpha = alpha;
  a.out = out;
  a.T_final = T_final;
  a.n_done = n_done;
  a.sh_sg = sh_stride_g;
  a.sh_sc = sh_stride_c;
  a.thresh = thresh;
  a.write_empty = 1;
  a.views = views;
  a.n_views = B;
  a.n_tiles_total = n_tiles_total;
  a.bviews = block_views;
  a.sstart = bstart;
  a.send = bend;
  a.sids = bids;
  a.rtl = tl;
  a.rbr = br;
  a.super_shift = shift;
  a.foff = foff;
  a.view_fids = view_fids;
  a.fend = fend;
  return launch_fwd_two(a, C, exact, (hipStream_t)stream);
}

// Forward of a render-server batch (declared in serve_internal.h): as
// gsmi_render_views_sh_two_level with each view's SH rows at view_sh[b] and
// image written to view_out[b] ([H, W, 3]), no T_final / n_done; with work
// (device int[8]) and grid_cap > 0 the tiles are taken from per-XCD work
// queues by at most grid_cap workgroups.
int gsmi_internal_render_served(int64_t V, int B, const gsmi_view* views, const gsmi_view* block_views,
                                  int n_tiles_total, const float* mean, const float* cov,
                                  const float* const* view_sh, int64_t sh_stride_g, int64_t sh_stride_c,
                                  const int32_t* sh_index, const float* alpha, const int32_t* tl, const int32_t* br,
                                  const int32_t* bstart, const int32_t* bend, const int32_t* bids, int shift,
                                  float* const* view_out, int C, float thresh, int exact, int* work, int grid_cap,
                                  hipStream_t st) {
  int rc = check_views(V, 0, B, views, n_tiles_total);
  if (rc) return rc;
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (shift < 0 || shift > 4 || !block_views || !bstart || !bend || !view_sh || !view_out) return GSMI_ERR_ARG;
  if (V > 0 && (!mean || !cov || !alpha || !tl || !br || !bids || !sh_index)) return GSMI_ERR_ARG;
  if (sh_stride_c < C * C || sh_stride_g < 3 * sh_stride_c) return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh_index = sh_index;
  a.alpha = alpha;
  a.sh_sg = sh_stride_g;
  a.sh_sc = sh_stride_c;
  a.thresh = thresh;
  a.write_empty = 1;
  a.views = views;
  a.n_views = B;
  a.n_tiles_total = n_tiles_total;
  a.bviews = block_views;
  a.sstart = bstart;
  a.send = bend;
  a.sids = bids;
  a.rtl = tl;
  a.rbr = br;
  a.super_shift = shift;
  a.view_sh = view_sh;
  a.view_out = view_out;
  a.work = grid_cap > 0 ? work : nullptr;
  return launch_fwd_two(a, C, exact, st, grid_cap);
}

extern "C" size_t gsmi_two_level_offsets_workspace_size(int n_tiles) {
  return scan_workspace_bytes(n_tiles < 0 ? 0 : n_tiles, 4);
}

extern "C" int gsmi_two_level_offsets(int n_tiles_h, int n_tiles_w, int shift, const int32_t* sstart,
                                      const int32_t* send, int32_t* foff, void* workspace, size_t workspace_bytes,
                                      void* stream) {
  if (n_tiles_h <= 0 || n_tiles_w <= 0 || shift < 0 || shift > 4 || !sstart || !send || !foff) return GSMI_ERR_ARG;
  const int n = n_tiles_h * n_tiles_w;
  if (!workspace || workspace_bytes < gsmi_two_level_offsets_workspace_size(n)) return GSMI_ERR_WORKSPACE;
  const int n_super_w = (n_tiles_w + (1 << shift) - 1) >> shift;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_tile_caps, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n_tiles_h, n_tiles_w, shift,
                     n_super_w, sstart, send, foff);
  GSMI_CHECK_LAUNCH();
  return exclusive_scan<int32_t>(foff, foff, n, nullptr, workspace, st);
}

extern "C" int gsmi_render_sh_two_level(int64_t V, const float* mean, const float* cov, const float* sh,
                                        int64_t sh_stride_g, int64_t sh_stride_c, const int32_t* sh_index,
                                        const float* alpha, const int32_t* tl, const int32_t* br,
                                        const int32_t* sstart, const int32_t* send, const int32_t* sids,
                                        int shift, const int32_t* foff, int32_t* fids, int32_t* fend, float* out,
                                        const float* topleft, const float* c2w, int n_tiles_h, int n_tiles_w,
                                        float pixel_size_x, float pixel_size_y, int H, int W, int C, float thresh,
                                        float* T_final, int32_t* n_done, int exact, void* stream) {
  int rc = check_common(V, 0, kTile, n_tiles_h, n_tiles_w, H, W);
  if (rc) return rc;
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (shift < 0 || shift > 4) return GSMI_ERR_ARG;
  if (!sstart || !send || !foff || !fend || !out || !topleft || !c2w) return GSMI_ERR_ARG;
  if (V > 0 && (!mean || !cov || !sh || !alpha || !tl || !br || !sids || !fids)) return GSMI_ERR_ARG;
  if (sh_stride_c < C * C || sh_stride_g < 3 * sh_stride_c) return GSMI_ERR_ARG;
  if (((uintptr_t)mean & 7) || ((uintptr_t)cov & 15) || ((uintptr_t)tl & 7) || ((uintptr_t)br & 7))
    return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh = sh;
  a.sh_index = sh_index;
  a.alpha = alpha;
  a.topleft = topleft;
  a.c2w = c2w;
  a.out = out;
  a.T_final = T_final;
  a.n_done = n_done;
  a.sh_sg = sh_stride_g;
  a.sh_sc = sh_stride_c;
  a.psx = pixel_size_x;
  a.psy = pixel_size_y;
  a.thresh = thresh;
  a.n_tiles_h = n_tiles_h;
  a.n_tiles_w = n_tiles_w;
  a.n_tiles_total = n_tiles_h * n_tiles_w;
  a.H = H;
  a.W = W;
  a.write_empty = 1;
  a.sstart = sstart;
  a.send = send;
  a.sids = sids;
  a.rtl = tl;
  a.rbr = br;
  a.foff = foff;
  a.fids = fids;
  a.fend = fend;
  a.super_shift = shift;
  a.n_super_w = (n_tiles_w + (1 << shift) - 1) >> shift;
  return launch_fwd_two(a, C, exact, (hipStream_t)stream);
}

extern "C" int gsmi_render_views_sh_two_level(int64_t V, int B, const gsmi_view* views,
                                              const gsmi_view* block_views, int n_tiles_total, const float* mean,
                                              const float* cov, const float* sh, int64_t sh_stride_g,
                                              int64_t sh_stride_c, const int32_t* sh_index, const float* alpha,
                                              const int32_t* tl, const int32_t* br, const int32_t* bstart,
                                              const int32_t* bend, const int32_t* bids, int shift, float* out,
                                              int C, float thresh, float* T_final, int32_t* n_done, int exact,
                                              void* stream) {
  int rc = check_views(V, 0, B, views, n_tiles_total);
  if (rc) return rc;
  if (C < 1 || C > kMaxC) return GSMI_ERR_UNSUPPORTED;
  if (shift < 0 || shift > 4) return GSMI_ERR_ARG;
  if (!block_views || ((uintptr_t)block_views & 7) || !bstart || !bend || !out) return GSMI_ERR_ARG;
  if (V > 0 && (!mean || !cov || !sh || !alpha || !tl || !br || !bids)) return GSMI_ERR_ARG;
  if (sh_stride_c < C * C || sh_stride_g < 3 * sh_stride_c) return GSMI_ERR_ARG;
  if (((uintptr_t)mean & 7) || ((uintptr_t)cov & 15) || ((uintptr_t)tl & 7) || ((uintptr_t)br & 7))
    return GSMI_ERR_ARG;
  RasterArgs a{};
  a.mean = mean;
  a.cov = cov;
  a.sh = sh;
  a.sh_index = sh_index;
  a.alpha = alpha;
  a.out = out;
  a.T_final = T_final;
  a.n_done = n_done;
  a.sh_sg = sh_stride_g;
  a.sh_sc = sh_stride_c;
  a.thresh = thresh;
  a.write_empty = 1;
  a.views = views;
  a.n_views = B;
  a.n_tiles_total = n_tiles_total;
  a.bviews = block_views;
  a.sstart = bstart;
  a.send = bend;
  a.sids = bids;
  a.rtl = tl;
  a.rbr = br;
  a.super_shift = shift;
  return launch_fwd_two(a, C, exact, (hipStream_t)stream);
}

// Test hook: the forward's quotient q / det from the staged reciprocal
// (radial_staged with y = staged_recip(det)) and, when gauss is given, its
// Gaussian value (gauss_staged), element-wise.
__global__ void k_div_staged(int64_t n, const float* __restrict__ q, const float* __restrict__ det,
                             float* __restrict__ out, float* __restrict__ gauss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float qi = q[i], di = det[i];
  const float y = staged_recip(di);
  out[i] = radial_staged(qi, di, y, qi * y);
  if (gauss) gauss[i] = gauss_staged(qi, di, y, qi * y);
}

// The backward's launch order for given tile costs (test hook): order[0..n)
// as the backward computes it from its tile costs; workspace as for the
// backward's order.
extern "C" int gsmi_tile_order_from_costs(int n, const uint32_t* cost, int32_t* order, void* workspace,
                                          size_t workspace_bytes, void* stream) {
  if (n < 0 || (n > 0 && (!cost || !order || !workspace))) return GSMI_ERR_ARG;
  if (n == 0) return GSMI_OK;
  if (workspace_bytes < order_ws_bytes(n)) return GSMI_ERR_WORKSPACE;
  uint32_t* cw = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(workspace) + 2 * align_up((size_t)n * 4));
  place_tile_order(cost, n, order, cw, (hipStream_t)stream);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_div_staged(int64_t n, const float* q, const float* det, float* out, float* gauss,
                               void* stream) {
  if (n < 0 || (n > 0 && (!q || !det || !out))) return GSMI_ERR_ARG;
  if (n == 0) return GSMI_OK;
  hipLaunchKernelGGL(k_div_staged, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, q, det,
                     out, gauss);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}
