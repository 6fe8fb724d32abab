// Device-side arithmetic shared by the gfx950 kernels.
//
// The whole library is compiled with -ffp-contract=off: every expression is
// evaluated as written, one rounding per operation, and fused multiply-adds
// appear only where written explicitly (fmaf).  Functions whose results feed
// a compositing decision (cull, tile bounds, 1/255 skip, transmittance break)
// follow the reference source's operation order exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gsmi.h"

namespace gsmi {

constexpr int kTile = 16;
constexpr int kTilePixels = kTile * kTile;  // one 256-thread block per tile
constexpr int kMaxC = 4;

// common.h:89  MIN_RENDER_ALPHA = 1 / 255.0f
__device__ constexpr float kMinRenderAlpha = 1.0f / 255.0f;

// Binary32 exponential, specified algorithm: Cody-Waite reduction
// x = k*ln2 + r (ln2 split into a 12-bit hi part, so k*hi is exact, and a lo
// part), degree-7 Taylor polynomial of e^r in Horner form with fused
// multiply-adds, scaling by 2^k with ldexp.  Within 1 ulp of the true value.
__device__ __forceinline__ float exp_spec(float x) {
  if (x != x) return x;
  if (x > 89.0f) return __builtin_huge_valf();
  if (x < -110.0f) return 0.0f;
  const float kf = rintf(x * 0x1.715476p+0f);
  float r = fmaf(kf, -0x1.62ep-1f, x);
  r = fmaf(kf, -0x1.0bfbe8p-15f, r);
  float p = 0x1.a01a02p-13f;
  p = fmaf(p, r, 0x1.6c16c2p-10f);
  p = fmaf(p, r, 0x1.111112p-7f);
  p = fmaf(p, r, 0x1.555556p-5f);
  p = fmaf(p, r, 0x1.555556p-3f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)kf);
}

// exp_spec for x <= 0 or NaN (a Gaussian's exponent, -r/2 with r >= 0):
// the same value, without the overflow test and with the NaN test as a
// select.  Inputs below -120 (including -inf) are clamped first, so every
// intermediate is finite; exp_spec's underflow test (x < -110 -> 0) needs
// no select, since from -120 to -110 the scaled polynomial is below 2^-158
// and ldexp rounds it to 0 as well.
__device__ __forceinline__ float exp_spec_nonpos(float x) {
  const float xc = fmaxf(x, -120.0f);  // NaN -> -120; the result is replaced below
  const float kf = rintf(xc * 0x1.715476p+0f);
  float r = fmaf(kf, -0x1.62ep-1f, xc);
  r = fmaf(kf, -0x1.0bfbe8p-15f, r);
  float p = 0x1.a01a02p-13f;
  p = fmaf(p, r, 0x1.6c16c2p-10f);
  p = fmaf(p, r, 0x1.111112p-7f);
  p = fmaf(p, r, 0x1.555556p-5f);
  p = fmaf(p, r, 0x1.555556p-3f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  const float e = ldexpf(p, (int)kf);
  return x != x ? x : e;
}

// Binary64 exponential, same scheme (32-bit hi part of ln2, degree-13
// Taylor polynomial).
__device__ __forceinline__ double exp_spec_d(double x) {
  if (x != x) return x;
  if (x > 710.0) return __builtin_huge_val();
  if (x < -746.0) return 0.0;
  const double kd = rint(x * 0x1.71547652b82fep+0);
  double r = fma(kd, -0x1.62e42fee00000p-1, x);
  r = fma(kd, -0x1.a39ef35793c76p-33, r);
  double p = 0x1.6124613a86d09p-33;
  p = fma(p, r, 0x1.1eed8eff8d898p-29);
  p = fma(p, r, 0x1.ae64567f544e4p-26);
  p = fma(p, r, 0x1.27e4fb7789f5cp-22);
  p = fma(p, r, 0x1.71de3a556c734p-19);
  p = fma(p, r, 0x1.a01a01a01a01ap-16);
  p = fma(p, r, 0x1.a01a01a01a01ap-13);
  p = fma(p, r, 0x1.6c16c16c16c17p-10);
  p = fma(p, r, 0x1.1111111111111p-7);
  p = fma(p, r, 0x1.5555555555555p-5);
  p = fma(p, r, 0x1.5555555555555p-3);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)kd);
}

// Reciprocal of a Gaussian's determinant as staged for div_staged: RN(1/det)
// where 2^-40 <= |det| <= 2^100 (then a normal number), else NaN, which
// makes div_staged return NaN and its caller take the IEEE quotient.
__device__ __forceinline__ float staged_recip(float det) {
  const float ad = fabsf(det);
  return (ad >= 0x1p-40f && ad <= 0x1p100f) ? 1.0f / det : __builtin_nanf("");
}

// q / det from y = staged_recip(det) and q0 = q * y, by two residual steps:
// q0 is within 1.5 ulp of the quotient, r0 = q - det q0 (one FMA) corrects
// it to a faithful q1, and with y within half an ulp of 1/det the second
// step rounds correctly (Markstein's theorem; the remainder of a faithful
// quotient is exact in an FMA while |q| >= 2^-103, where its bits stay
// above the subnormal spacing).  So for |q| >= 2^-100 and finite inputs the
// result is the IEEE quotient or NaN (an overflowed intermediate, a NaN y);
// smaller |q| is left to the caller.
__device__ __forceinline__ float div_staged(float q, float det, float y, float q0) {
  const float r0 = fmaf(-q0, det, q);
  const float q1 = fmaf(r0, y, q0);
  const float r1 = fmaf(-q1, det, q);
  return fmaf(r1, y, q1);
}

// shencoder.h:4 SIGMOID with the specified exponential (exact colour mode).
__device__ __forceinline__ float sigmoid_exact(float x) { return 1.0f / (1.0f + exp_spec(-x)); }

// Fast sigmoid for colour values only (never feeds a decision): hardware
// base-2 exponential and reciprocal.
__device__ __forceinline__ float sigmoid_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(x * -1.44269504088896341f);
  return __builtin_amdgcn_rcpf(1.0f + e);
}

// The same sigmoid of x from x' = -log2(e) x (the SH coefficients staged
// times -log2(e) give x' directly as the colour's dot product).
constexpr float kNegLog2e = -1.44269504088896341f;
__device__ __forceinline__ float sigmoid_scaled(float xs) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(xs));
}

// float -> int32 with truncation toward zero, saturation and NaN -> 0.
__device__ __forceinline__ int32_t f2i_sat(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return 2147483647;
  if (v <= -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)v;
}

// Real SH basis for C = degree + 1 in 1..4 (shencoder.h:13-56): closed-form
// normalisations with the reference's sign convention.
template <int C>
__device__ __forceinline__ void sh_basis(float x, float y, float z, float* o) {
  const float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
  o[0] = 0.28209479177387814f;  // 1/(2 sqrt(pi))
  if constexpr (C > 1) {
    o[1] = -0.48860251190291987f * y;  // sqrt(3)/(2 sqrt(pi))
    o[2] = 0.48860251190291987f * z;
    o[3] = -0.48860251190291987f * x;
  }
  if constexpr (C > 2) {
    o[4] = 1.0925484305920792f * xy;  // sqrt(15)/(2 sqrt(pi))
    o[5] = -1.0925484305920792f * yz;
    o[6] = 0.94617469575755997f * z2 - 0.31539156525251999f;  // sqrt(5)(3z^2-1)/(4 sqrt(pi))
    o[7] = -1.0925484305920792f * xz;
    o[8] = 0.54627421529603959f * x2 - 0.54627421529603959f * y2;  // sqrt(15)/(4 sqrt(pi))
  }
  if constexpr (C > 3) {
    o[9] = 0.59004358992664352f * y * (-3.0f * x2 + y2);   // sqrt(70)/(8 sqrt(pi))
    o[10] = 2.8906114426405538f * xy * z;                  // sqrt(105)/(2 sqrt(pi))
    o[11] = 0.45704579946446572f * y * (1.0f - 5.0f * z2);  // sqrt(42)/(8 sqrt(pi))
    o[12] = 0.3731763325901154f * z * (5.0f * z2 - 3.0f);   // sqrt(7)/(4 sqrt(pi))
    o[13] = 0.45704579946446572f * x * (1.0f - 5.0f * z2);
    o[14] = 1.4453057213202769f * z * (x2 - y2);            // sqrt(105)/(4 sqrt(pi))
    o[15] = 0.59004358992664352f * x * (-x2 + 3.0f * y2);
  }
}

// Pixel ray direction as the reference computes it (vol_render_sh.h:48-65):
// the 3x4 c2w buffer is read as three consecutive float3 "rows".
__device__ __forceinline__ void pixel_direction(const float* m, float px, float py, float* d) {
  const float d0 = m[0] * px + m[1] * py + m[2] * 1.0f;
  const float d1 = m[3] * px + m[4] * py + m[5] * 1.0f;
  const float d2 = m[6] * px + m[7] * py + m[8] * 1.0f;
  const float len = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
  d[0] = d0 / len;
  d[1] = d1 / len;
  d[2] = d2 / len;
}

// XCD-aware block -> tile mapping.  Blocks are dispatched round-robin over
// the 8 XCDs; mapping block b to tile (b % 8) * per_xcd + b / 8 gives each
// XCD (and its L2) a contiguous band of tiles, which share Gaussians.
__device__ __forceinline__ int xcd_tile(int b, int n_tiles) {
  const int per_xcd = (n_tiles + 7) >> 3;
  return (b & 7) * per_xcd + (b >> 3);
}
inline int xcd_grid(int n_tiles) { return ((n_tiles + 7) >> 3) << 3; }

__device__ __forceinline__ int lane_id() { return __lane_id(); }
// Whether p holds in some active lane: a ballot of the condition mask
// itself (HIP's __any widens p to an int and compares it again per lane).
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

static_assert(sizeof(gsmi_view) == 256, "gsmi_view is a 256-byte record");

// View owning global tile `tile`: the largest b with views[b].tile_base <= tile.
__device__ __forceinline__ int view_of_tile(const gsmi_view* __restrict__ views, int n_views, int tile) {
  int lo = 0, hi = n_views - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (views[mid].tile_base <= tile) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace gsmi

// Launch helper: returns GSMI_ERR_LAUNCH on a launch failure.
#define GSMI_CHECK_LAUNCH()                                   \
  do {                                                        \
    if (hipGetLastError() != hipSuccess) return GSMI_ERR_LAUNCH; \
  } while (0)
