/*
 * gs_oracle.c — CPU restatement of the reference rasteriser's float arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the smoke
 * check in __graft_entry__.py and the CPU-baseline leg of bench.py may load it,
 * as the checker.  The product (gaussian_splatting_3d_amd/) never links, loads
 * or calls it.
 *
 * Arithmetic convention (the "oracle semantics", which the HIP kernels mirror):
 *   - IEEE binary32, round-to-nearest-even, expressions evaluated as written in
 *     the reference source, left to right, with NO multiply-add contraction
 *     (this file is compiled with -ffp-contract=off);
 *   - division and sqrt correctly rounded (x86-64 SSE);
 *   - expf(x) / exp(x) are the specified algorithms orc_expf / orc_exp_d
 *     below (own Cody-Waite + Taylor evaluation, within 1 ulp of the true
 *     value), which the HIP kernels evaluate with the identical operation
 *     sequence;
 *   - float -> int32 conversion truncates toward zero, saturates out of range
 *     and maps NaN to 0 (the conversion the reference's CUDA build performs in
 *     `tensor.to(torch.int32)`, utils/camera.py:301).
 * Gradient sums are accumulated in double here (the reference accumulates with
 * float atomics in an unspecified order, vol_render.h:28-46), then rounded.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_C 4

/* common.h:89 — MIN_RENDER_ALPHA = 1 / 255.0f */
static const float kMinRenderAlpha = 1.0f / 255.0f;

/*
 * Binary32 exponential, specified algorithm (the GPU kernels evaluate the
 * same operation sequence, so every compositing decision that depends on an
 * exponential is reproduced bit for bit).  Cody & Waite argument reduction
 * x = k ln2 + r with ln2 split as hi (12 significant bits, so k*hi is exact)
 * + lo, then the degree-7 Taylor polynomial of e^r in Horner form with fused
 * multiply-adds, then scaling by 2^k.  Pinned against the C library's expf
 * (tests/test_oracle_math.py).
 */
float orc_expf(float x) {
  if (x != x) return x;
  if (x > 89.0f) return INFINITY;
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

/*
 * Binary64 exponential, same scheme: ln2 split with a 32-bit hi part,
 * degree-13 Taylor polynomial.  Used where the reference evaluates the
 * Gaussian in double (kernels.h:195-224).
 */
double orc_exp_d(double x) {
  if (x != x) return x;
  if (x > 710.0) return INFINITY;
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

/* shencoder.h:4 — SIGMOID(x) = 1 / (1 + expf(-x)) */
static float orc_sigmoid(float x) { return 1.0f / (1.0f + orc_expf(-x)); }

/* CUDA-style float -> int32: trunc toward zero, saturate, NaN -> 0. */
int32_t orc_f2i_sat(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return (int32_t)v;
}

/*
 * Real spherical-harmonic basis for C = degree+1 in 1..4 (C*C outputs).
 * shencoder.h:13-56.  Constants are the closed-form normalisations
 *   Y00 = 1/(2 sqrt(pi)), Y1m = sqrt(3)/(2 sqrt(pi)) * {-y, z, -x}, ...
 * with the reference's sign convention; tests check each literal against its
 * closed form and against scipy's complex SH.
 */
void orc_sh_basis(const float *dir, float *o, int C) {
  float x = dir[0], y = dir[1], z = dir[2];
  float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
  o[0] = 0.28209479177387814f;
  if (C <= 1) return;
  o[1] = -0.48860251190291987f * y;
  o[2] = 0.48860251190291987f * z;
  o[3] = -0.48860251190291987f * x;
  if (C <= 2) return;
  o[4] = 1.0925484305920792f * xy;
  o[5] = -1.0925484305920792f * yz;
  o[6] = 0.94617469575755997f * z2 - 0.31539156525251999f;
  o[7] = -1.0925484305920792f * xz;
  o[8] = 0.54627421529603959f * x2 - 0.54627421529603959f * y2;
  if (C <= 3) return;
  o[9] = 0.59004358992664352f * y * (-3.0f * x2 + y2);
  o[10] = 2.8906114426405538f * xy * z;
  o[11] = 0.45704579946446572f * y * (1.0f - 5.0f * z2);
  o[12] = 0.3731763325901154f * z * (5.0f * z2 - 3.0f);
  o[13] = 0.45704579946446572f * x * (1.0f - 5.0f * z2);
  o[14] = 1.4453057213202769f * z * (x2 - y2);
  o[15] = 0.59004358992664352f * x * (-x2 + 3.0f * y2);
}

/*
 * vol_render_sh.h:48-65 calc_direction.  The reference reinterprets the c2w
 * pointer as three consecutive float3, so with a row-major 3x4 c2w the "rows"
 * are c2w[0..2], c2w[3..5], c2w[6..8] (translation leaks in).  pos = (px,py,1).
 */
void orc_direction(const float *c2w, float px, float py, float *d) {
  float d0 = c2w[0] * px + c2w[1] * py + c2w[2] * 1.0f;
  float d1 = c2w[3] * px + c2w[4] * py + c2w[5] * 1.0f;
  float d2 = c2w[6] * px + c2w[7] * py + c2w[8] * 1.0f;
  float len = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
  d[0] = d0 / len;
  d[1] = d1 / len;
  d[2] = d2 / len;
}

/* kernels.h:172-193 kernel_gaussian_2d_float (float arithmetic). */
float orc_gauss2d_f(const float *m, const float *c, float px, float py) {
  float c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
  float det = c0 * c3 - c1 * c2;
  float x = px - m[0];
  float y = py - m[1];
  float tmpx = x * c3 - y * c2;
  float tmpy = -x * c1 + y * c0;
  float radial = tmpx * x + tmpy * y;
  radial /= det;
  if (radial < 0.0f) radial = 1000.0f;
  return orc_expf(-0.5f * radial);
}

/* kernels.h:195-224 kernel_gaussian_2d (double arithmetic, float result). */
float orc_gauss2d_d(const float *m, const float *c, float px, float py) {
  double c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
  double det = c0 * c3 - c1 * c2;
  double x = (double)(px - m[0]);
  double y = (double)(py - m[1]);
  double tmpx = x * c3 - y * c2;
  double tmpy = -x * c1 + y * c0;
  double radial = tmpx * x + tmpy * y;
  radial /= det;
  if (radial < 0.0) radial = 1000.0;
  return (float)orc_exp_d(-0.5 * radial);
}

/* vol_render_sh.h:18-26 sum_C. */
static float orc_dot(const float *a, const float *b, int n) {
  float s = 0.0f;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/*
 * Shared per-pixel setup: pixel position (vol_render_sh.h:213-214, no
 * half-pixel offset) and the SH basis of the pixel ray.
 */
static void orc_pixel(const float *topleft, float psx, float psy, int gx,
                      int gy, float *pos) {
  pos[0] = topleft[0] + (float)gx * psx;
  pos[1] = topleft[1] + (float)gy * psy;
}

/*
 * Forward SH compositing, one call for the whole image.
 * Reference: vol_render_sh.h:171-248 (entry) and :97-169 (batch loop); bg
 * variant vol_render_bg.h:12-100.  mode: 0 = render_sh, 1 = render_sh_bg.
 * out is written only for pixels of non-empty tiles (mode 0), as the
 * reference leaves the caller's zeros elsewhere; mode 1 writes bg into pixels
 * of empty tiles (the evident intent of vol_render_bg.h:28-37, which does not
 * compile as shipped).
 * Optional outputs (may be NULL): T_final[H*W] (transmittance after the loop),
 * n_done[H*W] (number of list entries examined before the T<thresh break),
 * pre_bg[H*W*3] (colour before the background blend).
 * Tiles are independent (each writes only its own pixels), so they run in
 * parallel (OpenMP); the result does not depend on the thread count.
 * Returns 0, or -1 for an unsupported C.
 */
int orc_render_sh_fwd(int C, const float *mean, const float *cov,
                      const float *sh, const float *alpha, const int *start,
                      const int *end, const int *ids, float *out,
                      const float *topleft, const float *c2w, int tile_size,
                      int n_tiles_h, int n_tiles_w, float psx, float psy, int H,
                      int W, float thresh, int mode, const float *bg,
                      float *T_final, int *n_done, float *pre_bg) {
  if (C < 1 || C > ORC_MAX_C) return -1;
  const int CC = C * C;
  const int n_tiles = n_tiles_h * n_tiles_w;
#pragma omp parallel for schedule(dynamic, 1)
  for (int tile = 0; tile < n_tiles; ++tile) {
    const int ty = tile / n_tiles_w, tx = tile % n_tiles_w;
    const int s0 = start[tile];
    const int empty = (s0 == -1) || (end[tile] - s0 == 0);
    for (int ly = 0; ly < tile_size; ++ly) {
      for (int lx = 0; lx < tile_size; ++lx) {
        const int gx = tx * tile_size + lx, gy = ty * tile_size + ly;
        if (gx >= W || gy >= H) continue;
        const size_t pix = (size_t)gy * W + gx;
        if (empty) {
          if (mode == 1 && s0 == -1) {
            out[3 * pix + 0] = bg[0];
            out[3 * pix + 1] = bg[1];
            out[3 * pix + 2] = bg[2];
          }
          if (T_final) T_final[pix] = 1.0f;
          if (n_done) n_done[pix] = 0;
          if (pre_bg) pre_bg[3 * pix] = pre_bg[3 * pix + 1] = pre_bg[3 * pix + 2] = 0.0f;
          continue;
        }
        const int n = end[tile] - s0;
        float pos[2], dir[3], Y[16];
        orc_pixel(topleft, psx, psy, gx, gy, pos);
        orc_direction(c2w, pos[0], pos[1], dir);
        orc_sh_basis(dir, Y, C);
        float o[3] = {0.0f, 0.0f, 0.0f};
        float T = 1.0f;
        int i = 0;
        for (; i < n; ++i) {
          if (T < thresh) break;
          const int g = ids[s0 + i];
          const float a = fminf(alpha[g], 0.99f);
          float coeff = a * T;
          const float G = orc_gauss2d_f(mean + 2 * g, cov + 4 * g, pos[0], pos[1]);
          coeff *= G;
          if (a * G < kMinRenderAlpha) continue;
          if (isnan(coeff)) coeff = 0.0f;
          float y[3];
          for (int ch = 0; ch < 3; ++ch)
            y[ch] = orc_sigmoid(orc_dot(sh + (size_t)(3 * g + ch) * CC, Y, CC));
          for (int ch = 0; ch < 3; ++ch)
            if (isnan(y[ch] * coeff)) y[ch] = 0.0f;
          for (int ch = 0; ch < 3; ++ch) o[ch] += coeff * y[ch];
          T *= (1 - a * G);
        }
        if (pre_bg) {
          pre_bg[3 * pix] = o[0];
          pre_bg[3 * pix + 1] = o[1];
          pre_bg[3 * pix + 2] = o[2];
        }
        if (mode == 1 && T > thresh) {
          /* vol_render_bg.h:90-94 */
          for (int ch = 0; ch < 3; ++ch) o[ch] = o[ch] * T + bg[ch] * (1.0f - T);
        }
        out[3 * pix + 0] = o[0];
        out[3 * pix + 1] = o[1];
        out[3 * pix + 2] = o[2];
        if (T_final) T_final[pix] = T;
        if (n_done) n_done[pix] = i;
      }
    }
  }
  return 0;
}

/*
 * Backward SH compositing, the reference's forward-order recompute.
 * vol_render_sh.h:353-455 (entry), :268-351 (batch), kernels.h:394-418 (FP64
 * Gaussian gradient helper).  `out_saved` is the forward output (for mode 1
 * the blended output, vol_render_bg.h:183-188).  Gradients are ADDED into the
 * grad_* arrays (callers pass zeros, renderer.py:760-763).  grad_cov receives
 * both off-diagonals, each 0.5 * P*a*G * s*u.
 * Optional abs_* outputs (double, may be NULL): for every gradient component
 * the sum over its per-pixel terms of the term's rounding scale.  For the SH
 * terms that is |term|.  The mean, cov and alpha terms carry the factor
 *   P = sum_ch g_ch (y_ch T - (F_ch - o_ch) / (1 - a G)),
 * whose difference F - o cancels late in a list (o has reached F and T is
 * small), so their scale is that of P's parts,
 *   |P| <= sum_ch |g_ch| (|y_ch| T + (|F_ch| + |o_ch|) / (1 - a G)),
 * times the remaining factors.  Any float evaluation of the reference's
 * formula (this one included) differs from the exact value by a small
 * multiple of eps times this scale, and so may differ from the reverse-order
 * adjoint, which accumulates the colour behind instead of cancelling it.
 * Sums are formed in double, per tile over its pixels, and added to the
 * per-Gaussian totals in tile order, so the result does not depend on the
 * number of threads the tiles run on.
 * Returns 0, -1 for unsupported C, 1 if the recomputed colour differs from
 * out_saved (the reference asserts equality, vol_render_sh.h:452-454).
 */
int orc_render_sh_bwd(int C, const float *mean, const float *cov,
                      const float *sh, const float *alpha, const int *start,
                      const int *end, const int *ids, const float *out_saved,
                      float *grad_mean, float *grad_cov, float *grad_sh,
                      float *grad_alpha, const float *grad_out,
                      const float *topleft, const float *c2w, int tile_size,
                      int n_tiles_h, int n_tiles_w, float psx, float psy, int H,
                      int W, float thresh, int mode, const float *bg, int V,
                      double *abs_mean, double *abs_cov, double *abs_sh,
                      double *abs_alpha) {
  if (C < 1 || C > ORC_MAX_C) return -1;
  const int CC = C * C;
  /* per entry: mean 2, cov 4, sh 3*CC, alpha 1 */
  const int NC = 7 + 3 * CC;
  const int n_tiles = n_tiles_h * n_tiles_w;
  int max_n = 0;
  for (int t = 0; t < n_tiles; ++t)
    if (start[t] != -1 && end[t] - start[t] > max_n) max_n = end[t] - start[t];
  double *acc = calloc((size_t)V * NC, sizeof(double));
  double *acc_abs = calloc((size_t)V * NC, sizeof(double));
  int mismatch = 0;
#pragma omp parallel reduction(| : mismatch)
  {
    /* this thread's per-tile sums, indexed by list position */
    double *loc = calloc((size_t)(max_n > 0 ? max_n : 1) * NC, sizeof(double));
    double *loc_abs = calloc((size_t)(max_n > 0 ? max_n : 1) * NC, sizeof(double));
#pragma omp for ordered schedule(dynamic, 1)
    for (int tile = 0; tile < n_tiles; ++tile) {
      const int ty = tile / n_tiles_w, tx = tile % n_tiles_w;
      const int s0 = start[tile];
      int used = 0; /* entries of the tile list with a term */
      if (s0 != -1 && end[tile] - s0 != 0) {
        const int n = end[tile] - s0;
        for (int ly = 0; ly < tile_size; ++ly) {
          for (int lx = 0; lx < tile_size; ++lx) {
            const int gx = tx * tile_size + lx, gy = ty * tile_size + ly;
            if (gx >= W || gy >= H) continue;
            const size_t pix = (size_t)gy * W + gx;
            float pos[2], dir[3], Y[16];
            orc_pixel(topleft, psx, psy, gx, gy, pos);
            orc_direction(c2w, pos[0], pos[1], dir);
            orc_sh_basis(dir, Y, C);
            const float *g = grad_out + 3 * pix;
            const float *F = out_saved + 3 * pix;
            float o[3] = {0.0f, 0.0f, 0.0f};
            float T = 1.0f;
            for (int i = 0; i < n; ++i) {
              if (T < thresh) break;
              const int gid = ids[s0 + i];
              const float a = fminf(alpha[gid], 0.99f);
              const float G = orc_gauss2d_f(mean + 2 * gid, cov + 4 * gid, pos[0], pos[1]);
              if (a * G < kMinRenderAlpha) continue;
              float coeff = a * T * G;
              if (isnan(coeff)) coeff = 0.0f;
              float y[3];
              for (int ch = 0; ch < 3; ++ch)
                y[ch] = orc_sigmoid(orc_dot(sh + (size_t)(3 * gid + ch) * CC, Y, CC));
              for (int ch = 0; ch < 3; ++ch)
                if (isnan(y[ch] * coeff)) y[ch] = 0.0f;
              for (int ch = 0; ch < 3; ++ch) o[ch] += coeff * y[ch];
              double *L = loc + (size_t)i * NC, *LA = loc_abs + (size_t)i * NC;
              if (i + 1 > used) used = i + 1;
              /* backward_C: grad_sh += (coeff * y(1-y) * g) * Y_k */
              for (int ch = 0; ch < 3; ++ch) {
                const float gr = coeff * (y[ch] * (1.0f - y[ch])) * g[ch];
                for (int k = 0; k < CC; ++k) {
                  const double v = (double)(gr * Y[k]);
                  L[6 + ch * CC + k] += v;
                  LA[6 + ch * CC + k] += fabs(v);
                }
              }
              float paG = 0.0f;
              for (int ch = 0; ch < 3; ++ch)
                paG += g[ch] * (y[ch] * T - (F[ch] - o[ch]) / (1 - a * G));
              double pab = 0.0; /* rounding scale of paG */
              for (int ch = 0; ch < 3; ++ch)
                pab += fabs((double)g[ch]) * (fabs((double)y[ch]) * (double)T +
                                              (fabs((double)F[ch]) + fabs((double)o[ch])) / (double)(1 - a * G));
              /* kernels.h:394-418 in double */
              {
                const float *m = mean + 2 * gid, *c = cov + 4 * gid;
                const double dg = (double)(paG * a * G);
                const double c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
                const double det = c0 * c3 - c1 * c2;
                const double x = (double)(pos[0] - m[0]);
                const double yy = (double)(pos[1] - m[1]);
                const double tmpx = (x * c3 - yy * c2) / det;
                const double tmpy = (-x * c1 + yy * c0) / det;
                double v[6];
                v[0] = (double)(float)(dg * tmpx);
                v[1] = (double)(float)(dg * tmpy);
                v[2] = 0.5 * (double)(float)(dg * tmpx * tmpx);
                v[3] = 0.5 * (double)(float)(dg * tmpx * tmpy);
                v[4] = 0.5 * (double)(float)(dg * tmpy * tmpx);
                v[5] = 0.5 * (double)(float)(dg * tmpy * tmpy);
                const double dab = pab * (double)a * (double)G;
                const double va[6] = {dab * fabs(tmpx), dab * fabs(tmpy), 0.5 * dab * tmpx * tmpx,
                                      0.5 * dab * fabs(tmpx * tmpy), 0.5 * dab * fabs(tmpx * tmpy),
                                      0.5 * dab * tmpy * tmpy};
                for (int q = 0; q < 6; ++q) {
                  L[q] += v[q];
                  LA[q] += va[q];
                }
              }
              {
                const double v = (double)(paG * G);
                L[6 + 3 * CC] += v;
                LA[6 + 3 * CC] += pab * (double)G;
              }
              T *= (1 - a * G);
            }
            if (mode == 1 && T > thresh)
              for (int ch = 0; ch < 3; ++ch) o[ch] = o[ch] * T + bg[ch] * (1.0f - T);
            for (int ch = 0; ch < 3; ++ch)
              if (o[ch] != F[ch]) mismatch = 1;
          }
        }
      }
#pragma omp ordered
      {
        for (int i = 0; i < used; ++i) {
          const int gid = ids[s0 + i];
          double *A = acc + (size_t)gid * NC, *AA = acc_abs + (size_t)gid * NC;
          double *L = loc + (size_t)i * NC, *LA = loc_abs + (size_t)i * NC;
          for (int q = 0; q < NC; ++q) {
            A[q] += L[q];
            AA[q] += LA[q];
            L[q] = 0.0;
            LA[q] = 0.0;
          }
        }
      }
    }
    free(loc);
    free(loc_abs);
  }
  for (size_t v = 0; v < (size_t)V; ++v) {
    const double *A = acc + v * NC, *AA = acc_abs + v * NC;
    grad_mean[2 * v + 0] += (float)A[0];
    grad_mean[2 * v + 1] += (float)A[1];
    for (int q = 0; q < 4; ++q) grad_cov[4 * v + q] += (float)A[2 + q];
    for (int q = 0; q < 3 * CC; ++q) grad_sh[v * 3 * CC + q] += (float)A[6 + q];
    grad_alpha[v] += (float)A[6 + 3 * CC];
    if (abs_mean) {
      abs_mean[2 * v + 0] += AA[0];
      abs_mean[2 * v + 1] += AA[1];
    }
    if (abs_cov)
      for (int q = 0; q < 4; ++q) abs_cov[4 * v + q] += AA[2 + q];
    if (abs_sh)
      for (int q = 0; q < 3 * CC; ++q) abs_sh[v * 3 * CC + q] += AA[6 + q];
    if (abs_alpha) abs_alpha[v] += AA[6 + 3 * CC];
  }
  free(acc);
  free(acc_abs);
  return mismatch;
}

/*
 * The same adjoint evaluated in double (test infrastructure: the exact value
 * that the float evaluations above and on the GPU approximate).  The skip and
 * break decisions are the float forward's (orc_gauss2d_f, float T), as in
 * orc_render_sh_bwd; every value along the composited entries (Gaussian
 * value, sigmoid colour, transmittance, F, the colour behind) is formed in
 * double, with F - o taken in double after a first pass.  Per-tile sums are
 * added in tile order.  Gradients are ADDED into the double grad_* arrays.
 * mode 1 blends bg as vol_render_bg.h:183-188 (where the float T ends above
 * thresh).  Returns 0, or -1 for unsupported C.
 */
int orc_render_sh_bwd_f64(int C, const float *mean, const float *cov, const float *sh, const float *alpha,
                          const int *start, const int *end, const int *ids, double *grad_mean,
                          double *grad_cov, double *grad_sh, double *grad_alpha, const float *grad_out,
                          const float *topleft, const float *c2w, int tile_size, int n_tiles_h, int n_tiles_w,
                          float psx, float psy, int H, int W, float thresh, int mode, const float *bg, int V) {
  if (C < 1 || C > ORC_MAX_C) return -1;
  const int CC = C * C;
  const int NC = 7 + 3 * CC;
  const int n_tiles = n_tiles_h * n_tiles_w;
  int max_n = 0;
  for (int t = 0; t < n_tiles; ++t)
    if (start[t] != -1 && end[t] - start[t] > max_n) max_n = end[t] - start[t];
  const size_t cap = (size_t)(max_n > 0 ? max_n : 1);
  double *acc = calloc((size_t)V * NC, sizeof(double));
#pragma omp parallel
  {
    double *loc = calloc(cap * NC, sizeof(double));
    int *sel = malloc(cap * sizeof(int));         /* composited list positions */
    double *val = malloc(cap * 6 * sizeof(double)); /* G, T before, y0..y2, a */
#pragma omp for ordered schedule(dynamic, 1)
    for (int tile = 0; tile < n_tiles; ++tile) {
      const int ty = tile / n_tiles_w, tx = tile % n_tiles_w;
      const int s0 = start[tile];
      int used = 0;
      if (s0 != -1 && end[tile] - s0 != 0) {
        const int n = end[tile] - s0;
        for (int ly = 0; ly < tile_size; ++ly) {
          for (int lx = 0; lx < tile_size; ++lx) {
            const int gx = tx * tile_size + lx, gy = ty * tile_size + ly;
            if (gx >= W || gy >= H) continue;
            const size_t pix = (size_t)gy * W + gx;
            float pos[2], dir[3], Y[16];
            orc_pixel(topleft, psx, psy, gx, gy, pos);
            orc_direction(c2w, pos[0], pos[1], dir);
            orc_sh_basis(dir, Y, C);
            const float *g = grad_out + 3 * pix;
            float Tf = 1.0f;
            double Td = 1.0, o[3] = {0.0, 0.0, 0.0};
            int m = 0;
            for (int i = 0; i < n; ++i) {
              if (Tf < thresh) break;
              const int gid = ids[s0 + i];
              const float af = fminf(alpha[gid], 0.99f);
              const float Gf = orc_gauss2d_f(mean + 2 * gid, cov + 4 * gid, pos[0], pos[1]);
              if (af * Gf < kMinRenderAlpha) continue;
              const float *mm = mean + 2 * gid, *c = cov + 4 * gid;
              const double c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3], det = c0 * c3 - c1 * c2;
              const double x = (double)pos[0] - mm[0], yy = (double)pos[1] - mm[1];
              double r = ((x * c3 - yy * c2) * x + (-x * c1 + yy * c0) * yy) / det;
              if (r < 0.0) r = 1000.0;
              const double G = exp(-0.5 * r), a = (double)af;
              double *vv = val + (size_t)m * 6;
              vv[0] = G;
              vv[1] = Td;
              vv[5] = a;
              for (int ch = 0; ch < 3; ++ch) {
                double d = 0.0;
                for (int k = 0; k < CC; ++k) d += (double)sh[(size_t)(3 * gid + ch) * CC + k] * (double)Y[k];
                vv[2 + ch] = 1.0 / (1.0 + exp(-d));
                o[ch] += a * Td * G * vv[2 + ch];
              }
              sel[m++] = i;
              Tf *= (1 - af * Gf);
              Td *= 1.0 - a * G;
            }
            double F[3] = {o[0], o[1], o[2]};
            if (mode == 1 && Tf > thresh)
              for (int ch = 0; ch < 3; ++ch) F[ch] = o[ch] * Td + (double)bg[ch] * (1.0 - Td);
            double oa[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < m; ++q) {
              const int i = sel[q];
              const int gid = ids[s0 + i];
              const double *vv = val + (size_t)q * 6;
              const double G = vv[0], T = vv[1], a = vv[5], w = a * T * G;
              double *L = loc + (size_t)i * NC;
              if (i + 1 > used) used = i + 1;
              double P = 0.0;
              for (int ch = 0; ch < 3; ++ch) {
                const double y = vv[2 + ch];
                oa[ch] += w * y;
                const double gr = w * y * (1.0 - y) * (double)g[ch];
                for (int k = 0; k < CC; ++k) L[6 + ch * CC + k] += gr * (double)Y[k];
                P += (double)g[ch] * (y * T - (F[ch] - oa[ch]) / (1.0 - a * G));
              }
              const float *mm = mean + 2 * gid, *c = cov + 4 * gid;
              const double c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3], det = c0 * c3 - c1 * c2;
              const double x = (double)pos[0] - mm[0], yy = (double)pos[1] - mm[1];
              const double su = (x * c3 - yy * c2) / det, uu = (-x * c1 + yy * c0) / det;
              const double D = P * a * G;
              L[0] += D * su;
              L[1] += D * uu;
              L[2] += 0.5 * D * su * su;
              L[3] += 0.5 * D * su * uu;
              L[4] += 0.5 * D * uu * su;
              L[5] += 0.5 * D * uu * uu;
              L[6 + 3 * CC] += P * G;
            }
          }
        }
      }
#pragma omp ordered
      {
        for (int i = 0; i < used; ++i) {
          double *A = acc + (size_t)ids[s0 + i] * NC, *L = loc + (size_t)i * NC;
          for (int q = 0; q < NC; ++q) {
            A[q] += L[q];
            L[q] = 0.0;
          }
        }
      }
    }
    free(loc);
    free(sel);
    free(val);
  }
  for (size_t v = 0; v < (size_t)V; ++v) {
    const double *A = acc + v * NC;
    grad_mean[2 * v + 0] += A[0];
    grad_mean[2 * v + 1] += A[1];
    for (int q = 0; q < 4; ++q) grad_cov[4 * v + q] += A[2 + q];
    for (int q = 0; q < 3 * CC; ++q) grad_sh[v * 3 * CC + q] += A[6 + q];
    grad_alpha[v] += A[6 + 3 * CC];
  }
  free(acc);
  return 0;
}

/* Threads the parallel oracle loops use (OpenMP); n <= 0 keeps the default. */
void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/*
 * RGB compositing (render_start_end): vol_render.h:716-781 with the batch
 * loop vol_render.h:169-250; the Gaussian is evaluated in double
 * (kernels.h:195-224).  Colour is used as given (no SH, no sigmoid).
 */
int orc_render_rgb_fwd(const float *mean, const float *cov, const float *color,
                       const float *alpha, const int *start, const int *end,
                       const int *ids, float *out, const float *topleft,
                       int tile_size, int n_tiles_h, int n_tiles_w, float psx,
                       float psy, int H, int W, float thresh, float *T_final,
                       int *n_done) {
  for (int ty = 0; ty < n_tiles_h; ++ty) {
    for (int tx = 0; tx < n_tiles_w; ++tx) {
      const int tile = ty * n_tiles_w + tx;
      const int s0 = start[tile];
      const int empty = (s0 == -1) || (end[tile] - s0 == 0);
      for (int ly = 0; ly < tile_size; ++ly) {
        for (int lx = 0; lx < tile_size; ++lx) {
          const int gx = tx * tile_size + lx, gy = ty * tile_size + ly;
          if (gx >= W || gy >= H) continue;
          const size_t pix = (size_t)gy * W + gx;
          if (empty) {
            if (T_final) T_final[pix] = 1.0f;
            if (n_done) n_done[pix] = 0;
            continue;
          }
          const int n = end[tile] - s0;
          float pos[2];
          orc_pixel(topleft, psx, psy, gx, gy, pos);
          float o[3] = {0.0f, 0.0f, 0.0f};
          float T = 1.0f;
          int i = 0;
          for (; i < n; ++i) {
            if (T < thresh) break;
            const int g = ids[s0 + i];
            const float a = fminf(alpha[g], 0.99f);
            float coeff = a * T;
            const float G = orc_gauss2d_d(mean + 2 * g, cov + 4 * g, pos[0], pos[1]);
            coeff *= G;
            if (a * G < kMinRenderAlpha) continue;
            for (int ch = 0; ch < 3; ++ch) o[ch] += color[3 * g + ch] * coeff;
            T *= (1 - a * G);
          }
          out[3 * pix + 0] = o[0];
          out[3 * pix + 1] = o[1];
          out[3 * pix + 2] = o[2];
          if (T_final) T_final[pix] = T;
          if (n_done) n_done[pix] = i;
        }
      }
    }
  }
  return 0;
}

/*
 * RGB backward (render_start_end): vol_render.h:800-905 with the batch loop
 * vol_render.h:252-352.  partial_aG is accumulated in double there.
 */
int orc_render_rgb_bwd(const float *mean, const float *cov, const float *color,
                       const float *alpha, const int *start, const int *end,
                       const int *ids, const float *out_saved, float *grad_mean,
                       float *grad_cov, float *grad_color, float *grad_alpha,
                       const float *grad_out, const float *topleft,
                       int tile_size, int n_tiles_h, int n_tiles_w, float psx,
                       float psy, int H, int W, float thresh, int V) {
  double *am = calloc((size_t)V * 2, sizeof(double));
  double *ac = calloc((size_t)V * 4, sizeof(double));
  double *acol = calloc((size_t)V * 3, sizeof(double));
  double *aa = calloc((size_t)V, sizeof(double));
  int mismatch = 0;
  for (int ty = 0; ty < n_tiles_h; ++ty) {
    for (int tx = 0; tx < n_tiles_w; ++tx) {
      const int tile = ty * n_tiles_w + tx;
      const int s0 = start[tile];
      if (s0 == -1 || end[tile] - s0 == 0) continue;
      const int n = end[tile] - s0;
      for (int ly = 0; ly < tile_size; ++ly) {
        for (int lx = 0; lx < tile_size; ++lx) {
          const int gx = tx * tile_size + lx, gy = ty * tile_size + ly;
          if (gx >= W || gy >= H) continue;
          const size_t pix = (size_t)gy * W + gx;
          float pos[2];
          orc_pixel(topleft, psx, psy, gx, gy, pos);
          const float *g = grad_out + 3 * pix;
          const float *F = out_saved + 3 * pix;
          float o[3] = {0.0f, 0.0f, 0.0f};
          float T = 1.0f;
          for (int i = 0; i < n; ++i) {
            if (T < thresh) break;
            const int gid = ids[s0 + i];
            const float a = fminf(alpha[gid], 0.99f);
            const float G = orc_gauss2d_d(mean + 2 * gid, cov + 4 * gid, pos[0], pos[1]);
            if (a * G < kMinRenderAlpha) continue;
            const float coeff = a * T * G;
            for (int ch = 0; ch < 3; ++ch) o[ch] += color[3 * gid + ch] * coeff;
            for (int ch = 0; ch < 3; ++ch) acol[3 * gid + ch] += (double)(coeff * g[ch]);
            double paG = 0.0;
            for (int ch = 0; ch < 3; ++ch)
              paG += (double)((color[3 * gid + ch] * T - (F[ch] - o[ch]) / (1 - a * G)) * g[ch]);
            {
              const float *m = mean + 2 * gid, *c = cov + 4 * gid;
              const double dg = (double)(float)(paG * a * G);
              const double c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
              const double det = c0 * c3 - c1 * c2;
              const double x = (double)(pos[0] - m[0]);
              const double yy = (double)(pos[1] - m[1]);
              const double tmpx = (x * c3 - yy * c2) / det;
              const double tmpy = (-x * c1 + yy * c0) / det;
              am[2 * gid + 0] += (double)(float)(dg * tmpx);
              am[2 * gid + 1] += (double)(float)(dg * tmpy);
              ac[4 * gid + 0] += 0.5 * (double)(float)(dg * tmpx * tmpx);
              ac[4 * gid + 1] += 0.5 * (double)(float)(dg * tmpx * tmpy);
              ac[4 * gid + 2] += 0.5 * (double)(float)(dg * tmpy * tmpx);
              ac[4 * gid + 3] += 0.5 * (double)(float)(dg * tmpy * tmpy);
            }
            aa[gid] += (double)(float)(paG * G);
            T *= (1 - a * G);
          }
          for (int ch = 0; ch < 3; ++ch)
            if (o[ch] != F[ch]) mismatch = 1;
        }
      }
    }
  }
  for (size_t i = 0; i < (size_t)V * 2; ++i) grad_mean[i] += (float)am[i];
  for (size_t i = 0; i < (size_t)V * 4; ++i) grad_cov[i] += (float)ac[i];
  for (size_t i = 0; i < (size_t)V * 3; ++i) grad_color[i] += (float)acol[i];
  for (size_t i = 0; i < (size_t)V; ++i) grad_alpha[i] += (float)aa[i];
  free(am);
  free(ac);
  free(acol);
  free(aa);
  return mismatch;
}

/*
 * Frustum cull: culling.h:11-20 with kernels.h:156-170.  r = max(s)*thresh;
 * keep iff for all six planes dot(mu - pts_k, n_k) > -r.  (qvec is unused.)
 */
void orc_cull(int N, const float *mean, const float *svec, const float *normals,
              const float *pts, uint8_t *mask, float thresh) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    const float *m = mean + 3 * i, *s = svec + 3 * i;
    const float r = fmaxf(fmaxf(s[0], s[1]), s[2]) * thresh;
    int keep = 1;
    for (int k = 0; k < 6; ++k) {
      const float *p = pts + 3 * k, *n = normals + 3 * k;
      const float dx = m[0] - p[0], dy = m[1] - p[1], dz = m[2] - p[2];
      const float d = dx * n[0] + dy * n[1] + dz * n[2];
      if (!(d > -r)) {
        keep = 0;
        break;
      }
    }
    mask[i] = (uint8_t)keep;
  }
}

/*
 * Activations, sh_renderer.py:318-324 with utils/activations.py:13-19:
 * s = exp(svec_before_activation), alpha = sigmoid(alpha_before_activation).
 */
void orc_activate(int N, const float *log_s, const float *logit, float *s,
                  float *a) {
  for (int i = 0; i < 3 * N; ++i) s[i] = orc_expf(log_s[i]);
  for (int i = 0; i < N; ++i) a[i] = 1.0f / (1.0f + orc_expf(-logit[i]));
}

/*
 * Unit quaternion (w,x,y,z) -> rotation matrix, row-major, after
 * normalisation q / max(|q|, 1e-12).  utils/transforms.py:31-45 calls
 * kornia.geometry.conversions.quaternion_to_rotation_matrix(q, WXYZ), which
 * normalises and then applies the textbook formula below.
 */
void orc_quat_to_rot(const float *q, float *R) {
  const float nrm = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const float den = fmaxf(nrm, 1e-12f);
  const float w = q[0] / den, x = q[1] / den, y = q[2] / den, z = q[3] / den;
  const float tx = 2.0f * x, ty = 2.0f * y, tz = 2.0f * z;
  const float twx = tx * w, twy = ty * w, twz = tz * w;
  const float txx = tx * x, txy = ty * x, txz = tz * x;
  const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
  R[0] = 1.0f - (tyy + tzz);
  R[1] = txy - twz;
  R[2] = txz + twy;
  R[3] = txy + twz;
  R[4] = 1.0f - (txx + tzz);
  R[5] = tyz - twx;
  R[6] = txz - twy;
  R[7] = tyz + twx;
  R[8] = 1.0f - (txx + tyy);
}

/*
 * Projection, renderer.py:380-419 (project_pts, jacobian, project_gaussians)
 * with c2w = [R|t] row-major 3x4:
 *   u = R^T (mu - t);  M_ij = R(q)_ij * s_j;  Sigma = M M^T;
 *   J = [[1/z, 0, -x/z/z], [0, 1/z, -y/z/z], ...];  A = J R^T;
 *   cov2d = (A Sigma A^T)[:2,:2] (all four entries);  depth = z;
 *   mean2d = (x, y) / depth.
 * Sums run over the contracted index in increasing order.
 */
void orc_project(int V, const float *mean, const float *qvec, const float *svec,
                 const float *c2w, float *mean2d, float *cov2d, float *depth) {
  float Wm[9]; /* W = R^T */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Wm[3 * i + j] = c2w[4 * j + i];
  const float t0 = c2w[3], t1 = c2w[7], t2 = c2w[11];
#pragma omp parallel for schedule(static)
  for (int v = 0; v < V; ++v) {
    const float *m = mean + 3 * v;
    const float d0 = m[0] + (-t0), d1 = m[1] + (-t1), d2 = m[2] + (-t2);
    float u[3];
    for (int i = 0; i < 3; ++i)
      u[i] = Wm[3 * i + 0] * d0 + Wm[3 * i + 1] * d1 + Wm[3 * i + 2] * d2;
    float R[9], M[9], S[9];
    orc_quat_to_rot(qvec + 4 * v, R);
    const float *s = svec + 3 * v;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M[3 * i + j] = s[j] * R[3 * i + j];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        S[3 * a + b] = M[3 * a + 0] * M[3 * b + 0] + M[3 * a + 1] * M[3 * b + 1] +
                       M[3 * a + 2] * M[3 * b + 2];
    const float x = u[0], y = u[1], z = u[2];
    const float J00 = 1.0f / z, J02 = -x / z / z;
    const float J11 = 1.0f / z, J12 = -y / z / z;
    float A[6];
    for (int k = 0; k < 3; ++k) {
      A[k] = J00 * Wm[0 * 3 + k] + J02 * Wm[2 * 3 + k];
      A[3 + k] = J11 * Wm[1 * 3 + k] + J12 * Wm[2 * 3 + k];
    }
    float T1[6];
    for (int i = 0; i < 2; ++i)
      for (int b = 0; b < 3; ++b)
        T1[3 * i + b] = A[3 * i + 0] * S[0 * 3 + b] + A[3 * i + 1] * S[1 * 3 + b] +
                        A[3 * i + 2] * S[2 * 3 + b];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        cov2d[4 * v + 2 * i + j] = T1[3 * i + 0] * A[3 * j + 0] +
                                   T1[3 * i + 1] * A[3 * j + 1] +
                                   T1[3 * i + 2] * A[3 * j + 2];
    depth[v] = z;
    mean2d[2 * v + 0] = x / z;
    mean2d[2 * v + 1] = y / z;
  }
}

/*
 * Tile AABB: gs/culling.py:8-37 with utils/camera.py:290-303, all in float32:
 *   e = sqrt(D * cov_ii); tl/br = mean2d -/+ e;
 *   pix = int32(v * f + c) (multiply, then add, as two rounded ops);
 *   clamp to [0, W-1] x [0, H-1]; floor-divide by the tile size.
 * Returns the pair count P = sum (brx-tlx+1)(bry-tly+1).
 */
int64_t orc_aabb(int V, const float *mean2d, const float *cov2d, float fx,
                 float fy, float cx, float cy, int W, int H, float D,
                 int tile_size, int32_t *tl, int32_t *br) {
  int64_t P = 0;
#pragma omp parallel for schedule(static) reduction(+ : P)
  for (int v = 0; v < V; ++v) {
    const float ex = sqrtf(D * cov2d[4 * v + 0]);
    const float ey = sqrtf(D * cov2d[4 * v + 3]);
    const float mx = mean2d[2 * v], my = mean2d[2 * v + 1];
    const float lo[2] = {mx - ex, my - ey}, hi[2] = {mx + ex, my + ey};
    const float f[2] = {fx, fy}, c[2] = {cx, cy};
    const int lim[2] = {W - 1, H - 1};
    for (int k = 0; k < 2; ++k) {
      const float pl = lo[k] * f[k];
      const float ph = hi[k] * f[k];
      int32_t il = orc_f2i_sat(pl + c[k]);
      int32_t ih = orc_f2i_sat(ph + c[k]);
      il = il < 0 ? 0 : (il > lim[k] ? lim[k] : il);
      ih = ih < 0 ? 0 : (ih > lim[k] ? lim[k] : ih);
      tl[2 * v + k] = il / tile_size; /* non-negative: floor == trunc */
      br[2 * v + k] = ih / tile_size;
    }
    P += (int64_t)(br[2 * v] - tl[2 * v] + 1) * (int64_t)(br[2 * v + 1] - tl[2 * v + 1] + 1);
  }
  return P;
}

/* Vectorised helpers so tests can sweep the scalar functions. */
void orc_expf_vec(int64_t n, const float *x, float *y) {
  for (int64_t i = 0; i < n; ++i) y[i] = orc_expf(x[i]);
}
void orc_exp_d_vec(int64_t n, const double *x, double *y) {
  for (int64_t i = 0; i < n; ++i) y[i] = orc_exp_d(x[i]);
}
/* Largest |orc_expf(x) - expf(x)| in units of the float spacing at expf(x)
 * over every binary32 x in [lo, hi], scanning bit patterns; n_diff counts
 * inputs whose results differ at all. */
double orc_expf_sweep(float lo, float hi, int64_t *n_checked, int64_t *n_diff) {
  double worst = 0.0;
  int64_t cnt = 0, nd = 0;
  for (float x = lo; x <= hi; x = nextafterf(x, INFINITY)) {
    const float a = orc_expf(x), b = expf(x);
    ++cnt;
    if (a != b) {
      ++nd;
      const double ulp = (double)nextafterf(b, INFINITY) - (double)b;
      const double e = fabs((double)a - (double)b) / (ulp > 0 ? ulp : 1e-45);
      if (e > worst) worst = e;
    }
    if (x == hi) break;
  }
  *n_checked = cnt;
  *n_diff = nd;
  return worst;
}
void orc_sh_basis_vec(int64_t n, const float *dirs, float *out, int C) {
  for (int64_t i = 0; i < n; ++i) orc_sh_basis(dirs + 3 * i, out + (size_t)i * C * C, C);
}
void orc_gauss2d_f_vec(int64_t n, const float *m, const float *c,
                       const float *p, float *out) {
  for (int64_t i = 0; i < n; ++i)
    out[i] = orc_gauss2d_f(m + 2 * i, c + 4 * i, p[2 * i], p[2 * i + 1]);
}
