// Photometric loss with the SSIM term on gfx950: utils/loss.py:14-21,
//   loss = m * ssim_loss(out, gt, win) + (1 - m) * base(out, gt),
// base = mean |out - gt| (l1) or mean (out - gt)^2 (l2), and the gradient of
// the loss with respect to out, for images [H, W, CH] (HWC, the layout the
// rasteriser writes; the reference moves the channels first, which changes
// nothing: SSIM is per channel).
//
// ssim_loss is kornia's (third party, unpinned in requirements.txt):
//   ssim_loss = mean over (CH, H, W) of (1 - SSIM) / 2,
//   SSIM = (2 mu1 mu2 + C1)(2 s12 + C2) / ((mu1^2 + mu2^2 + C1)(s11 + s22 + C2) + 1e-12),
// C1 = 0.01^2, C2 = 0.03^2, mu = G * img, s11 = G * img1^2 - mu1^2, s12 =
// G * (img1 img2) - mu1 mu2, where G is the win x win Gaussian of sigma 1.5
// (1-D weights exp(-x^2 / 4.5) / sum over x = -win/2 .. win/2, 2-D = outer
// product) applied after reflect padding by win/2 (torch's "reflect": the
// edge pixel is not repeated).
//
// Kernel 1 (one 16 x 16 tile per block, channel by channel): the tile plus a
// win/2 halo of both images in LDS (reflected at the image border), a
// horizontal then a vertical pass of the five filtered moments
// (mu1, mu2, e11 = G*img1^2, e22, e12 = G*(img1 img2)), SSIM, and the three
// adjoint maps per pixel
//   a = g dSSIM/dmu1, b = g dSSIM/de11, c = g dSSIM/de12
// (partial derivatives with the other moments fixed, g = -m / (2 H W CH)),
// and per-block partial sums of the loss in double.
// Kernel 2: dL/dout = G^T a + 2 out (G^T b) + gt (G^T c) + base gradient,
// where G^T is the adjoint of "reflect pad, then filter": along one axis of
// length n, output p reads input q with weight
//   W(q, p) = w(|q - p|) + [q >= 1, q + p <= R] w(q + p)
//             + [q <= n - 2, 2(n - 1) - q - p <= R] w(2(n - 1) - q - p),
// and every such p lies in [q - R, q + R], so the adjoint is again a window
// sum over the tile plus halo, with W in place of w.
#include <algorithm>

#include "gsmi_device.h"

namespace gsmi {

constexpr int kSsimTile = 16;
constexpr int kSsimMaxR = 10;  // windows up to 21
constexpr int kSsimSpan = kSsimTile + 2 * kSsimMaxR;
constexpr int kSsimBlock = kSsimTile * kSsimTile;

struct SsimWeights {
  float w[kSsimMaxR + 1];  // w[d] for |d| = 0..R
};

__device__ __forceinline__ int reflect_index(int i, int n) {
  if (n == 1) return 0;
  if (i < 0) i = -i;
  if (i >= n) i = 2 * (n - 1) - i;
  return i;
}

// W(q, p) of one axis (see the header comment); |q - p| <= R.
__device__ __forceinline__ float adj_weight(const SsimWeights& sw, int R, int q, int p, int n) {
  float v = sw.w[abs(q - p)];
  if (q >= 1 && q + p <= R) v += sw.w[q + p];
  const int r = 2 * (n - 1) - q - p;
  if (q <= n - 2 && r <= R) v += sw.w[r];
  return v;
}

// Kernel 1.  maps: [3][H*W*CH] (a, b, c); partial: one double per block.
__global__ __launch_bounds__(kSsimBlock) void k_ssim_fwd(const float* __restrict__ img1,
                                                         const float* __restrict__ img2, int H, int W, int CH, int R,
                                                         SsimWeights sw, float ssim_scale, float base_scale, int l2,
                                                         float* __restrict__ maps, double* __restrict__ partial) {
  __shared__ float s_x[kSsimSpan][kSsimSpan];
  __shared__ float s_y[kSsimSpan][kSsimSpan];
  __shared__ float s_h[5][kSsimSpan][kSsimTile];  // horizontal sums: x, y, xx, yy, xy
  __shared__ double s_red[kSsimBlock / 64];
  const int tx = threadIdx.x & (kSsimTile - 1), ty = threadIdx.x / kSsimTile;
  const int x0 = blockIdx.x * kSsimTile, y0 = blockIdx.y * kSsimTile;
  const int span = kSsimTile + 2 * R;
  const int64_t npix = (int64_t)H * W;
  const int gx = x0 + tx, gy = y0 + ty;
  const bool inside = gx < W && gy < H;
  const float C1 = 0.0001f, C2 = 0.0009f;
  double acc = 0.0;
  for (int ch = 0; ch < CH; ++ch) {
    __syncthreads();
    for (int i = threadIdx.x; i < span * span; i += kSsimBlock) {
      const int ly = i / span, lx = i - ly * span;
      const int sy = reflect_index(min(max(y0 - R + ly, -R), H - 1 + R), H);
      const int sx = reflect_index(min(max(x0 - R + lx, -R), W - 1 + R), W);
      const int64_t o = ((int64_t)sy * W + sx) * CH + ch;
      s_x[ly][lx] = img1[o];
      s_y[ly][lx] = img2[o];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < span * kSsimTile; i += kSsimBlock) {
      const int ly = i / kSsimTile, lx = i - ly * kSsimTile;
      float sx = 0.f, sy = 0.f, sxx = 0.f, syy = 0.f, sxy = 0.f;
      for (int d = -R; d <= R; ++d) {
        const float wd = sw.w[abs(d)];
        const float x = s_x[ly][lx + R + d], y = s_y[ly][lx + R + d];
        sx = fmaf(wd, x, sx);
        sy = fmaf(wd, y, sy);
        sxx = fmaf(wd, x * x, sxx);
        syy = fmaf(wd, y * y, syy);
        sxy = fmaf(wd, x * y, sxy);
      }
      s_h[0][ly][lx] = sx;
      s_h[1][ly][lx] = sy;
      s_h[2][ly][lx] = sxx;
      s_h[3][ly][lx] = syy;
      s_h[4][ly][lx] = sxy;
    }
    __syncthreads();
    if (inside) {
      float m1 = 0.f, m2 = 0.f, e11 = 0.f, e22 = 0.f, e12 = 0.f;
      for (int d = -R; d <= R; ++d) {
        const float wd = sw.w[abs(d)];
        const int ly = ty + R + d;
        m1 = fmaf(wd, s_h[0][ly][tx], m1);
        m2 = fmaf(wd, s_h[1][ly][tx], m2);
        e11 = fmaf(wd, s_h[2][ly][tx], e11);
        e22 = fmaf(wd, s_h[3][ly][tx], e22);
        e12 = fmaf(wd, s_h[4][ly][tx], e12);
      }
      const float s11 = e11 - m1 * m1, s22 = e22 - m2 * m2, s12 = e12 - m1 * m2;
      const float A1 = 2.0f * m1 * m2 + C1, A2 = 2.0f * s12 + C2;
      const float B1 = m1 * m1 + m2 * m2 + C1, B2 = s11 + s22 + C2;
      const float D = B1 * B2 + 1e-12f;
      const float ssim = (A1 * A2) / D;
      // dSSIM/dmu1 (through A1, s12, B1, s11), dSSIM/de11, dSSIM/de12
      const float iD = 1.0f / D;
      const float q = A1 * A2 * iD * iD;
      const float dmu = (2.0f * m2 * A2 - 2.0f * m2 * A1) * iD - q * (2.0f * m1 * B2 - 2.0f * m1 * B1);
      const float de11 = -q * B1;
      const float de12 = 2.0f * A1 * iD;
      const int64_t o = ((int64_t)gy * W + gx) * CH + ch;
      const int64_t n = npix * CH;
      maps[o] = ssim_scale * dmu;
      maps[n + o] = ssim_scale * de11;
      maps[2 * n + o] = ssim_scale * de12;
      const float dlt = img1[o] - img2[o];
      // (1 - ssim) / 2 weighted by the SSIM factor; the base term by its own
      acc += (double)(-ssim_scale) * (double)(1.0f - ssim) + (double)base_scale * (l2 ? (double)dlt * dlt
                                                                                       : (double)fabsf(dlt));
    }
  }
  for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
}

// Kernel 2.  grad = G^T a + 2 img1 G^T b + img2 G^T c + base gradient.
__global__ __launch_bounds__(kSsimBlock) void k_ssim_bwd(const float* __restrict__ img1,
                                                         const float* __restrict__ img2, int H, int W, int CH, int R,
                                                         SsimWeights sw, float base_scale, int l2,
                                                         const float* __restrict__ maps, float* __restrict__ grad) {
  __shared__ float s_m[3][kSsimSpan][kSsimSpan];   // a, b, c over the tile + halo (0 outside the image)
  __shared__ float s_h[3][kSsimSpan][kSsimTile];   // after the x adjoint
  const int tx = threadIdx.x & (kSsimTile - 1), ty = threadIdx.x / kSsimTile;
  const int x0 = blockIdx.x * kSsimTile, y0 = blockIdx.y * kSsimTile;
  const int span = kSsimTile + 2 * R;
  const int64_t n = (int64_t)H * W * CH;
  const int gx = x0 + tx, gy = y0 + ty;
  const bool inside = gx < W && gy < H;
  for (int ch = 0; ch < CH; ++ch) {
    __syncthreads();
    for (int i = threadIdx.x; i < span * span; i += kSsimBlock) {
      const int ly = i / span, lx = i - ly * span;
      const int sy = y0 - R + ly, sx = x0 - R + lx;
      const bool ok = sy >= 0 && sy < H && sx >= 0 && sx < W;
      const int64_t o = ((int64_t)(ok ? sy : 0) * W + (ok ? sx : 0)) * CH + ch;
      s_m[0][ly][lx] = ok ? maps[o] : 0.0f;
      s_m[1][ly][lx] = ok ? maps[n + o] : 0.0f;
      s_m[2][ly][lx] = ok ? maps[2 * n + o] : 0.0f;
    }
    __syncthreads();
    // x adjoint for every halo row: column qx = x0 + lx of this tile
    for (int i = threadIdx.x; i < span * kSsimTile; i += kSsimBlock) {
      const int ly = i / kSsimTile, lx = i - ly * kSsimTile;
      const int qx = x0 + lx;
      float ha = 0.f, hb = 0.f, hc = 0.f;
      if (qx < W) {
        for (int d = -R; d <= R; ++d) {
          const int px = qx + d;
          if (px < 0 || px >= W) continue;
          const float wq = adj_weight(sw, R, qx, px, W);
          ha = fmaf(wq, s_m[0][ly][lx + R + d], ha);
          hb = fmaf(wq, s_m[1][ly][lx + R + d], hb);
          hc = fmaf(wq, s_m[2][ly][lx + R + d], hc);
        }
      }
      s_h[0][ly][lx] = ha;
      s_h[1][ly][lx] = hb;
      s_h[2][ly][lx] = hc;
    }
    __syncthreads();
    if (inside) {
      float ta = 0.f, tb = 0.f, tc = 0.f;
      for (int d = -R; d <= R; ++d) {
        const int py = gy + d;
        if (py < 0 || py >= H) continue;
        const float wq = adj_weight(sw, R, gy, py, H);
        ta = fmaf(wq, s_h[0][ty + R + d][tx], ta);
        tb = fmaf(wq, s_h[1][ty + R + d][tx], tb);
        tc = fmaf(wq, s_h[2][ty + R + d][tx], tc);
      }
      const int64_t o = ((int64_t)gy * W + gx) * CH + ch;
      const float x = img1[o], y = img2[o];
      const float dlt = x - y;
      const float gb = l2 ? 2.0f * dlt * base_scale : (dlt > 0.0f ? base_scale : (dlt < 0.0f ? -base_scale : 0.0f));
      grad[o] = ta + 2.0f * x * tb + y * tc + gb;
    }
  }
}

__global__ void k_ssim_sum(const double* __restrict__ partial, int nb, double offset, float* __restrict__ loss) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) acc += partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = (float)(s[0] + offset);
}

}  // namespace gsmi

using namespace gsmi;

extern "C" size_t gsmi_ssim_loss_workspace_size(int H, int W, int CH) {
  if (H <= 0 || W <= 0 || CH <= 0) return 0;
  const size_t nb = (size_t)((W + kSsimTile - 1) / kSsimTile) * ((H + kSsimTile - 1) / kSsimTile);
  return 3 * (size_t)H * W * CH * sizeof(float) + ((nb * sizeof(double) + 255) & ~size_t(255)) + 256;
}

extern "C" int gsmi_photometric_ssim_loss(int H, int W, int CH, const float* out, const float* gt, int win,
                                          float ssim_mult, int l2, float* grad, float* loss, void* workspace,
                                          size_t workspace_bytes, void* stream) {
  if (H <= 0 || W <= 0 || CH <= 0 || CH > 4 || !out || !gt || !loss || !grad) return GSMI_ERR_ARG;
  if (win < 1 || win % 2 == 0 || win / 2 > kSsimMaxR || win / 2 >= std::max(H, W)) return GSMI_ERR_ARG;
  if (!(ssim_mult >= 0.0f && ssim_mult <= 1.0f)) return GSMI_ERR_ARG;
  if (!workspace || workspace_bytes < gsmi_ssim_loss_workspace_size(H, W, CH)) return GSMI_ERR_WORKSPACE;
  const int R = win / 2;
  if (R >= H || R >= W) return GSMI_ERR_ARG;  // one reflection must suffice
  // 1-D weights as kornia builds them: float32 exp(-x^2 / (2 sigma^2)), normalised by their float32 sum
  SsimWeights sw{};
  float raw[2 * kSsimMaxR + 1];
  float sum = 0.0f;
  for (int i = 0; i < win; ++i) {
    const float x = (float)(i - R);
    raw[i] = std::exp(-(x * x) / 4.5f);
  }
  for (int i = 0; i < win; ++i) sum += raw[i];
  for (int d = 0; d <= R; ++d) sw.w[d] = raw[R + d] / sum;
  const int64_t n = (int64_t)H * W * CH;
  float* maps = reinterpret_cast<float*>(workspace);
  double* partial = reinterpret_cast<double*>(reinterpret_cast<char*>(workspace) +
                                              ((3 * (size_t)n * sizeof(float) + 255) & ~size_t(255)));
  const dim3 grid((W + kSsimTile - 1) / kSsimTile, (H + kSsimTile - 1) / kSsimTile);
  const int nb = (int)(grid.x * grid.y);
  // loss = m * mean((1 - ssim) / 2) + (1 - m) * mean(base) = sum over the n values of
  //   (m / (2n)) (1 - ssim) + ((1 - m) / n) base;  d loss / d ssim = -m / (2n)
  const float ssim_scale = -ssim_mult / (2.0f * (float)n);
  const float base_scale = (1.0f - ssim_mult) / (float)n;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ssim_fwd, grid, dim3(kSsimBlock), 0, st, out, gt, H, W, CH, R, sw, ssim_scale, base_scale,
                     l2, maps, partial);
  GSMI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ssim_bwd, grid, dim3(kSsimBlock), 0, st, out, gt, H, W, CH, R, sw, base_scale, l2, maps,
                     grad);
  GSMI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ssim_sum, dim3(1), dim3(256), 0, st, partial, nb, 0.0, loss);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}
