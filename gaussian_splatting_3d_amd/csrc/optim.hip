// Training-step stages after the adjoint, on gfx950: photometric loss with
// its gradient seed, and the Adam update of the five parameter groups.
//
// Loss (utils/loss.py:5-22 base term): L1 = mean |out - gt| or L2 =
// mean (out - gt)^2 over all image values; the gradient of the mean w.r.t.
// out is sign(out - gt) / n (L1, sign(0) = 0) or 2 (out - gt) / n (L2).  One
// pass writes the gradient and per-block partial sums; a second launch sums
// the partials in a fixed order (deterministic).
//
// Adam (torch.optim.Adam with betas (0.9, 0.99), as gs/sh_renderer.py:720-729
// builds it): per element, in torch's operation order,
//   m = m + (1 - b1) (g - m);   v = v b2 + (1 - b2) g g;
//   p = p - (lr / (1 - b1^t)) * m / (sqrt(v) / sqrt(1 - b2^t) + eps).
// All groups go through one launch (each block belongs to one group).  With
// `fresh` the moments start at zero and t = 1 without being read or stored:
// the reference rebuilds its optimizer every iteration (main_sh.py:238), so
// every step it takes is such a first step.
#include <algorithm>
#include <cmath>

#include "gsmi_device.h"

namespace gsmi {

constexpr int kLossBlock = 256;
constexpr int kLossItems = 16;

__global__ __launch_bounds__(kLossBlock) void k_loss_grad(const float* __restrict__ out, const float* __restrict__ gt,
                                                         int64_t n, int l2, float scale, float* __restrict__ grad,
                                                         double* __restrict__ partial) {
  __shared__ double s_w[kLossBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kLossBlock * kLossItems;
  double acc = 0.0;
#pragma unroll 4
  for (int k = 0; k < kLossItems; ++k) {
    const int64_t i = base + (int64_t)k * kLossBlock + threadIdx.x;
    if (i < n) {
      const float d = out[i] - gt[i];
      float g;
      if (l2) {
        acc += (double)d * (double)d;
        g = 2.0f * d * scale;
      } else {
        acc += (double)fabsf(d);
        g = d > 0.0f ? scale : (d < 0.0f ? -scale : 0.0f);
      }
      if (grad) grad[i] = g;
    }
  }
  for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partial, int nb, double scale,
                                                      float* __restrict__ loss) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) acc += partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = (float)(s[0] * scale);
}

constexpr int kAdamGroups = 8;
constexpr int kAdamBlock = 256;
constexpr int kAdamItems = 4;
constexpr int64_t kAdamChunk = kAdamBlock * kAdamItems;

struct AdamGroups {
  float* p[kAdamGroups];
  const float* g[kAdamGroups];
  float* m[kAdamGroups];
  float* v[kAdamGroups];
  int64_t n[kAdamGroups];
  int64_t first_block[kAdamGroups + 1];  // block prefix over groups
  float step_size[kAdamGroups];  // lr / (1 - b1^t)
  float bc2_sqrt[kAdamGroups];   // sqrt(1 - b2^t)
  int count;
};

// One element in torch's operation order.
__device__ __forceinline__ void adam_one(float& p, float gr, float& m, float& v, float w1, float b2, float w2,
                                         float eps, float ss, float bs) {
  const float m1 = m + w1 * (gr - m);
  const float v1 = v * b2 + w2 * gr * gr;
  const float denom = sqrtf(v1) / bs + eps;
  p = p + (-ss) * (m1 / denom);
  m = m1;
  v = v1;
}

// w1 = 1 - b1, w2 = 1 - b2 rounded from double, as torch passes them.  Each
// thread takes kAdamItems consecutive elements, as 16-byte vectors when they
// are all in range and every array of the group is 16-byte aligned.
__global__ __launch_bounds__(kAdamBlock) void k_adam(AdamGroups G, float w1, float b2, float w2, float eps,
                                                    int fresh) {
  static_assert(kAdamItems == 4, "one float4 per thread");
  int gi = 0;
  while (gi + 1 < G.count && (int64_t)blockIdx.x >= G.first_block[gi + 1]) ++gi;
  const int64_t i0 = ((int64_t)blockIdx.x - G.first_block[gi]) * kAdamChunk + (int64_t)kAdamItems * threadIdx.x;
  const int64_t n = G.n[gi];
  if (i0 >= n) return;
  float* __restrict__ p = G.p[gi];
  const float* __restrict__ g = G.g[gi];
  float* __restrict__ m = G.m[gi];
  float* __restrict__ v = G.v[gi];
  const float ss = G.step_size[gi], bs = G.bc2_sqrt[gi];
  uintptr_t al = reinterpret_cast<uintptr_t>(p) | reinterpret_cast<uintptr_t>(g);
  if (!fresh) al |= reinterpret_cast<uintptr_t>(m) | reinterpret_cast<uintptr_t>(v);
  if (i0 + kAdamItems <= n && (al & 15) == 0) {
    const float4 gr = *reinterpret_cast<const float4*>(g + i0);
    float4 pp = *reinterpret_cast<const float4*>(p + i0);
    float4 mm = make_float4(0.f, 0.f, 0.f, 0.f), vv = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!fresh) {
      mm = *reinterpret_cast<const float4*>(m + i0);
      vv = *reinterpret_cast<const float4*>(v + i0);
    }
    adam_one(pp.x, gr.x, mm.x, vv.x, w1, b2, w2, eps, ss, bs);
    adam_one(pp.y, gr.y, mm.y, vv.y, w1, b2, w2, eps, ss, bs);
    adam_one(pp.z, gr.z, mm.z, vv.z, w1, b2, w2, eps, ss, bs);
    adam_one(pp.w, gr.w, mm.w, vv.w, w1, b2, w2, eps, ss, bs);
    *reinterpret_cast<float4*>(p + i0) = pp;
    if (!fresh) {
      *reinterpret_cast<float4*>(m + i0) = mm;
      *reinterpret_cast<float4*>(v + i0) = vv;
    }
    return;
  }
  for (int k = 0; k < kAdamItems && i0 + k < n; ++k) {
    const int64_t i = i0 + k;
    float pp = p[i], mm = fresh ? 0.0f : m[i], vv = fresh ? 0.0f : v[i];
    adam_one(pp, g[i], mm, vv, w1, b2, w2, eps, ss, bs);
    p[i] = pp;
    if (!fresh) {
      m[i] = mm;
      v[i] = vv;
    }
  }
}

}  // namespace gsmi

using namespace gsmi;

extern "C" size_t gsmi_loss_workspace_size(int64_t n) {
  const int64_t nb = (n + kLossBlock * kLossItems - 1) / (kLossBlock * kLossItems);
  return (size_t)std::max<int64_t>(nb, 1) * sizeof(double);
}

extern "C" int gsmi_photometric_loss(int64_t n, const float* out, const float* gt, int l2, float* grad, float* loss,
                                     void* workspace, size_t workspace_bytes, void* stream) {
  if (n <= 0 || !out || !gt || !loss) return GSMI_ERR_ARG;
  if (!workspace || workspace_bytes < gsmi_loss_workspace_size(n)) return GSMI_ERR_WORKSPACE;
  const int nb = (int)((n + kLossBlock * kLossItems - 1) / (kLossBlock * kLossItems));
  const float scale = 1.0f / (float)n;
  hipStream_t st = (hipStream_t)stream;
  double* partial = reinterpret_cast<double*>(workspace);
  hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(kLossBlock), 0, st, out, gt, n, l2, scale, grad, partial);
  GSMI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial, nb, 1.0 / (double)n, loss);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_adam_step(int n_groups, float* const* params, const float* const* grads, float* const* exp_avg,
                              float* const* exp_avg_sq, const int64_t* numel, const double* lr, int step,
                              double beta1, double beta2, double eps, int fresh, void* stream) {
  if (n_groups < 1 || n_groups > kAdamGroups || !params || !grads || !numel || !lr || step < 1) return GSMI_ERR_ARG;
  if (!fresh && (!exp_avg || !exp_avg_sq)) return GSMI_ERR_ARG;
  AdamGroups G{};
  G.count = n_groups;
  int64_t blocks = 0;
  const double bc1 = 1.0 - std::pow(beta1, (double)step);
  const double bc2 = 1.0 - std::pow(beta2, (double)step);
  for (int k = 0; k < n_groups; ++k) {
    if (numel[k] < 0 || (numel[k] > 0 && (!params[k] || !grads[k]))) return GSMI_ERR_ARG;
    if (!fresh && numel[k] > 0 && (!exp_avg[k] || !exp_avg_sq[k])) return GSMI_ERR_ARG;
    G.p[k] = params[k];
    G.g[k] = grads[k];
    G.m[k] = fresh ? nullptr : exp_avg[k];
    G.v[k] = fresh ? nullptr : exp_avg_sq[k];
    G.n[k] = numel[k];
    G.first_block[k] = blocks;
    blocks += (numel[k] + kAdamChunk - 1) / kAdamChunk;
    G.step_size[k] = (float)(lr[k] / bc1);
    G.bc2_sqrt[k] = (float)std::sqrt(bc2);
  }
  G.first_block[n_groups] = blocks;
  if (blocks == 0) return GSMI_OK;
  if (blocks >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks), dim3(kAdamBlock), 0, (hipStream_t)stream, G,
                     (float)(1.0 - beta1), (float)beta2, (float)(1.0 - beta2), (float)eps, fresh);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}
