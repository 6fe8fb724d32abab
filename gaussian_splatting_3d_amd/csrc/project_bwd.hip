// Projection adjoint on gfx950: one thread per raw Gaussian, chaining the
// compositing gradients (mean2d, cov2d, alpha at the compacted index) back
// through project_gaussians, the quaternion/scale activations and the masked
// gathers, with the reference's detaches (gs/renderer.py:365-419: the
// Jacobian is built under no_grad and depth is detached, so the mean receives
// gradient only through xy / depth).
//
// Values here feed gradients only (checked to 1e-3 against the float64
// graph), so the exponentials and quotients use the hardware exp2 and
// reciprocal (about 1e-6 relative over the inputs' range) instead of the
// forward's exactly specified sequences, and every load of a thread is
// issued before the visibility branch.
#include "gsmi_device.h"

namespace gsmi {

struct Cam12 {
  float cw[12];
};

__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Per-view terms of the adjoint at a raw Gaussian with mean mu: adds dL/dmu
// (3) and dL/dSigma (3x3) for its compacted record seen through camera cw.
__device__ __forceinline__ void view_terms(const float (&mu)[3], const float* cw, float2 gm, float4 G,
                                           float* __restrict__ g_mean, float* __restrict__ gS) {
  // W = Rc^T, u = W (mu - t)
  float Wm[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) Wm[3 * a + b] = cw[4 * b + a];
  const float d0 = mu[0] - cw[3], d1 = mu[1] - cw[7], d2 = mu[2] - cw[11];
  float u[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) u[a] = Wm[3 * a] * d0 + Wm[3 * a + 1] * d1 + Wm[3 * a + 2] * d2;
  const float z = u[2];
  const float iz = fast_rcp(z);
  // mean: dL/du = (g.x / z, g.y / z, 0); dL/dmu = W^T dL/du.
  const float gu0 = gm.x * iz, gu1 = gm.y * iz;
#pragma unroll
  for (int b = 0; b < 3; ++b) g_mean[b] += Wm[b] * gu0 + Wm[3 + b] * gu1;
  // A = J W (first two rows), J00 = J11 = 1/z, J02 = -x/z^2, J12 = -y/z^2.
  const float J02 = -u[0] * iz * iz, J12 = -u[1] * iz * iz;
  float A[6];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    A[k] = iz * Wm[k] + J02 * Wm[6 + k];
    A[3 + k] = iz * Wm[3 + k] + J12 * Wm[6 + k];
  }
  // dL/dSigma = A^T G A with G the 2x2 cov2d gradient (all four entries).
  const float Gm[4] = {G.x, G.y, G.z, G.w};
  float GA[6];  // G A : 2x3
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int b = 0; b < 3; ++b) GA[3 * r + b] = Gm[2 * r] * A[b] + Gm[2 * r + 1] * A[3 + b];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) gS[3 * a + b] += A[a] * GA[b] + A[3 + a] * GA[3 + b];
}

// Raw parameters of Gaussian i, loaded together.
struct RawParams {
  float mu[3];
  float4 q;
  float ls[3];
  float lg;
};

__device__ __forceinline__ RawParams load_params(int64_t i, const float* __restrict__ mean,
                                                 const float* __restrict__ qvec, const float* __restrict__ log_s,
                                                 const float* __restrict__ logit) {
  RawParams r;
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    r.mu[b] = mean[3 * i + b];
    r.ls[b] = log_s[3 * i + b];
  }
  r.q = reinterpret_cast<const float4*>(qvec)[i];
  r.lg = logit[i];
  return r;
}

// dL/dSigma and dL/dalpha -> gradients of qvec, log-scale and logit of raw
// Gaussian i (writes the outputs of i).
__device__ __forceinline__ void param_chain(int64_t i, const RawParams& rp, const float* gS, float g_alpha,
                                            float* __restrict__ o_q, float* __restrict__ o_ls,
                                            float* __restrict__ o_logit) {
  const float4 q = rp.q;
  const float nq = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  const float den = fmaxf(nq, 1e-12f);
  const float iden = fast_rcp(den);
  const float w = q.x * iden, x = q.y * iden, y = q.z * iden, zq = q.w * iden;
  float R[9];
  R[0] = 1.0f - 2.0f * (y * y + zq * zq);
  R[1] = 2.0f * (x * y - w * zq);
  R[2] = 2.0f * (x * zq + w * y);
  R[3] = 2.0f * (x * y + w * zq);
  R[4] = 1.0f - 2.0f * (x * x + zq * zq);
  R[5] = 2.0f * (y * zq - w * x);
  R[6] = 2.0f * (x * zq - w * y);
  R[7] = 2.0f * (y * zq + w * x);
  R[8] = 1.0f - 2.0f * (x * x + y * y);
  const float s[3] = {fast_exp(rp.ls[0]), fast_exp(rp.ls[1]), fast_exp(rp.ls[2])};
  float M[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) M[3 * a + b] = R[3 * a + b] * s[b];
  // Sigma = M M^T  ->  dL/dM = (gS + gS^T) M
  float gM[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float acc = 0.0f;
#pragma unroll
      for (int b = 0; b < 3; ++b) acc = fmaf(gS[3 * a + b] + gS[3 * b + a], M[3 * b + c], acc);
      gM[3 * a + c] = acc;
    }
  // M = R diag(s)
  float gR[9], gs[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      gR[3 * a + b] = gM[3 * a + b] * s[b];
      gs[b] = fmaf(gM[3 * a + b], R[3 * a + b], gs[b]);
    }
  // d R / d (w,x,y,z) of the unit-quaternion matrix.
  const float gw = 2.0f * (-zq * gR[1] + y * gR[2] + zq * gR[3] - x * gR[5] - y * gR[6] + x * gR[7]);
  const float gx = 2.0f * (y * gR[1] + zq * gR[2] + y * gR[3] - 2.0f * x * gR[4] - w * gR[5] + zq * gR[6] +
                           w * gR[7] - 2.0f * x * gR[8]);
  const float gy = 2.0f * (-2.0f * y * gR[0] + x * gR[1] + w * gR[2] + x * gR[3] + zq * gR[5] - w * gR[6] +
                           zq * gR[7] - 2.0f * y * gR[8]);
  const float gz = 2.0f * (-2.0f * zq * gR[0] - w * gR[1] + x * gR[2] + w * gR[3] - 2.0f * zq * gR[4] +
                           y * gR[5] + x * gR[6] + y * gR[7]);
  // q_hat = q / max(|q|, eps): dq = (dq_hat - q_hat (q_hat . dq_hat)) / |q|.
  // (den == nq when nq > 1e-12, so one reciprocal serves both cases)
  const float dotp = nq > 1e-12f ? w * gw + x * gx + y * gy + zq * gz : 0.0f;
  const float4 gq = make_float4((gw - w * dotp) * iden, (gx - x * dotp) * iden, (gy - y * dotp) * iden,
                                (gz - zq * dotp) * iden);
  reinterpret_cast<float4*>(o_q)[i] = gq;
  // s = exp(log_s)
#pragma unroll
  for (int b = 0; b < 3; ++b) o_ls[3 * i + b] = gs[b] * s[b];
  // alpha = sigmoid(logit)
  const float al = fast_rcp(1.0f + fast_exp(-rp.lg));
  o_logit[i] = g_alpha * (al * (1.0f - al));
}

__device__ __forceinline__ void write_zero(int64_t i, float* __restrict__ o_mean, float* __restrict__ o_q,
                                           float* __restrict__ o_ls, float* __restrict__ o_logit) {
  o_mean[3 * i] = o_mean[3 * i + 1] = o_mean[3 * i + 2] = 0.0f;
  reinterpret_cast<float4*>(o_q)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  o_ls[3 * i] = o_ls[3 * i + 1] = o_ls[3 * i + 2] = 0.0f;
  o_logit[i] = 0.0f;
}

// Single view: c2w by value.
__global__ __launch_bounds__(256) void k_project_bwd(int64_t N, const float* __restrict__ mean,
                                                     const float* __restrict__ qvec,
                                                     const float* __restrict__ log_s,
                                                     const float* __restrict__ logit, Cam12 cam,
                                                     const uint8_t* __restrict__ vis,
                                                     const int32_t* __restrict__ voff,
                                                     const float* __restrict__ g_m2d,
                                                     const float* __restrict__ g_cov,
                                                     const float* __restrict__ g_alpha, float* __restrict__ o_mean,
                                                     float* __restrict__ o_q, float* __restrict__ o_ls,
                                                     float* __restrict__ o_logit) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  // vis, the record index and the raw parameters in one round of loads; the
  // record's gradients (index from voff, valid for visible i) in the next.
  const bool visible = vis[i] != 0;
  const int64_t v = voff[i];
  const RawParams rp = load_params(i, mean, qvec, log_s, logit);
  if (!visible) {
    write_zero(i, o_mean, o_q, o_ls, o_logit);
    return;
  }
  const float2 g2 = reinterpret_cast<const float2*>(g_m2d)[v];
  const float4 gc = reinterpret_cast<const float4*>(g_cov)[v];
  const float ga = g_alpha[v];
  float gm[3] = {0.f, 0.f, 0.f}, gS[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  view_terms(rp.mu, cam.cw, g2, gc, gm, gS);
#pragma unroll
  for (int b = 0; b < 3; ++b) o_mean[3 * i + b] = gm[b];
  param_chain(i, rp, gS, ga, o_q, o_ls, o_logit);
}

// Multi-view: the gradients of every view that sees Gaussian i are summed
// (views in order, so the sum is deterministic) before the parameter chain,
// which is linear in dL/dSigma and dL/dalpha.
__global__ __launch_bounds__(256) void k_project_bwd_views(
    int64_t N, int B, const float* __restrict__ mean, const float* __restrict__ qvec,
    const float* __restrict__ log_s, const float* __restrict__ logit, const gsmi_view* __restrict__ views,
    const uint8_t* __restrict__ vis, const int32_t* __restrict__ voff, const float* __restrict__ g_m2d,
    const float* __restrict__ g_cov, const float* __restrict__ g_alpha, float* __restrict__ o_mean,
    float* __restrict__ o_q, float* __restrict__ o_ls, float* __restrict__ o_logit) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const RawParams rp = load_params(i, mean, qvec, log_s, logit);
  float gm[3] = {0.f, 0.f, 0.f}, gS[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float ga = 0.0f;
  bool any = false;
  for (int b = 0; b < B; ++b) {
    const int64_t e = (int64_t)b * N + i;
    if (!vis[e]) continue;
    any = true;
    const int64_t v = voff[e];
    float cw[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) cw[k] = views[b].c2w[k];
    view_terms(rp.mu, cw, reinterpret_cast<const float2*>(g_m2d)[v], reinterpret_cast<const float4*>(g_cov)[v], gm,
               gS);
    ga += g_alpha[v];
  }
  if (!any) {
    write_zero(i, o_mean, o_q, o_ls, o_logit);
    return;
  }
#pragma unroll
  for (int b = 0; b < 3; ++b) o_mean[3 * i + b] = gm[b];
  param_chain(i, rp, gS, ga, o_q, o_ls, o_logit);
}

}  // namespace gsmi

using namespace gsmi;

extern "C" int gsmi_project_backward(int64_t N, const float* mean, const float* qvec, const float* log_s,
                                     const float* logit, const float* c2w_host, const uint8_t* vis,
                                     const int32_t* vis_off, const float* grad_mean2d, const float* grad_cov,
                                     const float* grad_alpha, float* grad_mean, float* grad_qvec,
                                     float* grad_log_s, float* grad_logit, void* stream) {
  if (N < 0 || !c2w_host) return GSMI_ERR_ARG;
  if (N == 0) return GSMI_OK;
  if (!mean || !qvec || !log_s || !logit || !vis || !vis_off || !grad_mean2d || !grad_cov || !grad_alpha ||
      !grad_mean || !grad_qvec || !grad_log_s || !grad_logit)
    return GSMI_ERR_ARG;
  if (((uintptr_t)qvec & 15) || ((uintptr_t)grad_qvec & 15) || ((uintptr_t)grad_cov & 15) ||
      ((uintptr_t)grad_mean2d & 7))
    return GSMI_ERR_ARG;
  Cam12 cam;
  for (int k = 0; k < 12; ++k) cam.cw[k] = c2w_host[k];
  hipLaunchKernelGGL(k_project_bwd, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream, N, mean,
                     qvec, log_s, logit, cam, vis, vis_off, grad_mean2d, grad_cov, grad_alpha, grad_mean, grad_qvec,
                     grad_log_s, grad_logit);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_project_backward_views(int64_t N, int B, const gsmi_view* views, const float* mean,
                                           const float* qvec, const float* log_s, const float* logit,
                                           const uint8_t* vis, const int32_t* vis_off, const float* grad_mean2d,
                                           const float* grad_cov, const float* grad_alpha, float* grad_mean,
                                           float* grad_qvec, float* grad_log_s, float* grad_logit, void* stream) {
  if (N < 0 || B <= 0 || !views) return GSMI_ERR_ARG;
  if (N == 0) return GSMI_OK;
  if (!mean || !qvec || !log_s || !logit || !vis || !vis_off || !grad_mean2d || !grad_cov || !grad_alpha ||
      !grad_mean || !grad_qvec || !grad_log_s || !grad_logit)
    return GSMI_ERR_ARG;
  if (((uintptr_t)qvec & 15) || ((uintptr_t)grad_qvec & 15) || ((uintptr_t)grad_cov & 15) ||
      ((uintptr_t)grad_mean2d & 7) || ((uintptr_t)views & 7))
    return GSMI_ERR_ARG;
  hipLaunchKernelGGL(k_project_bwd_views, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream, N,
                     B, mean, qvec, log_s, logit, views, vis, vis_off, grad_mean2d, grad_cov, grad_alpha, grad_mean,
                     grad_qvec, grad_log_s, grad_logit);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}
