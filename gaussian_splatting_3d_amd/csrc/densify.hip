// Adaptive density control on gfx950 (gs/sh_renderer.py:423-661): gradient
// statistics, split / clone and pruning as stream compaction over the raw
// structure-of-arrays parameters.
//
// Statistics (update_grads, :602-623, with the visibility count of
// forward, :215-216): acc[i] = max(acc[i], |g_i|) or acc[i] + |g_i| with
// g_i the mean gradient of Gaussian i (3D: [N,3]; 2D: the compacted [V,2]
// gradient of the visible ones), cnt[i] += vis[i].
//
// Split / clone (split_gaussians, :426-534): grad = acc / (cnt + 1e-5)
// ("mean") or acc ("max"); densify = grad > pos_grad_thresh; split =
// densify and any(exp(log_s) > split_scale_thresh); clone = densify and not
// split.  New order: the N - k unsplit Gaussians (including clone sources)
// in index order, then the clones in index order, then the k split ones
// twice (rows 0..k-1 then again 0..k-1, as .repeat(2, 1) lays them out),
// each with mean + R^T (noise * s) (R = rotation of the normalised
// quaternion; the reference applies the transpose) and log-scale
// log(s / shrink); the noise [2k, 3] is an input (torch.randn in the
// reference).
//
// Pruning (remove_low_alpha_gaussians, :536-558): keep sigmoid(logit) >=
// alpha_thresh, order preserved.
#include "project.h"
#include "scan.h"

namespace gsmi {

__global__ void k_grad_stats(int64_t n, const float* __restrict__ g, int dim, const int32_t* __restrict__ idx,
                             int use_max, float* __restrict__ acc) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float s = 0.0f;
  for (int k = 0; k < dim; ++k) s = s + g[dim * r + k] * g[dim * r + k];
  const float nrm = sqrtf(s);
  const int64_t i = idx ? idx[r] : r;
  acc[i] = use_max ? fmaxf(acc[i], nrm) : acc[i] + nrm;
}

__global__ void k_count_vis(int64_t N, const uint8_t* __restrict__ vis, int32_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N && vis[i]) cnt[i] += 1;
}

// Per Gaussian: Int2 {keep-in-place flag, extra rows} with class split = 2
// extra rows minus the removed original, clone = 1 extra row.
__global__ void k_split_class(int64_t N, const float* __restrict__ acc, const int32_t* __restrict__ cnt,
                              int use_mean, float grad_thresh, const float* __restrict__ log_s, float scale_thresh,
                              uint8_t* __restrict__ cls, Int2* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float gr = use_mean ? acc[i] / ((float)cnt[i] + 1e-5f) : acc[i];
  const bool densify = gr > grad_thresh;
  bool big = false;
  for (int k = 0; k < 3; ++k) big = big || (exp_spec(log_s[3 * i + k]) > scale_thresh);
  const uint8_t c = densify ? (big ? 2 : 1) : 0;  // 0 keep, 1 clone, 2 split
  cls[i] = c;
  flags[i] = Int2{c == 2 ? 0 : 1, c == 1 ? 1 : 0};  // (unsplit rank, clone rank)
}

__global__ void k_split_rank(int64_t N, const uint8_t* __restrict__ cls, int32_t* __restrict__ split_rank) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) split_rank[i] = cls[i] == 2 ? 1 : 0;
}

struct Soa {
  const float *mean, *qvec, *log_s, *logit, *sh;
};
struct SoaOut {
  float *mean, *qvec, *log_s, *logit, *sh;
};

__device__ __forceinline__ void copy_row(const Soa& in, int64_t i, const SoaOut& out, int64_t o, int sh_row) {
  for (int k = 0; k < 3; ++k) out.mean[3 * o + k] = in.mean[3 * i + k];
  reinterpret_cast<float4*>(out.qvec)[o] = reinterpret_cast<const float4*>(in.qvec)[i];
  for (int k = 0; k < 3; ++k) out.log_s[3 * o + k] = in.log_s[3 * i + k];
  out.logit[o] = in.logit[i];
  for (int k = 0; k < sh_row; ++k) out.sh[(int64_t)sh_row * o + k] = in.sh[(int64_t)sh_row * i + k];
}

// flags = exclusive scan of (unsplit, clone); srank = exclusive scan of split.
__global__ void k_split_write(int64_t N, Soa in, SoaOut out, int sh_row, const uint8_t* __restrict__ cls,
                              const Int2* __restrict__ off, const Int2* __restrict__ tot,
                              const int32_t* __restrict__ srank, const int32_t* __restrict__ stot,
                              const float* __restrict__ noise, float shrink) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint8_t c = cls[i];
  const int64_t n_unsplit = tot->x, n_clone = tot->y, k = *stot;
  if (c != 2) copy_row(in, i, out, off[i].x, sh_row);
  if (c == 1) copy_row(in, i, out, n_unsplit + off[i].y, sh_row);
  if (c == 2) {
    const int64_t r = srank[i];
    const float4 q = reinterpret_cast<const float4*>(in.qvec)[i];
    float R[9];
    quat_to_rot(q, R);
    float s[3], ls[3];
    for (int d = 0; d < 3; ++d) {
      s[d] = exp_spec(in.log_s[3 * i + d]);
      ls[d] = logf(s[d] / shrink);
    }
    for (int rep = 0; rep < 2; ++rep) {
      const int64_t row = rep * k + r;  // .repeat(2, 1): all k, then all k again
      const int64_t o = n_unsplit + n_clone + row;
      copy_row(in, i, out, o, sh_row);
      float gn[3];
      for (int d = 0; d < 3; ++d) gn[d] = noise[3 * row + d] * s[d];
      for (int a = 0; a < 3; ++a)  // (R^T gn)_a = sum_b R[b][a] gn[b]
        out.mean[3 * o + a] = in.mean[3 * i + a] + (R[0 * 3 + a] * gn[0] + R[1 * 3 + a] * gn[1] + R[2 * 3 + a] * gn[2]);
      for (int d = 0; d < 3; ++d) out.log_s[3 * o + d] = ls[d];
    }
  }
}

__global__ void k_alpha_keep(int64_t N, const float* __restrict__ logit, float alpha_thresh, int32_t* __restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) keep[i] = (1.0f / (1.0f + exp_spec(-logit[i]))) >= alpha_thresh ? 1 : 0;
}

__global__ void k_compact(int64_t N, Soa in, SoaOut out, int sh_row, const int32_t* __restrict__ keep_flag,
                          const int32_t* __restrict__ off) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N && keep_flag[i]) copy_row(in, i, out, off[i], sh_row);
}

static size_t au(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace gsmi

using namespace gsmi;

extern "C" int gsmi_grad_stats(int64_t n, const float* grad, int dim, const int32_t* index, int use_max, float* acc,
                               void* stream) {
  if (n < 0 || (dim != 2 && dim != 3) || (n > 0 && (!grad || !acc))) return GSMI_ERR_ARG;
  if (n == 0) return GSMI_OK;
  hipLaunchKernelGGL(k_grad_stats, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, grad, dim,
                     index, use_max, acc);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_count_visible(int64_t N, const uint8_t* vis, int32_t* cnt, void* stream) {
  if (N < 0 || (N > 0 && (!vis || !cnt))) return GSMI_ERR_ARG;
  if (N == 0) return GSMI_OK;
  hipLaunchKernelGGL(k_count_vis, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream, N, vis, cnt);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" size_t gsmi_densify_workspace_size(int64_t N) {
  if (N < 0) N = 0;
  return au(N) + au(N * 8) + au(N * 4) + au(scan_workspace_bytes(N, 8)) + 512;
}

// Phase 1: classify and count; synchronises and writes the host array
// counts[3] = (unsplit, clone, split): N_new = unsplit + clone + 2 split.
extern "C" int gsmi_split_clone_count(int64_t N, const float* acc, const int32_t* cnt, int use_mean,
                                      float grad_thresh, const float* log_s, float scale_thresh, int64_t* counts,
                                      void* workspace, size_t workspace_bytes, void* stream) {
  if (N < 0 || !counts || (N > 0 && (!acc || !log_s || (use_mean && !cnt)))) return GSMI_ERR_ARG;
  if (N >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  if (!workspace || workspace_bytes < gsmi_densify_workspace_size(N)) return GSMI_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* p = reinterpret_cast<char*>(workspace);
  uint8_t* cls = reinterpret_cast<uint8_t*>(p);
  p += au(N);
  Int2* flags = reinterpret_cast<Int2*>(p);
  p += au(N * 8);
  int32_t* srank = reinterpret_cast<int32_t*>(p);
  p += au(N * 4);
  void* scan_ws = p;
  p += au(scan_workspace_bytes(N, 8));
  Int2* tot = reinterpret_cast<Int2*>(p);
  int32_t* stot = reinterpret_cast<int32_t*>(p + 256);
  const unsigned g = (unsigned)((N + 255) / 256);
  if (N > 0) {
    hipLaunchKernelGGL(k_split_class, dim3(g), dim3(256), 0, st, N, acc, cnt, use_mean, grad_thresh, log_s,
                       scale_thresh, cls, flags);
    hipLaunchKernelGGL(k_split_rank, dim3(g), dim3(256), 0, st, N, cls, srank);
    GSMI_CHECK_LAUNCH();
  }
  int rc = exclusive_scan<Int2>(flags, flags, N, tot, scan_ws, st);
  if (!rc) rc = exclusive_scan<int32_t>(srank, srank, N, stot, scan_ws, st);
  if (rc) return rc;
  // counts = (unsplit, clone, split), read back like the reference's .item()
  Int2 h_tot;
  int32_t h_stot;
  if (hipMemcpyAsync(&h_tot, tot, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&h_stot, stot, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return GSMI_ERR_LAUNCH;
  counts[0] = h_tot.x;
  counts[1] = h_tot.y;
  counts[2] = h_stot;
  return GSMI_OK;
}

// Phase 2: write the new arrays [N + clone + split rows] (uses the workspace
// of phase 1 unchanged).  noise: [2 * split, 3] standard normal samples.
extern "C" int gsmi_split_clone_write(int64_t N, const float* mean, const float* qvec, const float* log_s,
                                      const float* logit, const float* sh, int sh_row, const float* noise,
                                      float shrink, float* out_mean, float* out_qvec, float* out_log_s,
                                      float* out_logit, float* out_sh, void* workspace, size_t workspace_bytes,
                                      void* stream) {
  if (N < 0 || sh_row <= 0 || !(shrink > 0.0f)) return GSMI_ERR_ARG;
  if (!workspace || workspace_bytes < gsmi_densify_workspace_size(N)) return GSMI_ERR_WORKSPACE;
  if (N == 0) return GSMI_OK;
  if (!mean || !qvec || !log_s || !logit || !sh || !out_mean || !out_qvec || !out_log_s || !out_logit || !out_sh)
    return GSMI_ERR_ARG;
  if (((uintptr_t)qvec & 15) || ((uintptr_t)out_qvec & 15)) return GSMI_ERR_ARG;
  char* p = reinterpret_cast<char*>(workspace);
  const uint8_t* cls = reinterpret_cast<const uint8_t*>(p);
  p += au(N);
  const Int2* flags = reinterpret_cast<const Int2*>(p);
  p += au(N * 8);
  const int32_t* srank = reinterpret_cast<const int32_t*>(p);
  p += au(N * 4);
  p += au(scan_workspace_bytes(N, 8));
  const Int2* tot = reinterpret_cast<const Int2*>(p);
  const int32_t* stot = reinterpret_cast<const int32_t*>(p + 256);
  hipLaunchKernelGGL(k_split_write, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream, N,
                     Soa{mean, qvec, log_s, logit, sh}, SoaOut{out_mean, out_qvec, out_log_s, out_logit, out_sh},
                     sh_row, cls, flags, tot, srank, stot, noise, shrink);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

// Pruning by opacity: phase 1 counts (synchronises; host *kept), phase 2 writes.
extern "C" int gsmi_prune_alpha_count(int64_t N, const float* logit, float alpha_thresh, int64_t* kept,
                                      void* workspace, size_t workspace_bytes, void* stream) {
  if (N < 0 || !kept || (N > 0 && !logit)) return GSMI_ERR_ARG;
  if (N >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  if (!workspace || workspace_bytes < gsmi_densify_workspace_size(N)) return GSMI_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* p = reinterpret_cast<char*>(workspace);
  int32_t* keep = reinterpret_cast<int32_t*>(p + au(N));
  int32_t* off = reinterpret_cast<int32_t*>(p + au(N) + au(N * 8));
  void* scan_ws = p + au(N) + au(N * 8) + au(N * 4);
  int32_t* tot = reinterpret_cast<int32_t*>(p + au(N) + au(N * 8) + au(N * 4) + au(scan_workspace_bytes(N, 8)));
  if (N > 0) {
    hipLaunchKernelGGL(k_alpha_keep, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, N, logit, alpha_thresh,
                       keep);
    GSMI_CHECK_LAUNCH();
  }
  int rc = exclusive_scan<int32_t>(keep, off, N, tot, scan_ws, st);
  if (rc) return rc;
  int32_t h = 0;
  if (hipMemcpyAsync(&h, tot, 4, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return GSMI_ERR_LAUNCH;
  *kept = h;
  return GSMI_OK;
}

extern "C" int gsmi_prune_alpha_write(int64_t N, const float* mean, const float* qvec, const float* log_s,
                                      const float* logit, const float* sh, int sh_row, float* out_mean,
                                      float* out_qvec, float* out_log_s, float* out_logit, float* out_sh,
                                      void* workspace, size_t workspace_bytes, void* stream) {
  if (N < 0 || sh_row <= 0) return GSMI_ERR_ARG;
  if (!workspace || workspace_bytes < gsmi_densify_workspace_size(N)) return GSMI_ERR_WORKSPACE;
  if (N == 0) return GSMI_OK;
  if (!mean || !qvec || !log_s || !logit || !sh || !out_mean || !out_qvec || !out_log_s || !out_logit || !out_sh)
    return GSMI_ERR_ARG;
  if (((uintptr_t)qvec & 15) || ((uintptr_t)out_qvec & 15)) return GSMI_ERR_ARG;
  char* p = reinterpret_cast<char*>(workspace);
  const int32_t* keep = reinterpret_cast<const int32_t*>(p + au(N));
  const int32_t* off = reinterpret_cast<const int32_t*>(p + au(N) + au(N * 8));
  hipLaunchKernelGGL(k_compact, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream, N,
                     Soa{mean, qvec, log_s, logit, sh}, SoaOut{out_mean, out_qvec, out_log_s, out_logit, out_sh},
                     sh_row, keep, off);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}
