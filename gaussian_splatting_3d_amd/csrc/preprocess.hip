// Per-Gaussian preprocessing on gfx950: frustum cull, projection, tile
// bounds, pair counts, compaction and key emission, in structure-of-arrays
// passes over the N raw Gaussians.  Every value that feeds a decision or a
// key follows the reference's operation order (gs/renderer.py:365-419,
// gs/culling.py:8-37, culling.h:11-20) with the specified exponential, so
// masks, tile bounds, depths and keys are bit-identical to the oracle's.
#include "scan.h"

namespace gsmi {

struct Planes {
  float n[18];
  float p[18];
};

// kernels.h:156-170 intersect_sphere_frustum.
__device__ __forceinline__ bool in_frustum(float mx, float my, float mz, float r, const float* __restrict__ nrm,
                                           const float* __restrict__ pts) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const float dx = mx - pts[3 * k], dy = my - pts[3 * k + 1], dz = mz - pts[3 * k + 2];
    const float d = dx * nrm[3 * k] + dy * nrm[3 * k + 1] + dz * nrm[3 * k + 2];
    if (!(d > -r)) return false;
  }
  return true;
}

__global__ void k_cull(int64_t N, const float* __restrict__ mean, const float* __restrict__ svec,
                       const float* __restrict__ nrm, const float* __restrict__ pts, uint8_t* __restrict__ mask,
                       float thresh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float r = fmaxf(fmaxf(svec[3 * i], svec[3 * i + 1]), svec[3 * i + 2]) * thresh;
  mask[i] = in_frustum(mean[3 * i], mean[3 * i + 1], mean[3 * i + 2], r, nrm, pts) ? 1 : 0;
}

// utils/transforms.py:31-45 (kornia WXYZ): normalise, then the unit
// quaternion rotation matrix, row-major.
__device__ __forceinline__ void quat_to_rot(float4 q, float* R) {
  const float nrm = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  const float den = fmaxf(nrm, 1e-12f);
  const float w = q.x / den, x = q.y / den, y = q.z / den, z = q.w / den;
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

struct Proj {
  float mx, my;  // mean2d
  float c[4];    // cov2d row-major
  float depth;
};

// gs/renderer.py:380-419 with c2w = [Rc | t]: u = Rc^T (mu - t),
// Sigma = (R(q) diag s)(R(q) diag s)^T, A = J Rc^T (J detached),
// cov2d = (A Sigma A^T)[:2,:2], depth = u_z, mean2d = u_xy / depth.
__device__ __forceinline__ Proj project_one(float m0, float m1, float m2, float4 q, float s0, float s1, float s2,
                                            const float* __restrict__ cw) {
  float Wm[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Wm[3 * i + j] = cw[4 * j + i];
  const float d0 = m0 + (-cw[3]), d1 = m1 + (-cw[7]), d2 = m2 + (-cw[11]);
  float u[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) u[i] = Wm[3 * i + 0] * d0 + Wm[3 * i + 1] * d1 + Wm[3 * i + 2] * d2;
  float R[9], M[9], S[9];
  quat_to_rot(q, R);
  const float s[3] = {s0, s1, s2};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[3 * i + j] = s[j] * R[3 * i + j];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      S[3 * a + b] = M[3 * a + 0] * M[3 * b + 0] + M[3 * a + 1] * M[3 * b + 1] + M[3 * a + 2] * M[3 * b + 2];
  const float x = u[0], y = u[1], z = u[2];
  const float J00 = 1.0f / z, J02 = -x / z / z;
  const float J11 = 1.0f / z, J12 = -y / z / z;
  float A[6];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    A[k] = J00 * Wm[0 * 3 + k] + J02 * Wm[2 * 3 + k];
    A[3 + k] = J11 * Wm[1 * 3 + k] + J12 * Wm[2 * 3 + k];
  }
  float T1[6];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      T1[3 * i + b] = A[3 * i + 0] * S[0 * 3 + b] + A[3 * i + 1] * S[1 * 3 + b] + A[3 * i + 2] * S[2 * 3 + b];
  Proj p;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
      p.c[2 * i + j] = T1[3 * i + 0] * A[3 * j + 0] + T1[3 * i + 1] * A[3 * j + 1] + T1[3 * i + 2] * A[3 * j + 2];
  p.depth = z;
  p.mx = x / z;
  p.my = y / z;
  return p;
}

// gs/culling.py:16-31 with utils/camera.py:296-301 for one axis.
__device__ __forceinline__ void aabb_axis(float m, float cii, float D, float f, float c, int lim, int tile,
                                          int* lo, int* hi) {
  const float e = sqrtf(D * cii);
  const float pl = (m - e) * f;
  const float ph = (m + e) * f;
  int il = f2i_sat(pl + c);
  int ih = f2i_sat(ph + c);
  il = il < 0 ? 0 : (il > lim ? lim : il);
  ih = ih < 0 ? 0 : (ih > lim ? lim : ih);
  *lo = il / tile;
  *hi = ih / tile;
}

struct CamArgs {
  float cw[12];
  float fx, fy, cx, cy, D, frustum_radius;
  int W, H, tile, skip_culling, n_tiles_w;
};

__global__ void k_aabb(int64_t V, const float* __restrict__ mean2d, const float* __restrict__ cov2d, CamArgs c,
                       int32_t* __restrict__ tl, int32_t* __restrict__ br, int32_t* __restrict__ cnt) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const float2 m = reinterpret_cast<const float2*>(mean2d)[v];
  const float4 cv = reinterpret_cast<const float4*>(cov2d)[v];
  int lx, hx, ly, hy;
  aabb_axis(m.x, cv.x, c.D, c.fx, c.cx, c.W - 1, c.tile, &lx, &hx);
  aabb_axis(m.y, cv.w, c.D, c.fy, c.cy, c.H - 1, c.tile, &ly, &hy);
  reinterpret_cast<int2*>(tl)[v] = make_int2(lx, ly);
  reinterpret_cast<int2*>(br)[v] = make_int2(hx, hy);
  cnt[v] = (hx - lx + 1) * (hy - ly + 1);
}

// Pass 1 over raw Gaussians: visibility and pair count.
__global__ __launch_bounds__(256) void k_pre_count_planes(int64_t N, const float* __restrict__ mean,
                                                          const float* __restrict__ qvec,
                                                          const float* __restrict__ log_s, Planes pl, CamArgs c,
                                                          uint8_t* __restrict__ vis, Int2* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float m0 = mean[3 * i], m1 = mean[3 * i + 1], m2 = mean[3 * i + 2];
  const float s0 = exp_spec(log_s[3 * i]), s1 = exp_spec(log_s[3 * i + 1]), s2 = exp_spec(log_s[3 * i + 2]);
  bool keep = true;
  if (!c.skip_culling) {
    const float r = fmaxf(fmaxf(s0, s1), s2) * c.frustum_radius;
    keep = in_frustum(m0, m1, m2, r, pl.n, pl.p);
  }
  int pairs = 0;
  if (keep) {
    const float4 q = reinterpret_cast<const float4*>(qvec)[i];
    const Proj p = project_one(m0, m1, m2, q, s0, s1, s2, c.cw);
    int lx, hx, ly, hy;
    aabb_axis(p.mx, p.c[0], c.D, c.fx, c.cx, c.W - 1, c.tile, &lx, &hx);
    aabb_axis(p.my, p.c[3], c.D, c.fy, c.cy, c.H - 1, c.tile, &ly, &hy);
    pairs = (hx - lx + 1) * (hy - ly + 1);
  }
  vis[i] = keep ? 1 : 0;
  cnt[i] = Int2{keep ? 1 : 0, pairs};
}

// Pass 2: compacted per-Gaussian records.
__global__ __launch_bounds__(256) void k_pre_emit(int64_t N, const float* __restrict__ mean,
                                                  const float* __restrict__ qvec, const float* __restrict__ log_s,
                                                  const float* __restrict__ logit, CamArgs c,
                                                  const uint8_t* __restrict__ vis, const int32_t* __restrict__ voff,
                                                  float* __restrict__ mean2d, float* __restrict__ cov2d,
                                                  float* __restrict__ depth, float* __restrict__ alpha,
                                                  int32_t* __restrict__ tl, int32_t* __restrict__ br,
                                                  int32_t* __restrict__ orig) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || !vis[i]) return;
  const int64_t v = voff[i];
  const float m0 = mean[3 * i], m1 = mean[3 * i + 1], m2 = mean[3 * i + 2];
  const float s0 = exp_spec(log_s[3 * i]), s1 = exp_spec(log_s[3 * i + 1]), s2 = exp_spec(log_s[3 * i + 2]);
  const float4 q = reinterpret_cast<const float4*>(qvec)[i];
  const Proj p = project_one(m0, m1, m2, q, s0, s1, s2, c.cw);
  int lx, hx, ly, hy;
  aabb_axis(p.mx, p.c[0], c.D, c.fx, c.cx, c.W - 1, c.tile, &lx, &hx);
  aabb_axis(p.my, p.c[3], c.D, c.fy, c.cy, c.H - 1, c.tile, &ly, &hy);
  reinterpret_cast<float2*>(mean2d)[v] = make_float2(p.mx, p.my);
  reinterpret_cast<float4*>(cov2d)[v] = make_float4(p.c[0], p.c[1], p.c[2], p.c[3]);
  depth[v] = p.depth;
  alpha[v] = 1.0f / (1.0f + exp_spec(-logit[i]));
  reinterpret_cast<int2*>(tl)[v] = make_int2(lx, ly);
  reinterpret_cast<int2*>(br)[v] = make_int2(hx, hy);
  orig[v] = (int32_t)i;
}

__global__ void k_split_int2(const Int2* __restrict__ in, int64_t n, int32_t* __restrict__ a,
                             int32_t* __restrict__ b, const Int2* __restrict__ tot, int64_t* __restrict__ totals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a[i] = in[i].x;
    b[i] = in[i].y;
  }
  if (i == 0) {
    totals[0] = tot->x;
    totals[1] = tot->y;
  }
}

__global__ void k_exp_spec(int64_t n, const float* __restrict__ x, float* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = exp_spec(x[i]);
}

static size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

static CamArgs make_cam(const float* c2w_host, float fx, float fy, float cx, float cy, int W, int H, float D,
                        float frustum_radius, int skip, int tile) {
  CamArgs c{};
  if (c2w_host)
    for (int k = 0; k < 12; ++k) c.cw[k] = c2w_host[k];
  c.fx = fx;
  c.fy = fy;
  c.cx = cx;
  c.cy = cy;
  c.D = D;
  c.frustum_radius = frustum_radius;
  c.W = W;
  c.H = H;
  c.tile = tile;
  c.skip_culling = skip;
  c.n_tiles_w = (W + tile - 1) / tile;
  return c;
}

}  // namespace gsmi

using namespace gsmi;

extern "C" const char* gsmi_status_string(int s) {
  switch (s) {
    case GSMI_OK: return "ok";
    case GSMI_ERR_ARG: return "invalid argument";
    case GSMI_ERR_UNSUPPORTED: return "unsupported configuration";
    case GSMI_ERR_LAUNCH: return "kernel launch or HIP runtime call failed";
    case GSMI_ERR_WORKSPACE: return "workspace too small";
    case GSMI_ERR_COUNT: return "pair count differs from the size of gaussian_ids";
    default: return "unknown status";
  }
}

extern "C" int gsmi_version(void) { return 100; }

extern "C" int gsmi_culling_gaussian_bsphere(int64_t N, const float* mean, const float* qvec, const float* svec,
                                             const float* normals, const float* pts, uint8_t* mask, float thresh,
                                             void* stream) {
  (void)qvec;
  if (N < 0) return GSMI_ERR_ARG;
  if (N == 0) return GSMI_OK;
  if (!mean || !svec || !normals || !pts || !mask) return GSMI_ERR_ARG;
  hipLaunchKernelGGL(k_cull, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream, N, mean, svec,
                     normals, pts, mask, thresh);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_exp_spec(int64_t n, const float* x, float* y, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y))) return GSMI_ERR_ARG;
  if (n == 0) return GSMI_OK;
  hipLaunchKernelGGL(k_exp_spec, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, x, y);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

// Workspace of gsmi_tile_culling_aabb_count: counts [V] + scan partials + total.
extern "C" size_t gsmi_aabb_count_workspace_size(int64_t V) {
  if (V < 0) V = 0;
  return align_up(V * 4) + align_up(scan_workspace_bytes(V, 4)) + 256;
}

__global__ void k_total_to_i64(const int32_t* __restrict__ in, int64_t* __restrict__ out) { *out = *in; }

extern "C" int gsmi_tile_culling_aabb_count(int64_t V, const float* mean2d, const float* cov2d, float fx, float fy,
                                            float cx, float cy, int W, int H, float D, int tile_size, int32_t* tl,
                                            int32_t* br, int64_t* d_pair_count, void* workspace,
                                            size_t workspace_bytes, void* stream) {
  if (V < 0 || W <= 0 || H <= 0 || tile_size <= 0 || !d_pair_count) return GSMI_ERR_ARG;
  if (V > 0 && (!mean2d || !cov2d || !tl || !br)) return GSMI_ERR_ARG;
  if (V >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  if (!workspace || workspace_bytes < gsmi_aabb_count_workspace_size(V)) return GSMI_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* p = reinterpret_cast<char*>(workspace);
  int32_t* cnt = reinterpret_cast<int32_t*>(p);
  p += align_up(V * 4);
  void* scan_ws = p;
  p += align_up(scan_workspace_bytes(V, 4));
  int32_t* total = reinterpret_cast<int32_t*>(p);
  const CamArgs c = make_cam(nullptr, fx, fy, cx, cy, W, H, D, 0.0f, 0, tile_size);
  if (V > 0) {
    hipLaunchKernelGGL(k_aabb, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, st, V, mean2d, cov2d, c, tl, br,
                       cnt);
    GSMI_CHECK_LAUNCH();
  }
  int rc = exclusive_scan<int32_t>(cnt, cnt, V, total, scan_ws, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_total_to_i64, dim3(1), dim3(1), 0, st, total, d_pair_count);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

// Fused preprocessing workspace: Int2 counts [N] + scan partials + total.
extern "C" size_t gsmi_preprocess_workspace_size(int64_t N) {
  if (N < 0) N = 0;
  return align_up(N * 8) + align_up(scan_workspace_bytes(N, 8)) + 256;
}

extern "C" int gsmi_preprocess_count(int64_t N, const float* mean, const float* qvec, const float* log_s,
                                     const float* c2w_host, const float* normals_host, const float* pts_host,
                                     float fx, float fy, float cx, float cy, int W, int H, float D,
                                     float frustum_radius, int skip_culling, int tile_size, uint8_t* vis,
                                     int32_t* vis_off, int32_t* pair_off, int64_t* totals, void* workspace,
                                     size_t workspace_bytes, void* stream) {
  if (N < 0 || W <= 0 || H <= 0 || !c2w_host || !totals) return GSMI_ERR_ARG;
  if (!skip_culling && (!normals_host || !pts_host)) return GSMI_ERR_ARG;
  if (N > 0 && (!mean || !qvec || !log_s || !vis || !vis_off || !pair_off)) return GSMI_ERR_ARG;
  if (tile_size != kTile) return GSMI_ERR_UNSUPPORTED;
  if (N >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  if (((uintptr_t)qvec & 15) != 0) return GSMI_ERR_ARG;
  if (!workspace || workspace_bytes < gsmi_preprocess_workspace_size(N)) return GSMI_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* p = reinterpret_cast<char*>(workspace);
  Int2* cnt = reinterpret_cast<Int2*>(p);
  p += align_up(N * 8);
  void* scan_ws = p;
  p += align_up(scan_workspace_bytes(N, 8));
  Int2* total = reinterpret_cast<Int2*>(p);
  const CamArgs c = make_cam(c2w_host, fx, fy, cx, cy, W, H, D, frustum_radius, skip_culling, tile_size);
  Planes pl{};
  if (!skip_culling)
    for (int k = 0; k < 18; ++k) {
      pl.n[k] = normals_host[k];
      pl.p[k] = pts_host[k];
    }
  if (N > 0) {
    hipLaunchKernelGGL(k_pre_count_planes, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, N, mean, qvec,
                       log_s, pl, c, vis, cnt);
    GSMI_CHECK_LAUNCH();
  }
  int rc = exclusive_scan<Int2>(cnt, cnt, N, total, scan_ws, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_split_int2, dim3((unsigned)((N + 255) / 256 + 1)), dim3(256), 0, st, cnt, N, vis_off,
                     pair_off, total, totals);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_preprocess_emit(int64_t N, const float* mean, const float* qvec, const float* log_s,
                                    const float* logit, const float* c2w_host, float fx, float fy, float cx,
                                    float cy, int W, int H, float D, int tile_size, const uint8_t* vis,
                                    const int32_t* vis_off, float* mean2d, float* cov2d, float* depth, float* alpha,
                                    int32_t* tl, int32_t* br, int32_t* orig, void* stream) {
  if (N < 0 || W <= 0 || H <= 0 || !c2w_host) return GSMI_ERR_ARG;
  if (tile_size != kTile) return GSMI_ERR_UNSUPPORTED;
  if (N == 0) return GSMI_OK;
  if (!mean || !qvec || !log_s || !logit || !vis || !vis_off) return GSMI_ERR_ARG;
  if (((uintptr_t)qvec & 15) != 0 || ((uintptr_t)cov2d & 15) != 0 || ((uintptr_t)mean2d & 7) != 0 ||
      ((uintptr_t)tl & 7) != 0 || ((uintptr_t)br & 7) != 0)
    return GSMI_ERR_ARG;
  const CamArgs c = make_cam(c2w_host, fx, fy, cx, cy, W, H, D, 0.0f, 0, tile_size);
  hipLaunchKernelGGL(k_pre_emit, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream, N, mean,
                     qvec, log_s, logit, c, vis, vis_off, mean2d, cov2d, depth, alpha, tl, br, orig);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}
