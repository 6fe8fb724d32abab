// Per-Gaussian preprocessing on gfx950: frustum cull, projection, tile
// bounds, pair counts, compaction and key emission, in structure-of-arrays
// passes over the N raw Gaussians.  Every value that feeds a decision or a
// key follows the reference's operation order (gs/renderer.py:365-419,
// gs/culling.py:8-37, culling.h:11-20) with the specified exponential, so
// masks, tile bounds, depths and keys are bit-identical to the oracle's.
#include "project.h"
#include "scan.h"

namespace gsmi {

__global__ void k_cull(int64_t N, const float* __restrict__ mean, const float* __restrict__ svec,
                       const float* __restrict__ nrm, const float* __restrict__ pts, uint8_t* __restrict__ mask,
                       float thresh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float r = fmaxf(fmaxf(svec[3 * i], svec[3 * i + 1]), svec[3 * i + 2]) * thresh;
  mask[i] = in_frustum(mean[3 * i], mean[3 * i + 1], mean[3 * i + 2], r, nrm, pts) ? 1 : 0;
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

template <bool NONPOS>
__global__ void k_exp_spec(int64_t n, const float* __restrict__ x, float* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = x[i];
  y[i] = NONPOS && !(v > 0.0f) ? exp_spec_nonpos(v) : exp_spec(v);
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

static int launch_exp_spec(bool nonpos, int64_t n, const float* x, float* y, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y))) return GSMI_ERR_ARG;
  if (n == 0) return GSMI_OK;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (nonpos)
    hipLaunchKernelGGL(k_exp_spec<true>, grid, dim3(256), 0, (hipStream_t)stream, n, x, y);
  else
    hipLaunchKernelGGL(k_exp_spec<false>, grid, dim3(256), 0, (hipStream_t)stream, n, x, y);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_exp_spec(int64_t n, const float* x, float* y, void* stream) {
  return launch_exp_spec(false, n, x, y, stream);
}

extern "C" int gsmi_exp_spec_nonpos(int64_t n, const float* x, float* y, void* stream) {
  return launch_exp_spec(true, n, x, y, stream);
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
  // 16-pixel tiles; 32 or 64 count the pairs of 2x2 / 4x4-tile blocks (the
  // floor division of the same pixel bounds), for two-level binning
  if (tile_size != kTile && tile_size != 2 * kTile && tile_size != 4 * kTile) return GSMI_ERR_UNSUPPORTED;
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

// ---------------------------------------------------------------------------
// Two-level binning: tile bounds -> block bounds (blocks of 2^shift x 2^shift
// tiles); the floor division of non-negative tile indices is a shift.
// ---------------------------------------------------------------------------
namespace gsmi {
__global__ void k_shift_rects(int64_t V, const int2* __restrict__ tl, const int2* __restrict__ br, int shift,
                              int2* __restrict__ tl_s, int2* __restrict__ br_s) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int2 a = tl[v], b = br[v];
  tl_s[v] = make_int2(a.x >> shift, a.y >> shift);
  br_s[v] = make_int2(b.x >> shift, b.y >> shift);
}
}  // namespace gsmi

extern "C" int gsmi_shift_rects(int64_t V, const int32_t* tl, const int32_t* br, int shift, int32_t* tl_s,
                                int32_t* br_s, void* stream) {
  if (V < 0 || shift < 0 || shift > 4) return GSMI_ERR_ARG;
  if (V == 0) return GSMI_OK;
  if (!tl || !br || !tl_s || !br_s) return GSMI_ERR_ARG;
  if (((uintptr_t)tl | (uintptr_t)br | (uintptr_t)tl_s | (uintptr_t)br_s) & 7) return GSMI_ERR_ARG;
  hipLaunchKernelGGL(k_shift_rects, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, (hipStream_t)stream, V,
                     reinterpret_cast<const int2*>(tl), reinterpret_cast<const int2*>(br), shift,
                     reinterpret_cast<int2*>(tl_s), reinterpret_cast<int2*>(br_s));
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

// ---------------------------------------------------------------------------
// Multi-view batches: B cameras (gsmi_view records in device memory) over the
// same N Gaussians.  Entry (b, i) of the [B, N] arrays is Gaussian i seen by
// view b; one exclusive scan over the flattened array numbers the visible
// entries view-major, so every view's records are contiguous and the whole
// batch goes through one binning and one compositing launch.  Per entry the
// arithmetic is the single-view kernels' (same functions, same order), so
// every view's masks, records and keys equal its single-view ones.
// ---------------------------------------------------------------------------
namespace gsmi {

__global__ __launch_bounds__(256) void k_mv_count(int64_t N, int B, const float* __restrict__ mean,
                                                  const float* __restrict__ qvec, const float* __restrict__ log_s,
                                                  const gsmi_view* __restrict__ views, float D,
                                                  float frustum_radius, int skip_culling, int shift,
                                                  uint8_t* __restrict__ vis, Int2* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float m0 = mean[3 * i], m1 = mean[3 * i + 1], m2 = mean[3 * i + 2];
  const float s0 = exp_spec(log_s[3 * i]), s1 = exp_spec(log_s[3 * i + 1]), s2 = exp_spec(log_s[3 * i + 2]);
  const float4 q = reinterpret_cast<const float4*>(qvec)[i];
  const float r = fmaxf(fmaxf(s0, s1), s2) * frustum_radius;
  for (int b = 0; b < B; ++b) {
    const gsmi_view& v = views[b];
    const bool keep = skip_culling || in_frustum(m0, m1, m2, r, v.normals, v.points);
    int pairs = 0;
    if (keep) {
      const Proj p = project_one(m0, m1, m2, q, s0, s1, s2, v.c2w);
      int lx, hx, ly, hy;
      aabb_axis(p.mx, p.c[0], D, v.fx, v.cx, v.W - 1, kTile << shift, &lx, &hx);
      aabb_axis(p.my, p.c[3], D, v.fy, v.cy, v.H - 1, kTile << shift, &ly, &hy);
      pairs = (hx - lx + 1) * (hy - ly + 1);
    }
    const int64_t e = (int64_t)b * N + i;
    vis[e] = keep ? 1 : 0;
    cnt[e] = Int2{keep ? 1 : 0, pairs};
  }
}

// Visible offsets (x of the scanned counts) and per-view bases:
// bases[2b] / bases[2b+1] = first visible entry / first pair of view b,
// bases[2B] / bases[2B+1] = totals.
__global__ void k_mv_offsets(const Int2* __restrict__ cnt, int64_t N, int B, const Int2* __restrict__ total,
                             int32_t* __restrict__ vis_off, int64_t* __restrict__ bases) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (int64_t)B * N) vis_off[e] = cnt[e].x;
  if (e <= B) {
    const Int2 o = e < B ? (N > 0 ? cnt[e * N] : *total) : *total;
    bases[2 * e] = o.x;
    bases[2 * e + 1] = o.y;
  }
}

__global__ __launch_bounds__(256) void k_mv_emit(int64_t N, int B, const float* __restrict__ mean,
                                                 const float* __restrict__ qvec, const float* __restrict__ log_s,
                                                 const float* __restrict__ logit,
                                                 const gsmi_view* __restrict__ views, float D,
                                                 const uint8_t* __restrict__ vis, const int32_t* __restrict__ voff,
                                                 float* __restrict__ mean2d, float* __restrict__ cov2d,
                                                 float* __restrict__ depth, float* __restrict__ alpha,
                                                 int32_t* __restrict__ tl, int32_t* __restrict__ br,
                                                 int32_t* __restrict__ orig) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  bool any = false;
  for (int b = 0; b < B; ++b) any = any || vis[(int64_t)b * N + i];
  if (!any) return;
  const float m0 = mean[3 * i], m1 = mean[3 * i + 1], m2 = mean[3 * i + 2];
  const float s0 = exp_spec(log_s[3 * i]), s1 = exp_spec(log_s[3 * i + 1]), s2 = exp_spec(log_s[3 * i + 2]);
  const float4 q = reinterpret_cast<const float4*>(qvec)[i];
  const float al = 1.0f / (1.0f + exp_spec(-logit[i]));
  for (int b = 0; b < B; ++b) {
    const int64_t e = (int64_t)b * N + i;
    if (!vis[e]) continue;
    const gsmi_view& v = views[b];
    const int64_t o = voff[e];
    const Proj p = project_one(m0, m1, m2, q, s0, s1, s2, v.c2w);
    int lx, hx, ly, hy;
    aabb_axis(p.mx, p.c[0], D, v.fx, v.cx, v.W - 1, kTile, &lx, &hx);
    aabb_axis(p.my, p.c[3], D, v.fy, v.cy, v.H - 1, kTile, &ly, &hy);
    reinterpret_cast<float2*>(mean2d)[o] = make_float2(p.mx, p.my);
    reinterpret_cast<float4*>(cov2d)[o] = make_float4(p.c[0], p.c[1], p.c[2], p.c[3]);
    depth[o] = p.depth;
    alpha[o] = al;
    reinterpret_cast<int2*>(tl)[o] = make_int2(lx, ly);
    reinterpret_cast<int2*>(br)[o] = make_int2(hx, hy);
    orig[o] = (int32_t)i;
  }
}

}  // namespace gsmi

extern "C" size_t gsmi_preprocess_views_workspace_size(int64_t N, int B) {
  if (N < 0) N = 0;
  if (B < 0) B = 0;
  const int64_t n = N * B;
  return align_up(n * 8) + align_up(scan_workspace_bytes(n, 8)) + 256;
}

extern "C" int gsmi_preprocess_views_count_blocks(int64_t N, int B, const gsmi_view* views, const float* mean,
                                                  const float* qvec, const float* log_s, float D,
                                                  float frustum_radius, int skip_culling, int shift, uint8_t* vis,
                                                  int32_t* vis_off, int64_t* bases, void* workspace,
                                                  size_t workspace_bytes, void* stream) {
  if (N < 0 || B <= 0 || !views || !bases || shift < 0 || shift > 4) return GSMI_ERR_ARG;
  if (N > 0 && (!mean || !qvec || !log_s || !vis || !vis_off)) return GSMI_ERR_ARG;
  if (N * B >= (int64_t(1) << 31)) return GSMI_ERR_UNSUPPORTED;
  if (((uintptr_t)qvec & 15) != 0 || ((uintptr_t)views & 7) != 0) return GSMI_ERR_ARG;
  if (!workspace || workspace_bytes < gsmi_preprocess_views_workspace_size(N, B)) return GSMI_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = N * B;
  char* p = reinterpret_cast<char*>(workspace);
  Int2* cnt = reinterpret_cast<Int2*>(p);
  p += align_up(n * 8);
  void* scan_ws = p;
  p += align_up(scan_workspace_bytes(n, 8));
  Int2* total = reinterpret_cast<Int2*>(p);
  if (N > 0) {
    hipLaunchKernelGGL(k_mv_count, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, N, B, mean, qvec, log_s,
                       views, D, frustum_radius, skip_culling, shift, vis, cnt);
    GSMI_CHECK_LAUNCH();
  }
  int rc = exclusive_scan<Int2>(cnt, cnt, n, total, scan_ws, st);
  if (rc) return rc;
  const int64_t m = std::max<int64_t>(n, B + 1);
  hipLaunchKernelGGL(k_mv_offsets, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, cnt, N, B, total, vis_off,
                     bases);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}

extern "C" int gsmi_preprocess_views_count(int64_t N, int B, const gsmi_view* views, const float* mean,
                                           const float* qvec, const float* log_s, float D, float frustum_radius,
                                           int skip_culling, uint8_t* vis, int32_t* vis_off, int64_t* bases,
                                           void* workspace, size_t workspace_bytes, void* stream) {
  return gsmi_preprocess_views_count_blocks(N, B, views, mean, qvec, log_s, D, frustum_radius, skip_culling, 0, vis,
                                            vis_off, bases, workspace, workspace_bytes, stream);
}

extern "C" int gsmi_preprocess_views_emit(int64_t N, int B, const gsmi_view* views, const float* mean,
                                          const float* qvec, const float* log_s, const float* logit, float D,
                                          const uint8_t* vis, const int32_t* vis_off, float* mean2d, float* cov2d,
                                          float* depth, float* alpha, int32_t* tl, int32_t* br, int32_t* orig,
                                          void* stream) {
  if (N < 0 || B <= 0 || !views) return GSMI_ERR_ARG;
  if (N == 0) return GSMI_OK;
  if (!mean || !qvec || !log_s || !logit || !vis || !vis_off) return GSMI_ERR_ARG;
  if (((uintptr_t)qvec & 15) != 0 || ((uintptr_t)cov2d & 15) != 0 || ((uintptr_t)mean2d & 7) != 0 ||
      ((uintptr_t)tl & 7) != 0 || ((uintptr_t)br & 7) != 0 || ((uintptr_t)views & 7) != 0)
    return GSMI_ERR_ARG;
  hipLaunchKernelGGL(k_mv_emit, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream, N, B, mean,
                     qvec, log_s, logit, views, D, vis, vis_off, mean2d, cov2d, depth, alpha, tl, br, orig);
  GSMI_CHECK_LAUNCH();
  return GSMI_OK;
}
