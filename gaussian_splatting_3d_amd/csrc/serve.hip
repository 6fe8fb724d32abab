// Continuous-batching render server on one MI355X (host code).
//
// The reference serves viewers by calling SHRenderer.forward once per client
// in turn (utils/viewer/viser_viewer.py:118-158, vis.py:18-32).  Here scenes
// stay resident in HBM (caller-owned device tensors) and camera requests are
// queued; a dispatcher thread packs up to max_views pending requests of one
// scene into a batch and runs the multi-view pipeline on one of n_slots HIP
// streams, so a new batch is counted, binned and composited while earlier
// batches are still on the GPU.  Every batch draws its buffers from one
// device arena allocated at start-up (no allocation per request):
//   phase A (before the pair count is known): view records, [B,N] visibility
//     and offsets, count workspace;
//   phase B (after one readback of the per-view bases): compacted records,
//     tile lists, binning workspace, the flat image.
// Admission: after the count, the batch keeps the longest prefix of its views
// whose phase-B bytes fit the arena (and whose pair count fits int32); the
// other views go back to the front of the queue.  A request that cannot fit
// even alone while nothing else holds memory completes with
// GSMI_ERR_WORKSPACE.  Each view's image is exactly what the single-view
// pipeline returns for that camera: the batch runs the same per-view
// arithmetic, only the launches are shared.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gsmi.h"

namespace gsmi_serve {

using Clock = std::chrono::steady_clock;

static size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// First-fit allocator over one device allocation (offsets, 256-byte aligned).
class Arena {
 public:
  void init(char* base, size_t size) {
    base_ = base;
    size_ = size;
    free_.clear();
    if (size) free_[0] = size;
  }
  bool alloc(size_t n, size_t* off) {
    n = align_up(std::max<size_t>(n, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n;
        const size_t at = it->first + n;
        free_.erase(it);
        if (rest) free_[at] = rest;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = align_up(std::max<size_t>(n, 1));
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t largest() const {
    size_t m = 0;
    for (auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }
  size_t size() const { return size_; }
  char* ptr(size_t off) const { return base_ + off; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0;
  std::map<size_t, size_t> free_;
};

struct Scene {
  int64_t N;
  int C;
  int64_t sh_stride_c;  // coefficients per channel in the sh tensor (C_max^2)
  const float *mean, *qvec, *log_s, *logit, *sh;
};

struct Request {
  int64_t id;
  int scene;
  gsmi_view view;  // tile_base / pix_base set when batched
  float* out;      // caller's device buffer [H*W*3]
  double t_submit;
};

struct Done {
  int64_t id;
  double t_submit, t_done;
  int status;
  int batch_views;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  gsmi_view* h_views = nullptr;  // pinned staging
  int64_t* h_bases = nullptr;    // pinned readback
  std::vector<Request> reqs;
  size_t offA = 0, bytesA = 0, offB = 0, bytesB = 0;
  bool busy = false;
  bool has_work = false;
  std::thread waiter;
};

// Pairs are binned per block of 2^kBlockShift x 2^kBlockShift tiles; each
// tile filters its block's list while compositing (the same list, so the
// same image as per-tile binning, with a sixth of the pairs to sort at 1080p).
constexpr int kBlockShift = 2;

// Bytes of phase A / phase B for a batch (layout carved in run_batch); P
// counts (record, block) pairs, n_blocks the blocks of all views.
static size_t phase_a_bytes(int64_t N, int B) {
  return align_up((size_t)(2 * B) * sizeof(gsmi_view)) + align_up((size_t)(B * N)) +
         align_up((size_t)(B * N) * 4) + align_up((size_t)(2 * (B + 1)) * 8) +
         align_up(gsmi_preprocess_views_workspace_size(N, B));
}

static size_t phase_b_bytes(int64_t V, int64_t P, int64_t n_blocks, int64_t n_pix) {
  return align_up(V * 8) + align_up(V * 16) + align_up(V * 4) * 2 + align_up(V * 8) * 4 + align_up(V * 4) +
         align_up(P * 4) + align_up(n_blocks * 4) * 2 + align_up(gsmi_binning_workspace_size(V, P, (int)n_blocks)) +
         align_up(n_pix * 12);
}

static int blocks_of(int n_tiles) { return (n_tiles + (1 << kBlockShift) - 1) >> kBlockShift; }

}  // namespace gsmi_serve

using namespace gsmi_serve;

struct gsmi_server {
  int device = 0;
  int max_views = 32;
  float D = 6.0f, frustum_radius = 1.0f, thresh = 1e-4f;
  int exact = 0;
  void* arena_mem = nullptr;
  Arena arena;
  std::vector<Scene> scenes;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request> pending;
  std::deque<Done> done;
  std::thread dispatcher;
  bool stop = false;
  int64_t next_id = 0;
  int in_flight = 0;
  Clock::time_point t0;
  int64_t stat_batches = 0, stat_views = 0, stat_max_batch = 0, stat_splits = 0, stat_failed = 0;

  double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

namespace gsmi_serve {

// Batches launched and not yet completed (their memory will be released).
static bool any_launched(gsmi_server* s) {
  for (auto& sl : s->slots)
    if (sl.has_work) return true;
  return false;
}

static void fail_requests(gsmi_server* s, const std::vector<Request>& reqs, int status) {
  const double t = s->now_ms();
  std::lock_guard<std::mutex> g(s->mu);
  for (auto& r : reqs) s->done.push_back(Done{r.id, r.t_submit, t, status, 0});
  s->stat_failed += (int64_t)reqs.size();
  s->cv.notify_all();
}

// Allocates n bytes from the arena, waiting for in-flight batches to release
// memory; false when it cannot fit even with nothing else in flight.
static bool arena_take(gsmi_server* s, size_t n, size_t* off) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    if (s->arena.alloc(n, off)) return true;
    if (!any_launched(s)) return false;
    s->cv.wait(lk);
  }
}

static void arena_give(gsmi_server* s, size_t off, size_t n) {
  std::lock_guard<std::mutex> g(s->mu);
  s->arena.release(off, n);
  s->cv.notify_all();
}

// Runs one batch on slot `sl` up to the launch of its last kernel; returns
// false when the requests were failed or re-queued instead.
static bool run_batch(gsmi_server* s, Slot& sl, std::vector<Request> reqs) {
  const Scene sc = s->scenes[reqs[0].scene];
  const int64_t N = sc.N;
  int B = (int)reqs.size();
  // tile / block / pixel numbering of the batch; h_views[B + b] is view b
  // with its block grid
  const int B0 = B;
  int64_t tile_base = 0, block_base = 0, pix_base = 0;
  for (int b = 0; b < B; ++b) {
    gsmi_view& v = reqs[b].view;
    v.tile_base = (int32_t)tile_base;
    v.pix_base = pix_base;
    tile_base += (int64_t)v.n_tiles_w * v.n_tiles_h;
    pix_base += (int64_t)v.W * v.H;
    sl.h_views[b] = v;
    gsmi_view bv = v;
    bv.n_tiles_w = blocks_of(v.n_tiles_w);
    bv.n_tiles_h = blocks_of(v.n_tiles_h);
    bv.tile_base = (int32_t)block_base;
    block_base += (int64_t)bv.n_tiles_w * bv.n_tiles_h;
    sl.h_views[B + b] = bv;
  }
  // ---- phase A: count
  const size_t bytesA = phase_a_bytes(N, B);
  size_t offA;
  if (!arena_take(s, bytesA, &offA)) {
    fail_requests(s, reqs, GSMI_ERR_WORKSPACE);
    return false;
  }
  char* p = s->arena.ptr(offA);
  gsmi_view* views_d = reinterpret_cast<gsmi_view*>(p);
  gsmi_view* bviews_d = views_d + B0;
  p += align_up((size_t)(2 * B) * sizeof(gsmi_view));
  uint8_t* vis = reinterpret_cast<uint8_t*>(p);
  p += align_up((size_t)(B * N));
  int32_t* vis_off = reinterpret_cast<int32_t*>(p);
  p += align_up((size_t)(B * N) * 4);
  int64_t* bases_d = reinterpret_cast<int64_t*>(p);
  p += align_up((size_t)(2 * (B + 1)) * 8);
  void* cws = p;
  const size_t cws_bytes = gsmi_preprocess_views_workspace_size(N, B);
  hipStream_t st = sl.stream;
  int rc = hipMemcpyAsync(views_d, sl.h_views, (size_t)(2 * B) * sizeof(gsmi_view), hipMemcpyHostToDevice, st) ==
                   hipSuccess
               ? GSMI_OK
               : GSMI_ERR_LAUNCH;
  if (!rc)
    rc = gsmi_preprocess_views_count_blocks(N, B, views_d, sc.mean, sc.qvec, sc.log_s, s->D, s->frustum_radius, 0,
                                            kBlockShift, vis, vis_off, bases_d, cws, cws_bytes, st);
  if (!rc && hipMemcpyAsync(sl.h_bases, bases_d, (size_t)(2 * (B + 1)) * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
    rc = GSMI_ERR_LAUNCH;
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = GSMI_ERR_LAUNCH;
  if (rc) {
    arena_give(s, offA, bytesA);
    fail_requests(s, reqs, rc);
    return false;
  }
  // ---- admission: longest prefix whose phase-B bytes fit
  size_t largest;
  {
    std::lock_guard<std::mutex> g(s->mu);
    largest = s->arena.largest();
  }
  int k = B;
  auto need = [&](int kk) {
    const int64_t V = sl.h_bases[2 * kk], P = sl.h_bases[2 * kk + 1];
    int64_t nt = 0, np = 0;
    for (int b = 0; b < kk; ++b) {
      nt += (int64_t)sl.h_views[B0 + b].n_tiles_w * sl.h_views[B0 + b].n_tiles_h;
      np += (int64_t)reqs[b].view.W * reqs[b].view.H;
    }
    return std::make_pair(P < (int64_t(1) << 31), phase_b_bytes(V, P, nt, np));
  };
  while (k > 1) {
    auto nb = need(k);
    if (nb.first && nb.second <= largest) break;
    --k;
  }
  if (k < B) {
    std::lock_guard<std::mutex> g(s->mu);
    for (int b = B - 1; b >= k; --b) s->pending.push_front(reqs[b]);
    reqs.resize(k);
    ++s->stat_splits;
    s->cv.notify_all();
  }
  B = k;
  auto nb = need(B);
  if (!nb.first) {
    arena_give(s, offA, bytesA);
    fail_requests(s, reqs, GSMI_ERR_UNSUPPORTED);
    return false;
  }
  const int64_t V = sl.h_bases[2 * B], P = sl.h_bases[2 * B + 1];
  int64_t n_tiles = 0, n_blocks = 0, n_pix = 0;
  for (int b = 0; b < B; ++b) {
    n_tiles += (int64_t)reqs[b].view.n_tiles_w * reqs[b].view.n_tiles_h;
    n_blocks += (int64_t)sl.h_views[B0 + b].n_tiles_w * sl.h_views[B0 + b].n_tiles_h;
    n_pix += (int64_t)reqs[b].view.W * reqs[b].view.H;
  }
  const size_t bytesB = nb.second;
  size_t offB;
  if (!arena_take(s, bytesB, &offB)) {
    arena_give(s, offA, bytesA);
    fail_requests(s, reqs, GSMI_ERR_WORKSPACE);
    return false;
  }
  // ---- phase B: emit, bin, composite, copy out
  p = s->arena.ptr(offB);
  auto take = [&p](size_t n) {
    char* r = p;
    p += align_up(n);
    return r;
  };
  float* mean2d = reinterpret_cast<float*>(take(V * 8));
  float* cov2d = reinterpret_cast<float*>(take(V * 16));
  float* depth = reinterpret_cast<float*>(take(V * 4));
  float* alpha = reinterpret_cast<float*>(take(V * 4));
  int32_t* tl = reinterpret_cast<int32_t*>(take(V * 8));
  int32_t* br = reinterpret_cast<int32_t*>(take(V * 8));
  int32_t* tl_b = reinterpret_cast<int32_t*>(take(V * 8));
  int32_t* br_b = reinterpret_cast<int32_t*>(take(V * 8));
  int32_t* orig = reinterpret_cast<int32_t*>(take(V * 4));
  int32_t* ids = reinterpret_cast<int32_t*>(take(P * 4));
  int32_t* start = reinterpret_cast<int32_t*>(take(n_blocks * 4));
  int32_t* end = reinterpret_cast<int32_t*>(take(n_blocks * 4));
  const size_t bws_bytes = gsmi_binning_workspace_size(V, P, (int)n_blocks);
  void* bws = take(bws_bytes);
  float* out = reinterpret_cast<float*>(take(n_pix * 12));
  rc = gsmi_preprocess_views_emit(N, B, views_d, sc.mean, sc.qvec, sc.log_s, sc.logit, s->D, vis, vis_off, mean2d,
                                  cov2d, depth, alpha, tl, br, orig, st);
  if (!rc) rc = gsmi_shift_rects(V, tl, br, kBlockShift, tl_b, br_b, st);
  if (!rc)
    rc = gsmi_bin_views(V, P, B, bviews_d, bases_d, (int)n_blocks, tl_b, br_b, depth, ids, start, end, bws,
                        bws_bytes, st);
  if (!rc)
    rc = gsmi_render_views_sh_two_level(V, B, views_d, bviews_d, (int)n_tiles, mean2d, cov2d, sc.sh,
                                        3 * sc.sh_stride_c, sc.sh_stride_c, orig, alpha, tl, br, start, end, ids,
                                        kBlockShift, out, sc.C, s->thresh, nullptr, nullptr, s->exact, st);
  for (int b = 0; !rc && b < B; ++b) {
    const gsmi_view& v = reqs[b].view;
    if (hipMemcpyAsync(reqs[b].out, out + 3 * v.pix_base, (size_t)v.W * v.H * 12, hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      rc = GSMI_ERR_LAUNCH;
  }
  if (!rc && hipEventRecord(sl.done, st) != hipSuccess) rc = GSMI_ERR_LAUNCH;
  if (rc) {
    (void)hipStreamSynchronize(st);
    arena_give(s, offB, bytesB);
    arena_give(s, offA, bytesA);
    fail_requests(s, reqs, rc);
    return false;
  }
  std::lock_guard<std::mutex> g(s->mu);
  sl.reqs = std::move(reqs);
  sl.offA = offA;
  sl.bytesA = bytesA;
  sl.offB = offB;
  sl.bytesB = bytesB;
  sl.has_work = true;
  ++s->stat_batches;
  s->stat_views += B;
  s->stat_max_batch = std::max<int64_t>(s->stat_max_batch, B);
  s->cv.notify_all();
  return true;
}

static void waiter_loop(gsmi_server* s, int idx) {
  (void)hipSetDevice(s->device);
  Slot& sl = s->slots[idx];
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return sl.has_work || (s->stop && !sl.busy); });
      if (!sl.has_work) return;
    }
    const int ok = hipEventSynchronize(sl.done) == hipSuccess ? GSMI_OK : GSMI_ERR_LAUNCH;
    const double t = s->now_ms();
    std::lock_guard<std::mutex> g(s->mu);
    for (auto& r : sl.reqs) s->done.push_back(Done{r.id, r.t_submit, t, ok, (int)sl.reqs.size()});
    s->arena.release(sl.offB, sl.bytesB);
    s->arena.release(sl.offA, sl.bytesA);
    sl.reqs.clear();
    sl.has_work = false;
    sl.busy = false;
    --s->in_flight;
    s->cv.notify_all();
  }
}

static void dispatcher_loop(gsmi_server* s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    int idx = -1;
    std::vector<Request> batch;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] {
        if (s->stop) return true;
        if (s->pending.empty()) return false;
        for (auto& sl : s->slots)
          if (!sl.busy) return true;
        return false;
      });
      if (s->stop) return;
      for (size_t i = 0; i < s->slots.size(); ++i)
        if (!s->slots[i].busy) {
          idx = (int)i;
          break;
        }
      // oldest request's scene; then the other pending requests of that
      // scene in arrival order
      const int scene = s->pending.front().scene;
      for (auto it = s->pending.begin(); it != s->pending.end() && (int)batch.size() < s->max_views;) {
        if (it->scene == scene) {
          batch.push_back(*it);
          it = s->pending.erase(it);
        } else {
          ++it;
        }
      }
      s->slots[idx].busy = true;
      ++s->in_flight;
    }
    if (!run_batch(s, s->slots[idx], std::move(batch))) {
      std::lock_guard<std::mutex> g(s->mu);
      s->slots[idx].busy = false;
      --s->in_flight;
      s->cv.notify_all();
    }
  }
}

}  // namespace gsmi_serve

// ---------------------------------------------------------------------------
// Camera record (host): the same float32/float64 steps as the Python
// CameraInfo (frustum planes of utils/camera.py:249-283, image constants of
// gs/sh_renderer.py:238-248), so records built here and there are identical.
// ---------------------------------------------------------------------------
static void cross3(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

extern "C" int gsmi_make_view(const float* c2w, double fx, double fy, double cx, double cy, int W, int H,
                              double near_plane, double far_plane, int tile_size, gsmi_view* out) {
  if (!c2w || !out || W <= 0 || H <= 0 || tile_size <= 0 || !(fx != 0.0) || !(fy != 0.0)) return GSMI_ERR_ARG;
  gsmi_view v;
  std::memset(&v, 0, sizeof(v));
  for (int k = 0; k < 12; ++k) v.c2w[k] = c2w[k];
  float right[3], up[3], look[3], t[3];
  for (int r = 0; r < 3; ++r) {
    right[r] = c2w[4 * r + 0];
    up[r] = -c2w[4 * r + 1];
    look[r] = c2w[4 * r + 2];
    t[r] = c2w[4 * r + 3];
  }
  const double yfov = 2.0 * std::atan((double)H / (2.0 * fy));
  const double aspect = (double)W / (double)H;
  const double half_v = far_plane * std::tan(yfov * 0.5);
  const double half_h = half_v * aspect;
  const float hv = (float)half_v, hh = (float)half_h;
  const float nearf = (float)near_plane, farf = (float)far_plane;
  float np_[3], fp[3], a[3], b[3];
  for (int r = 0; r < 3; ++r) {
    np_[r] = nearf * look[r];
    fp[r] = farf * look[r];
  }
  float n[6][3];
  for (int r = 0; r < 3; ++r) {
    n[0][r] = look[r];
    n[1][r] = -look[r];
  }
  for (int r = 0; r < 3; ++r) a[r] = fp[r] - hh * right[r];
  cross3(a, up, n[2]);
  for (int r = 0; r < 3; ++r) a[r] = fp[r] + hh * right[r];
  cross3(up, a, n[3]);
  for (int r = 0; r < 3; ++r) a[r] = fp[r] + hv * up[r];
  cross3(a, right, n[4]);
  for (int r = 0; r < 3; ++r) b[r] = fp[r] - hv * up[r];
  cross3(right, b, n[5]);
  for (int k = 0; k < 6; ++k) {
    const float len = std::sqrt(n[k][0] * n[k][0] + n[k][1] * n[k][1] + n[k][2] * n[k][2]);
    const float d = std::max(len, 1e-12f);
    for (int r = 0; r < 3; ++r) v.normals[3 * k + r] = n[k][r] / d;
  }
  for (int r = 0; r < 3; ++r) {
    v.points[r] = np_[r] + t[r];
    v.points[3 + r] = fp[r] + t[r];
    for (int k = 2; k < 6; ++k) v.points[3 * k + r] = t[r];
  }
  v.fx = (float)fx;
  v.fy = (float)fy;
  v.cx = (float)cx;
  v.cy = (float)cy;
  v.topleft_x = (float)(-cx / fx);
  v.topleft_y = (float)(-cy / fy);
  v.pixel_size_x = (float)(1.0 / fx);
  v.pixel_size_y = (float)(1.0 / fy);
  v.W = W;
  v.H = H;
  v.n_tiles_w = W / tile_size + (W % tile_size > 0);
  v.n_tiles_h = H / tile_size + (H % tile_size > 0);
  *out = v;
  return GSMI_OK;
}

// ---------------------------------------------------------------------------
// Server C ABI
// ---------------------------------------------------------------------------
extern "C" int gsmi_server_create(int device, int64_t arena_bytes, int max_views, int n_slots, float D,
                                  float frustum_radius, float thresh, int exact, gsmi_server** out) {
  if (!out || arena_bytes <= 0 || max_views < 1 || max_views > 256 || n_slots < 1 || n_slots > 8)
    return GSMI_ERR_ARG;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return GSMI_ERR_LAUNCH;
  gsmi_server* s = new gsmi_server();
  s->device = device;
  s->max_views = max_views;
  s->D = D;
  s->frustum_radius = frustum_radius;
  s->thresh = thresh;
  s->exact = exact;
  if (hipMalloc(&s->arena_mem, (size_t)arena_bytes) != hipSuccess) {
    delete s;
    return GSMI_ERR_WORKSPACE;
  }
  s->arena.init(reinterpret_cast<char*>(s->arena_mem), (size_t)arena_bytes);
  s->slots.resize(n_slots);
  for (auto& sl : s->slots) {
    bool ok = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess &&
              hipHostMalloc((void**)&sl.h_views, (size_t)(2 * max_views) * sizeof(gsmi_view), hipHostMallocDefault) ==
                  hipSuccess &&
              hipHostMalloc((void**)&sl.h_bases, (size_t)(2 * (max_views + 1)) * 8, hipHostMallocDefault) ==
                  hipSuccess;
    if (!ok) return GSMI_ERR_LAUNCH;  // resources of a failed create are not reclaimed
  }
  s->t0 = Clock::now();
  for (int i = 0; i < n_slots; ++i) s->slots[i].waiter = std::thread(waiter_loop, s, i);
  s->dispatcher = std::thread(dispatcher_loop, s);
  *out = s;
  return GSMI_OK;
}

extern "C" int gsmi_server_add_scene(gsmi_server* s, int64_t N, int C, int64_t sh_stride_c, const float* mean,
                                     const float* qvec, const float* log_s, const float* logit, const float* sh,
                                     int* scene_id) {
  if (!s || !scene_id || N < 0 || C < 1 || C > 4 || sh_stride_c < C * C) return GSMI_ERR_ARG;
  if (N > 0 && (!mean || !qvec || !log_s || !logit || !sh)) return GSMI_ERR_ARG;
  if ((uintptr_t)qvec & 15) return GSMI_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  s->scenes.push_back(Scene{N, C, sh_stride_c, mean, qvec, log_s, logit, sh});
  *scene_id = (int)s->scenes.size() - 1;
  return GSMI_OK;
}

extern "C" int gsmi_server_submit(gsmi_server* s, int scene_id, const float* c2w, double fx, double fy, double cx,
                                  double cy, int W, int H, double near_plane, double far_plane, float* out,
                                  int64_t* request_id) {
  if (!s || !out || !request_id) return GSMI_ERR_ARG;
  Request r;
  int rc = gsmi_make_view(c2w, fx, fy, cx, cy, W, H, near_plane, far_plane, 16, &r.view);
  if (rc) return rc;
  r.out = out;
  std::lock_guard<std::mutex> g(s->mu);
  if (scene_id < 0 || scene_id >= (int)s->scenes.size()) return GSMI_ERR_ARG;
  r.scene = scene_id;
  r.id = s->next_id++;
  r.t_submit = s->now_ms();
  s->pending.push_back(r);
  *request_id = r.id;
  s->cv.notify_all();
  return GSMI_OK;
}

// Completed requests in completion order: ids, submit / done times (ms since
// the server started), status, views in the batch that served it.
extern "C" int gsmi_server_poll(gsmi_server* s, int64_t* ids, double* t_submit, double* t_done, int32_t* status,
                                int32_t* batch_views, int max, int* n_out) {
  if (!s || !n_out || max < 0) return GSMI_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  int n = 0;
  while (n < max && !s->done.empty()) {
    const Done& d = s->done.front();
    if (ids) ids[n] = d.id;
    if (t_submit) t_submit[n] = d.t_submit;
    if (t_done) t_done[n] = d.t_done;
    if (status) status[n] = d.status;
    if (batch_views) batch_views[n] = d.batch_views;
    s->done.pop_front();
    ++n;
  }
  *n_out = n;
  return GSMI_OK;
}

// Blocks until no request is pending or in flight (or timeout_ms passes;
// returns GSMI_ERR_LAUNCH then).
extern "C" int gsmi_server_drain(gsmi_server* s, double timeout_ms) {
  if (!s) return GSMI_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  const bool ok = s->cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms),
                                 [&] { return s->pending.empty() && s->in_flight == 0; });
  return ok ? GSMI_OK : GSMI_ERR_LAUNCH;
}

// stats[0..5] = batches, views served, largest batch, admission splits,
// failed requests, arena bytes.
extern "C" int gsmi_server_stats(gsmi_server* s, int64_t* stats) {
  if (!s || !stats) return GSMI_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  stats[0] = s->stat_batches;
  stats[1] = s->stat_views;
  stats[2] = s->stat_max_batch;
  stats[3] = s->stat_splits;
  stats[4] = s->stat_failed;
  stats[5] = (int64_t)s->arena.size();
  return GSMI_OK;
}

extern "C" int gsmi_server_destroy(gsmi_server* s) {
  if (!s) return GSMI_ERR_ARG;
  gsmi_server_drain(s, 600000.0);
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->dispatcher.joinable()) s->dispatcher.join();
  for (auto& sl : s->slots)
    if (sl.waiter.joinable()) sl.waiter.join();
  for (auto& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.h_views) (void)hipHostFree(sl.h_views);
    if (sl.h_bases) (void)hipHostFree(sl.h_bases);
  }
  if (s->arena_mem) (void)hipFree(s->arena_mem);
  delete s;
  return GSMI_OK;
}
