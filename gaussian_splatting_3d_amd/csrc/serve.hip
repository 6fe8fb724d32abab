// Continuous-batching render server on one MI355X (host code).
//
// The reference serves viewers by calling SHRenderer.forward once per client
// in turn (utils/viewer/viser_viewer.py:118-158, vis.py:18-32).  Here scenes
// stay resident in HBM (caller-owned device tensors) and camera requests are
// queued.  n_slots worker threads each own a HIP stream; a free worker takes
// the oldest pending request and, in arrival order, the other pending
// requests whose scenes share its SH layout (scenes may mix), up to
// max_views views and max_batch_pixels pixels (the cap bounds how long one
// batch holds the GPU, which bounds the latency of the requests behind it).
// The batch runs the multi-view pipeline: the views are grouped by scene,
// each group is counted (cull, project, pairs per 64-px block) against its
// scene, one readback gives the pair counts, then each group's records are
// emitted into one record array, all views are binned together and one
// compositing launch over every tile of every view writes each image
// straight into its request's buffer (per-view SH rows and output pointers);
// that launch is a persistent kernel whose workgroups take tiles from
// per-XCD work queues.  Several workers keep several batches in flight on
// the GPU; no thread waits on a batch but the worker that owns it.
//
// Memory: one device arena allocated at start-up; a batch takes
//   phase A (before the pair counts are known): view records, per-group
//     [B_g, N_g] visibility and offsets, counts and count workspace;
//   phase B (after the readback): combined bases, per-view pointers, work
//     queues, compacted records, block lists, binning workspace.
// Admission: after the count the batch keeps the longest prefix of its
// views whose phase-B bytes fit the arena's largest free block (and whose
// pair count fits int32); the others go back to the front of the queue.  A
// request that cannot fit even alone while nothing else holds memory
// completes with GSMI_ERR_WORKSPACE.  Each image is exactly what the
// single-view pipeline returns for that camera: the batch runs the same
// per-view arithmetic, only the launches are shared.
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
#include "serve_internal.h"
#include "serve_sched.h"

namespace gsmi_serve {

using Clock = std::chrono::steady_clock;

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
  int64_t pixels;
};

struct Done {
  int64_t id;
  double t_submit, t_done;
  int status;
  int batch_views;
};

// Pairs are binned per block of 2^kBlockShift x 2^kBlockShift tiles; each
// tile filters its block's list while compositing (the same list, so the
// same image as per-tile binning, with a sixth of the pairs to sort at 1080p).
constexpr int kBlockShift = 2;

static int blocks_of(int n_tiles) { return (n_tiles + (1 << kBlockShift) - 1) >> kBlockShift; }

// A scene group of a batch: views [v0, v0 + nv) render scene `scene`.
struct Group {
  int scene, v0, nv;
  size_t vis_off, voff_off, bases_off, ws_off, ws_bytes;  // phase-A offsets
};

struct Worker {
  hipStream_t stream = nullptr;
  hipEvent_t counted = nullptr, done = nullptr;
  gsmi_view* h_views = nullptr;  // pinned: [B] views, then [B] block views
  int64_t* h_lbases = nullptr;   // pinned: per group 2 (nv + 1) local bases, groups back to back
  int64_t* h_bases = nullptr;    // pinned: 2 (B + 1) combined bases
  void** h_ptrs = nullptr;       // pinned: [B] SH pointers, then [B] output pointers
  std::thread th;
};

}  // namespace gsmi_serve

using namespace gsmi_serve;

struct gsmi_server {
  int device = 0;
  int max_views = 32;
  int64_t max_batch_pixels = int64_t(1) << 40;
  int queue_grid = 0;  // work-queue compositing: workgroups (0: one per tile)
  float D = 6.0f, frustum_radius = 1.0f, thresh = 1e-4f;
  int exact = 0;
  void* arena_mem = nullptr;
  Arena arena;
  std::vector<Scene> scenes;
  std::vector<Worker> workers;
  std::mutex mu;
  std::condition_variable cv;
  // Held by the one batch that is acquiring memory, from its phase-A
  // allocation to its phase-B allocation: every other batch in flight then
  // holds all its memory and will release it, so waiting cannot deadlock.
  std::mutex acquire_mu;
  std::deque<Request> pending;
  std::deque<Done> done;
  bool stop = false;
  int64_t next_id = 0;
  int in_flight = 0;  // batches taken by workers and not finished
  Clock::time_point t0;
  int64_t stat_batches = 0, stat_views = 0, stat_max_batch = 0, stat_splits = 0, stat_failed = 0, stat_mixed = 0;

  double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

namespace gsmi_serve {

static void fail_requests(gsmi_server* s, const std::vector<Request>& reqs, int status) {
  const double t = s->now_ms();
  std::lock_guard<std::mutex> g(s->mu);
  for (auto& r : reqs) s->done.push_back(Done{r.id, r.t_submit, t, status, 0});
  s->stat_failed += (int64_t)reqs.size();
  s->cv.notify_all();
}

// Allocates n bytes from the arena, waiting for other batches to release
// memory; false when it cannot fit even with no other batch in flight.
// Called with acquire_mu held.
static bool arena_take(gsmi_server* s, size_t n, size_t* off) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    if (s->arena.alloc(n, off)) return true;
    if (s->in_flight <= 1) return false;  // only this batch: waiting frees nothing
    s->cv.wait(lk);
  }
}

static void arena_give(gsmi_server* s, size_t off, size_t n) {
  std::lock_guard<std::mutex> g(s->mu);
  s->arena.release(off, n);
  s->cv.notify_all();
}

static bool same_layout(const Scene& a, const Scene& b) { return a.C == b.C && a.sh_stride_c == b.sh_stride_c; }

// Takes a batch from the queue (caller holds the lock; pending not empty);
// the policy is select_batch (serve_sched.h).
static std::vector<Request> take_batch(gsmi_server* s) {
  return select_batch(
      s->pending, s->max_views, s->max_batch_pixels, [](const Request& r) { return r.scene; },
      [](const Request& r) { return r.pixels; },
      [s](int a, int b) { return same_layout(s->scenes[a], s->scenes[b]); });
}

// Phase-B bytes of a batch prefix with its binning workspace.
static size_t batch_phase_b_bytes(int64_t V, int64_t P, int64_t n_blocks, int B) {
  return phase_b_bytes(V, P, n_blocks, B, gsmi_binning_workspace_size(V, P, (int)n_blocks));
}

// Runs one batch to completion on worker w (returns when its images are
// written or its requests failed / re-queued).
static void run_batch(gsmi_server* s, Worker& w, std::vector<Request> reqs) {
  hipStream_t st = w.stream;
  const int B0 = (int)reqs.size();
  // groups of consecutive views of one scene
  std::vector<Group> groups;
  for (int b = 0; b < B0; ++b)
    if (groups.empty() || groups.back().scene != reqs[b].scene)
      groups.push_back(Group{reqs[b].scene, b, 1, 0, 0, 0, 0, 0});
    else
      ++groups.back().nv;
  // tile / block numbering of the batch; h_views[B0 + b] is view b's block grid
  {
    int64_t tile_base = 0, block_base = 0, pix_base = 0;
    for (int b = 0; b < B0; ++b) {
      gsmi_view& v = reqs[b].view;
      v.tile_base = (int32_t)tile_base;
      v.pix_base = pix_base;
      tile_base += (int64_t)v.n_tiles_w * v.n_tiles_h;
      pix_base += (int64_t)v.W * v.H;
      w.h_views[b] = v;
      gsmi_view bv = v;
      bv.n_tiles_w = blocks_of(v.n_tiles_w);
      bv.n_tiles_h = blocks_of(v.n_tiles_h);
      bv.tile_base = (int32_t)block_base;
      block_base += (int64_t)bv.n_tiles_w * bv.n_tiles_h;
      w.h_views[B0 + b] = bv;
    }
  }
  // ---- phase A: count every group
  size_t bytesA = align_up((size_t)(2 * B0) * sizeof(gsmi_view));
  for (auto& g : groups) {
    const int64_t N = s->scenes[g.scene].N;
    g.vis_off = bytesA;
    bytesA += align_up((size_t)(g.nv * N));
    g.voff_off = bytesA;
    bytesA += align_up((size_t)(g.nv * N) * 4);
    g.bases_off = bytesA;
    bytesA += align_up((size_t)(2 * (g.nv + 1)) * 8);
    g.ws_off = bytesA;
    g.ws_bytes = gsmi_preprocess_views_workspace_size(N, g.nv);
    bytesA += align_up(g.ws_bytes);
  }
  std::unique_lock<std::mutex> acquiring(s->acquire_mu);
  size_t offA;
  if (!arena_take(s, bytesA, &offA)) {
    fail_requests(s, reqs, GSMI_ERR_WORKSPACE);
    return;
  }
  char* pa = s->arena.ptr(offA);
  gsmi_view* views_d = reinterpret_cast<gsmi_view*>(pa);
  gsmi_view* bviews_d = views_d + B0;
  int rc = hipMemcpyAsync(views_d, w.h_views, (size_t)(2 * B0) * sizeof(gsmi_view), hipMemcpyHostToDevice, st) ==
                   hipSuccess
               ? GSMI_OK
               : GSMI_ERR_LAUNCH;
  size_t lb = 0;  // offset of a group's local bases in h_lbases
  for (auto& g : groups) {
    const Scene& sc = s->scenes[g.scene];
    if (!rc)
      rc = gsmi_preprocess_views_count_blocks(sc.N, g.nv, views_d + g.v0, sc.mean, sc.qvec, sc.log_s, s->D,
                                              s->frustum_radius, 0, kBlockShift,
                                              reinterpret_cast<uint8_t*>(pa + g.vis_off),
                                              reinterpret_cast<int32_t*>(pa + g.voff_off),
                                              reinterpret_cast<int64_t*>(pa + g.bases_off), pa + g.ws_off,
                                              g.ws_bytes, st);
    if (!rc && hipMemcpyAsync(w.h_lbases + lb, pa + g.bases_off, (size_t)(2 * (g.nv + 1)) * 8,
                              hipMemcpyDeviceToHost, st) != hipSuccess)
      rc = GSMI_ERR_LAUNCH;
    lb += 2 * (g.nv + 1);
  }
  if (!rc && hipEventRecord(w.counted, st) != hipSuccess) rc = GSMI_ERR_LAUNCH;
  if (!rc && hipEventSynchronize(w.counted) != hipSuccess) rc = GSMI_ERR_LAUNCH;  // this worker only
  if (rc) {
    (void)hipStreamSynchronize(st);
    arena_give(s, offA, bytesA);
    fail_requests(s, reqs, rc);
    return;
  }
  // per-view (records, pairs) and the batch prefix sums
  std::vector<int64_t> vV(B0), vP(B0);
  lb = 0;
  for (auto& g : groups) {
    const int64_t* l = w.h_lbases + lb;
    for (int k = 0; k < g.nv; ++k) {
      vV[g.v0 + k] = l[2 * (k + 1)] - l[2 * k];
      vP[g.v0 + k] = l[2 * (k + 1) + 1] - l[2 * k + 1];
    }
    lb += 2 * (g.nv + 1);
  }
  // ---- admission: longest prefix whose phase-B bytes fit
  size_t largest;
  {
    std::lock_guard<std::mutex> g(s->mu);
    largest = s->arena.largest();
  }
  std::vector<int64_t> vNb(B0);
  for (int b = 0; b < B0; ++b) vNb[b] = (int64_t)w.h_views[B0 + b].n_tiles_w * w.h_views[B0 + b].n_tiles_h;
  bool fits_int = true;
  size_t bytesB = 0;
  const int B = admit_prefix(B0, vV.data(), vP.data(), vNb.data(), largest, batch_phase_b_bytes, &fits_int, &bytesB);
  if (B < B0) {
    std::lock_guard<std::mutex> g(s->mu);
    for (int b = B0 - 1; b >= B; --b) s->pending.push_front(reqs[b]);
    ++s->stat_splits;
    s->cv.notify_all();
  }
  reqs.resize(B);
  if (!fits_int) {
    arena_give(s, offA, bytesA);
    fail_requests(s, reqs, GSMI_ERR_UNSUPPORTED);
    return;
  }
  size_t offB;
  if (!arena_take(s, bytesB, &offB)) {
    arena_give(s, offA, bytesA);
    fail_requests(s, reqs, GSMI_ERR_WORKSPACE);
    return;
  }
  acquiring.unlock();
  // ---- phase B: emit per group, bin all views, composite into the outputs
  int64_t V = 0, P = 0, n_tiles = 0, n_blocks = 0;
  for (int b = 0; b < B; ++b) {
    w.h_bases[2 * b] = V;
    w.h_bases[2 * b + 1] = P;
    V += vV[b];
    P += vP[b];
    n_tiles += (int64_t)reqs[b].view.n_tiles_w * reqs[b].view.n_tiles_h;
    n_blocks += (int64_t)w.h_views[B0 + b].n_tiles_w * w.h_views[B0 + b].n_tiles_h;
    w.h_ptrs[b] = const_cast<float*>(s->scenes[reqs[b].scene].sh);
    w.h_ptrs[B + b] = reqs[b].out;
  }
  w.h_bases[2 * B] = V;
  w.h_bases[2 * B + 1] = P;
  char* p = s->arena.ptr(offB);
  auto take = [&p](size_t n) {
    char* r = p;
    p += align_up(n);
    return r;
  };
  int64_t* bases_d = reinterpret_cast<int64_t*>(take((size_t)(2 * (B + 1)) * 8));
  void** ptrs_d = reinterpret_cast<void**>(take((size_t)(2 * B) * sizeof(void*)));
  int* work_d = reinterpret_cast<int*>(take(8 * 4));
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
  rc = hipMemcpyAsync(bases_d, w.h_bases, (size_t)(2 * (B + 1)) * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
               hipMemcpyAsync(ptrs_d, w.h_ptrs, (size_t)(2 * B) * sizeof(void*), hipMemcpyHostToDevice, st) ==
                   hipSuccess
           ? GSMI_OK
           : GSMI_ERR_LAUNCH;
  for (auto& g : groups) {
    if (rc || g.v0 >= B) break;
    const Scene& sc = s->scenes[g.scene];
    const int nv = std::min(g.nv, B - g.v0);
    const int64_t v0 = w.h_bases[2 * g.v0];
    rc = gsmi_preprocess_views_emit(sc.N, nv, views_d + g.v0, sc.mean, sc.qvec, sc.log_s, sc.logit, s->D,
                                    reinterpret_cast<const uint8_t*>(pa + g.vis_off),
                                    reinterpret_cast<const int32_t*>(pa + g.voff_off), mean2d + 2 * v0,
                                    cov2d + 4 * v0, depth + v0, alpha + v0, tl + 2 * v0, br + 2 * v0, orig + v0, st);
  }
  if (!rc) rc = gsmi_shift_rects(V, tl, br, kBlockShift, tl_b, br_b, st);
  if (!rc)
    rc = gsmi_bin_views(V, P, B, bviews_d, bases_d, (int)n_blocks, tl_b, br_b, depth, ids, start, end, bws,
                        bws_bytes, st);
  if (!rc) {
    const Scene& sc = s->scenes[reqs[0].scene];
    rc = gsmi_internal_render_served(V, B, views_d, bviews_d, (int)n_tiles, mean2d, cov2d,
                                     reinterpret_cast<const float* const*>(ptrs_d), 3 * sc.sh_stride_c,
                                     sc.sh_stride_c, orig, alpha, tl, br, start, end, ids, kBlockShift,
                                     reinterpret_cast<float* const*>(ptrs_d + B), sc.C, s->thresh, s->exact, work_d,
                                     s->queue_grid, st);
  }
  if (!rc && hipEventRecord(w.done, st) != hipSuccess) rc = GSMI_ERR_LAUNCH;
  const int ok = !rc && hipEventSynchronize(w.done) == hipSuccess ? GSMI_OK : (rc ? rc : GSMI_ERR_LAUNCH);
  if (ok != GSMI_OK) (void)hipStreamSynchronize(st);
  const double t = s->now_ms();
  std::lock_guard<std::mutex> g(s->mu);
  for (auto& r : reqs) s->done.push_back(Done{r.id, r.t_submit, t, ok, B});
  s->arena.release(offB, bytesB);
  s->arena.release(offA, bytesA);
  if (ok == GSMI_OK) {
    s->stat_mixed += groups.size() > 1 && groups[1].v0 < B;
    ++s->stat_batches;
    s->stat_views += B;
    s->stat_max_batch = std::max<int64_t>(s->stat_max_batch, B);
  } else {
    s->stat_failed += B;
  }
  s->cv.notify_all();
}

static void worker_loop(gsmi_server* s, int idx) {
  (void)hipSetDevice(s->device);
  Worker& w = s->workers[idx];
  for (;;) {
    std::vector<Request> batch;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->pending.empty(); });
      if (s->pending.empty()) return;  // stopping with nothing queued
      batch = take_batch(s);
      ++s->in_flight;
    }
    run_batch(s, w, std::move(batch));
    std::lock_guard<std::mutex> g(s->mu);
    --s->in_flight;
    s->cv.notify_all();
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
  s->workers.resize(n_slots);
  for (auto& w : s->workers) {
    bool ok = hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&w.counted, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&w.done, hipEventDisableTiming) == hipSuccess &&
              hipHostMalloc((void**)&w.h_views, (size_t)(2 * max_views) * sizeof(gsmi_view), hipHostMallocDefault) ==
                  hipSuccess &&
              hipHostMalloc((void**)&w.h_lbases, (size_t)(4 * max_views + 2) * 8, hipHostMallocDefault) ==
                  hipSuccess &&
              hipHostMalloc((void**)&w.h_bases, (size_t)(2 * (max_views + 1)) * 8, hipHostMallocDefault) ==
                  hipSuccess &&
              hipHostMalloc((void**)&w.h_ptrs, (size_t)(2 * max_views) * sizeof(void*), hipHostMallocDefault) ==
                  hipSuccess;
    if (!ok) return GSMI_ERR_LAUNCH;  // resources of a failed create are not reclaimed
  }
  s->t0 = Clock::now();
  for (int i = 0; i < n_slots; ++i) s->workers[i].th = std::thread(worker_loop, s, i);
  *out = s;
  return GSMI_OK;
}

extern "C" int gsmi_server_configure(gsmi_server* s, int64_t max_batch_pixels, int queue_grid) {
  if (!s || max_batch_pixels <= 0 || queue_grid < 0 || queue_grid > (1 << 20)) return GSMI_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  s->max_batch_pixels = max_batch_pixels;
  s->queue_grid = queue_grid;
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
  r.pixels = (int64_t)W * H;
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

extern "C" int gsmi_server_stats_ex(gsmi_server* s, int64_t* stats, int n) {
  if (!s || !stats || n < 0 || n > 7) return GSMI_ERR_ARG;
  int64_t v[7];
  int rc = gsmi_server_stats(s, v);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> g(s->mu);
    v[6] = s->stat_mixed;
  }
  for (int k = 0; k < n; ++k) stats[k] = v[k];
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
  for (auto& w : s->workers)
    if (w.th.joinable()) w.th.join();
  for (auto& w : s->workers) {
    if (w.stream) (void)hipStreamSynchronize(w.stream);
    if (w.counted) (void)hipEventDestroy(w.counted);
    if (w.done) (void)hipEventDestroy(w.done);
    if (w.stream) (void)hipStreamDestroy(w.stream);
    if (w.h_views) (void)hipHostFree(w.h_views);
    if (w.h_lbases) (void)hipHostFree(w.h_lbases);
    if (w.h_bases) (void)hipHostFree(w.h_bases);
    if (w.h_ptrs) (void)hipHostFree(w.h_ptrs);
  }
  if (s->arena_mem) (void)hipFree(s->arena_mem);
  delete s;
  return GSMI_OK;
}

// ---------------------------------------------------------------------------
// Host-only test hooks of the scheduling policy (no device is touched).
// ---------------------------------------------------------------------------

extern "C" int gsmi_sched_arena_run(int64_t size, int n_ops, const int64_t* ops, int64_t* out) {
  if (size < 0 || n_ops < 0 || (n_ops > 0 && (!ops || !out))) return GSMI_ERR_ARG;
  Arena a;
  a.init(nullptr, (size_t)size);
  for (int i = 0; i < n_ops; ++i) {
    const int64_t op = ops[3 * i], x = ops[3 * i + 1], y = ops[3 * i + 2];
    if (op == 0) {
      size_t off = 0;
      out[i] = x >= 0 && a.alloc((size_t)x, &off) ? (int64_t)off : -1;
    } else if (op == 1) {
      if (x < 0 || y < 0) return GSMI_ERR_ARG;
      a.release((size_t)x, (size_t)y);
      out[i] = 0;
    } else if (op == 2) {
      out[i] = (int64_t)a.largest();
    } else {
      return GSMI_ERR_ARG;
    }
  }
  return GSMI_OK;
}

extern "C" int gsmi_sched_select_batch(int n, const int32_t* scene, const int64_t* pixels, int n_scenes,
                                       const int32_t* layout, int max_views, int64_t max_pixels, int32_t* batch,
                                       int32_t* n_batch, int32_t* rest) {
  if (n <= 0 || n_scenes <= 0 || max_views <= 0 || !scene || !pixels || !layout || !batch || !n_batch || !rest)
    return GSMI_ERR_ARG;
  std::deque<int> q;
  for (int i = 0; i < n; ++i) {
    if (scene[i] < 0 || scene[i] >= n_scenes || pixels[i] < 0) return GSMI_ERR_ARG;
    q.push_back(i);
  }
  const std::vector<int> b = select_batch(
      q, max_views, max_pixels, [&](int r) { return scene[r]; }, [&](int r) { return pixels[r]; },
      [&](int s0, int s1) { return layout[s0] == layout[s1]; });
  *n_batch = (int32_t)b.size();
  for (size_t k = 0; k < b.size(); ++k) batch[k] = b[k];
  int k = 0;
  for (int r : q) rest[k++] = r;
  return GSMI_OK;
}

extern "C" int gsmi_sched_admit(int B0, const int64_t* vV, const int64_t* vP, const int64_t* n_blocks,
                                uint64_t largest, int32_t* B, int32_t* fits_int, uint64_t* bytes) {
  if (B0 <= 0 || !vV || !vP || !n_blocks || !B || !fits_int || !bytes) return GSMI_ERR_ARG;
  bool fi = true;
  size_t by = 0;
  *B = admit_prefix(B0, vV, vP, n_blocks, (size_t)largest, batch_phase_b_bytes, &fi, &by);
  *fits_int = fi ? 1 : 0;
  *bytes = (uint64_t)by;
  return GSMI_OK;
}
