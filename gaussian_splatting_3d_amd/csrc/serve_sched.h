// Host-side scheduling of the render server: the device arena's allocator,
// the choice of a batch from the request queue and the admission of a
// counted batch.  Plain C++ (no HIP): the server in serve.hip uses these, and
// the host-only test hooks there expose them to CPU tests.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <map>
#include <vector>

namespace gsmi_serve {

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// First-fit allocator over one device allocation (offsets, 256-byte aligned
// sizes); freed ranges are merged with free neighbours.
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

// Batch choice over a queue in arrival order (Q: a sequence container of
// requests; scene_of(r) and pixels_of(r) read a request; same_layout(a, b)
// compares two scenes' SH layouts).  The oldest request is always taken;
// then, in arrival order, each request whose scene shares the oldest one's
// layout while the batch stays within max_views views and max_pixels pixels
// (a request over the pixel budget is skipped, later smaller ones may still
// join).  Taken requests leave the queue, the rest keep their order.  The
// batch is sorted by scene, stably, so each scene's views are contiguous.
template <typename Q, typename SceneOf, typename PixelsOf, typename SameLayout>
std::vector<typename Q::value_type> select_batch(Q& pending, int max_views, int64_t max_pixels, SceneOf scene_of,
                                                 PixelsOf pixels_of, SameLayout same_layout) {
  std::vector<typename Q::value_type> batch;
  if (pending.empty()) return batch;
  const int lead = scene_of(pending.front());
  int64_t pixels = 0;
  for (auto it = pending.begin(); it != pending.end() && (int)batch.size() < max_views;) {
    const bool fits = batch.empty() || pixels + pixels_of(*it) <= max_pixels;
    if (fits && same_layout(scene_of(*it), lead)) {
      pixels += pixels_of(*it);
      batch.push_back(*it);
      it = pending.erase(it);
    } else {
      ++it;
    }
  }
  std::stable_sort(batch.begin(), batch.end(),
                   [&](const typename Q::value_type& a, const typename Q::value_type& b) {
                     return scene_of(a) < scene_of(b);
                   });
  return batch;
}

// Device bytes of a batch's second phase (after the pair counts are known)
// for V records, P (block, Gaussian) pairs, n_blocks blocks and B views;
// binning_ws is the binning workspace for (V, P, n_blocks).
inline size_t phase_b_bytes(int64_t V, int64_t P, int64_t n_blocks, int B, size_t binning_ws) {
  return align_up((size_t)(2 * (B + 1)) * 8) + align_up((size_t)(2 * B) * sizeof(void*)) + align_up(8 * 4) +
         align_up(V * 8) + align_up(V * 16) + align_up(V * 4) * 2 + align_up(V * 8) * 4 + align_up(V * 4) +
         align_up(P * 4) + align_up(n_blocks * 4) * 2 + align_up(binning_ws);
}

// Admission of a counted batch of B0 views (records vV[b], pairs vP[b],
// blocks nb[b]): the longest prefix, at least one view, whose phase-B bytes
// fit `largest` and whose record and pair totals fit int32.  bytes_of(V, P,
// n_blocks, B) gives the phase-B bytes.  Returns the prefix length; *fits_int
// says whether its totals fit int32 and *bytes is its phase-B size.  A
// single view that does not fit `largest` is returned as it is (the caller
// then waits for memory or fails it).
template <typename BytesOf>
int admit_prefix(int B0, const int64_t* vV, const int64_t* vP, const int64_t* nb, size_t largest, BytesOf bytes_of,
                 bool* fits_int, size_t* bytes) {
  auto need = [&](int k, bool* fi) {
    int64_t V = 0, P = 0, n = 0;
    for (int b = 0; b < k; ++b) {
      V += vV[b];
      P += vP[b];
      n += nb[b];
    }
    *fi = P < (int64_t(1) << 31) && V < (int64_t(1) << 31);
    return bytes_of(V, P, n, k);
  };
  int B = B0;
  *bytes = need(B, fits_int);
  while (B > 1 && !(*fits_int && *bytes <= largest)) *bytes = need(--B, fits_int);
  return B;
}

}  // namespace gsmi_serve
