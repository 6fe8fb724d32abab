/*
 * gsmi.h — C ABI of the MI355X (gfx950) Gaussian-splatting rasteriser.
 *
 * Plain pointers and sizes only.  Every device pointer is a gfx950 device
 * address; `stream` is a hipStream_t passed as void* (NULL = the null
 * stream).  All calls are asynchronous on `stream` unless stated otherwise,
 * never allocate device memory, never exit the process, and return a status
 * code (GSMI_OK on success).  Outputs are caller-owned, as in the reference
 * (gs/renderer.py:695, 760-763; gs/sh_renderer.py:245-249).
 *
 * Float conventions shared with the kernels: binary32, no multiply-add
 * contraction where a value feeds a decision (culling, tile bounds, the
 * 1/255 skip, the transmittance break), and the exponential evaluated by the
 * specified algorithm documented in gsmi_exp_spec() below.
 */
#ifndef GSMI_H
#define GSMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GSMI_OK = 0,
  GSMI_ERR_ARG = 1,         /* null pointer, negative size, bad dimension */
  GSMI_ERR_UNSUPPORTED = 2, /* e.g. SH order C outside 1..4, tile size != 16 */
  GSMI_ERR_LAUNCH = 3,      /* a kernel launch or runtime call failed */
  GSMI_ERR_WORKSPACE = 4,   /* workspace smaller than the size query returned */
  GSMI_ERR_COUNT = 5        /* pair count differs from the one the caller sized */
};

/*
 * One camera of a multi-view batch (256 bytes; arrays of these live in device
 * memory).  Tiles of all views are numbered consecutively: view b owns global
 * tile ids [tile_base, tile_base + n_tiles_w * n_tiles_h); its pixels start
 * at pix_base in the flat [sum(H*W), 3] image.  topleft = (-cx/fx, -cy/fy),
 * pixel size = (1/fx, 1/fy) (gs/sh_renderer.py:239-248), frustum planes as
 * CameraInfo.get_frustum (utils/camera.py:249-283).
 */
typedef struct gsmi_view {
  float c2w[12]; /* row-major 3x4 camera-to-world */
  float normals[18];
  float points[18];
  float fx, fy, cx, cy;
  float topleft_x, topleft_y, pixel_size_x, pixel_size_y;
  int32_t W, H, n_tiles_w, n_tiles_h;
  int32_t tile_base;
  int32_t reserved;
  int64_t pix_base;
} gsmi_view;

/* Human-readable text for a status code (static storage). */
const char *gsmi_status_string(int status);

/* ABI version: major * 10000 + minor * 100 + patch. */
int gsmi_version(void);

/* ------------------------------------------------------------------------ */
/* Frustum culling                                                           */
/* ------------------------------------------------------------------------ */

/*
 * Replaces culling_gaussian_bsphere (gs/src/render.cu:15-43,
 * gs/src/include/culling.h:11-34; bound at gs/src/bindings.cpp:6).
 * mask[i] = 1 iff for all six planes k: dot(mean_i - pts_k, normals_k) >
 * -max(svec_i) * thresh.  mean, svec: [N,3]; qvec [N,4] is accepted and
 * unused, as in the reference; normals, pts: [6,3] device arrays.
 */
int gsmi_culling_gaussian_bsphere(int64_t N, const float *mean, const float *qvec, const float *svec,
                                  const float *normals, const float *pts, uint8_t *mask, float thresh,
                                  void *stream);

/* ------------------------------------------------------------------------ */
/* Tile bounds and binning                                                   */
/* ------------------------------------------------------------------------ */

/*
 * Replaces the PyTorch tile_culling_aabb_count (gs/culling.py:8-37 with
 * utils/camera.py:290-303).  For V projected Gaussians (mean2d [V,2], cov2d
 * [V,4] row-major 2x2) writes int32 tile bounds tl, br [V,2] and the pair
 * count P = sum (brx-tlx+1)(bry-tly+1) into *d_pair_count (device int64).
 * fx, fy, cx, cy, D are the camera/config scalars rounded to binary32.
 */
size_t gsmi_aabb_count_workspace_size(int64_t V);

int gsmi_tile_culling_aabb_count(int64_t V, const float *mean2d, const float *cov2d, float fx, float fy,
                                 float cx, float cy, int W, int H, float D, int tile_size, int32_t *tl,
                                 int32_t *br, int64_t *d_pair_count, void *workspace, size_t workspace_bytes,
                                 void *stream);

/* Workspace bytes needed by gsmi_tile_culling_aabb_start_end. */
size_t gsmi_binning_workspace_size(int64_t V, int64_t P, int n_tiles);

/*
 * Replaces tile_culling_aabb_start_end (gs/src/render.cu:380-397,
 * gs/src/include/aabb_culling.h:15-41, 70-103, 192-260; bound at
 * bindings.cpp:29).  Produces the order of a stable sort of the keys
 * (int64)(y*n_tiles_w + x) << 32 | float_bits(depth[v]) emitted per covered
 * tile in index order: gaussian_ids [P] (sorted values v), start/end
 * [n_tiles] (-1 for empty tiles) and, if sorted_keys != NULL, the sorted
 * 64-bit keys [P].  Computed depth-first: stable sort of the V Gaussians by
 * depth bits, pair expansion in that order, stable sort of the pairs by
 * tile.  flags bit 0: check that P equals the pair count of tl/br before any
 * pair is written (synchronises the stream; GSMI_ERR_COUNT on mismatch).
 * flags bits 8-10 (no counterpart in the reference): shift s <= 4; the
 * bounds are binned per block of 2^s x 2^s tiles (tl >> s, br >> s; P, the
 * grid n_tiles_h x n_tiles_w and start/end are then the blocks'), as the
 * two-level lists need.  Other bits must be zero.
 */
int gsmi_tile_culling_aabb_start_end(int64_t V, int64_t P, const int32_t *tl, const int32_t *br,
                                     const float *depth, int n_tiles_h, int n_tiles_w, int32_t *gaussian_ids,
                                     int32_t *start, int32_t *end, uint64_t *sorted_keys, void *workspace,
                                     size_t workspace_bytes, int flags, void *stream);

/* Workspace bytes for gsmi_sort_pairs (radix sort of P (uint64, int32) pairs). */
size_t gsmi_sort_workspace_size(int64_t P);

/*
 * Stable ascending sort of P (key, value) pairs on key bits [0, end_bit),
 * end_bit <= 64; the pairs are sorted in place (keys/values overwritten).
 * Replaces cub::DeviceRadixSort::SortPairs<int64_t,int>
 * (gs/src/include/aabb_culling.h:235-241).
 */
int gsmi_sort_pairs(int64_t P, uint64_t *keys, int32_t *values, int end_bit, void *workspace,
                    size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------------ */
/* Fused preprocessing (activation, cull, projection, tile bounds, keys)     */
/* ------------------------------------------------------------------------ */

/*
 * Replaces, as one coalesced pass over the N raw Gaussians, the chain
 * SHRenderer.forward (gs/sh_renderer.py:188-262): exp/sigmoid activations
 * (sh_renderer.py:318-324), culling_gaussian_bsphere, the masked gathers
 * (:205-209), project_gaussians (gs/renderer.py:365-419) and
 * tile_culling_aabb_count (gs/culling.py:8-37).
 *
 * Phase 1 (gsmi_preprocess_count): per Gaussian i, vis[i] = survives the
 * frustum cull; pairs[i] = its tile-pair count (0 when culled).  Then
 * exclusive prefix sums: vis_off[i], pair_off[i], and totals[0] = V,
 * totals[1] = P (device int64[2]).  The caller reads totals to size the
 * phase-2 outputs.
 * Camera: c2w_host [3,4] row-major and normals_host / pts_host [6,3]
 * (CameraInfo.get_frustum) are HOST arrays, copied into the kernel
 * arguments; fx, fy, cx, cy, W, H, D host scalars.  mean, qvec, log_s are
 * device arrays [N,3], [N,4] (16-byte aligned), [N,3].
 */
size_t gsmi_preprocess_workspace_size(int64_t N);

int gsmi_preprocess_count(int64_t N, const float *mean, const float *qvec, const float *log_s,
                          const float *c2w_host, const float *normals_host, const float *pts_host, float fx,
                          float fy, float cx, float cy, int W, int H, float D, float frustum_radius,
                          int skip_culling, int tile_size, uint8_t *vis, int32_t *vis_off, int32_t *pair_off,
                          int64_t *totals, void *workspace, size_t workspace_bytes, void *stream);

/*
 * Phase 2: for every visible i with v = vis_off[i], writes the compacted
 * mean2d[v] (2), cov2d[v] (4), depth[v], alpha[v] (sigmoid of logit),
 * tl[v], br[v] (2 each) and orig[v] = i.  Binning then runs on (tl, br,
 * depth) with gsmi_tile_culling_aabb_start_end.
 */
int gsmi_preprocess_emit(int64_t N, const float *mean, const float *qvec, const float *log_s, const float *logit,
                         const float *c2w_host, float fx, float fy, float cx, float cy, int W, int H, float D,
                         int tile_size, const uint8_t *vis, const int32_t *vis_off, float *mean2d, float *cov2d,
                         float *depth, float *alpha, int32_t *tl, int32_t *br, int32_t *orig, void *stream);

/*
 * Tile ranges from sorted keys: start/end [n_tiles], -1 when empty
 * (gs/src/include/aabb_culling.h:70-103, 248-249).
 */
int gsmi_tile_ranges(int64_t P, const uint64_t *sorted_keys, int n_tiles, int32_t *start, int32_t *end,
                     void *stream);

/* ------------------------------------------------------------------------ */
/* Compositing                                                               */
/* ------------------------------------------------------------------------ */

/*
 * The reference's compositing functions with the reference's arguments (V =
 * mean.size(0), P = gaussian_ids.size(0), then the tensors' device pointers
 * in the order of gs/src/render.h, then the stream).  sh_coeffs is the
 * contiguous [V, 3, C*C] slice (gs/sh_renderer.py:273), color [V,3], out the
 * flat [H*W*3] image (HWC) that the caller zeroed (gs/renderer.py:695), and
 * the grad_* arrays are zeroed or hold values to accumulate onto
 * (renderer.py:760-763).  grad_cov receives 0.5 * dL/dcov_offdiag in both
 * off-diagonals and grad_alpha the gradient w.r.t. the post-sigmoid alpha
 * straight through the 0.99 clamp, as the reference.  The backward entries
 * recompute the forward's transmittance inside the launch (the reference's
 * backward does the same, vol_render_sh.h:268-351); for SH without a
 * background and for RGB `out` is not read.  C outside 1..4 returns
 * GSMI_ERR_UNSUPPORTED (the reference's switch silently does nothing).
 *
 * tile_based_vol_rendering_sh: render.h:86-93, render.cu:483-544,
 *   vol_render_sh.h:97-262; bound at bindings.cpp:42.
 * tile_based_vol_rendering_backward_sh: render.h:95-101, render.cu:546-624,
 *   vol_render_sh.h:268-476; bindings.cpp:44.
 * tile_based_vol_rendering_sh_with_bg: render.h:119-124, render.cu:780-842,
 *   vol_render_bg.h:12-119 (bg_rgb: device float[3]).
 * tile_based_vol_rendering_backward_sh_with_bg: render.h:126-133,
 *   render.cu:844-925, vol_render_bg.h:121-258 (out: the blended output).
 * tile_based_vol_rendering_start_end: render.h:70-76, render.cu:399-441,
 *   vol_render.h:716-798; bindings.cpp:31 (Gaussian in double,
 *   kernels.h:195-224).
 * tile_based_vol_rendering_backward_start_end: render.h:78-84,
 *   render.cu:443-481, vol_render.h:800-925; bindings.cpp:35.
 */
int gsmi_tile_based_vol_rendering_sh(int64_t V, int64_t P, const float *mean, const float *cov,
                                     const float *sh_coeffs, const float *alpha, const int32_t *start,
                                     const int32_t *end, const int32_t *gaussian_ids, float *out,
                                     const float *topleft, const float *c2w, int tile_size, int n_tiles_h,
                                     int n_tiles_w, float pixel_size_x, float pixel_size_y, int H, int W, int C,
                                     float thresh, void *stream);
int gsmi_tile_based_vol_rendering_backward_sh(int64_t V, int64_t P, const float *mean, const float *cov,
                                              const float *sh_coeffs, const float *alpha, const int32_t *start,
                                              const int32_t *end, const int32_t *gaussian_ids, const float *out,
                                              float *grad_mean, float *grad_cov, float *grad_sh_coeffs,
                                              float *grad_alpha, const float *grad_out, const float *topleft,
                                              const float *c2w, int tile_size, int n_tiles_h, int n_tiles_w,
                                              float pixel_size_x, float pixel_size_y, int H, int W, int C,
                                              float thresh, void *stream);
int gsmi_tile_based_vol_rendering_sh_with_bg(int64_t V, int64_t P, const float *mean, const float *cov,
                                             const float *sh_coeffs, const float *alpha, const int32_t *start,
                                             const int32_t *end, const int32_t *gaussian_ids, float *out,
                                             const float *topleft, const float *c2w, int tile_size, int n_tiles_h,
                                             int n_tiles_w, float pixel_size_x, float pixel_size_y, int H, int W,
                                             int C, float thresh, const float *bg_rgb, void *stream);
int gsmi_tile_based_vol_rendering_backward_sh_with_bg(
    int64_t V, int64_t P, const float *mean, const float *cov, const float *sh_coeffs, const float *alpha,
    const int32_t *start, const int32_t *end, const int32_t *gaussian_ids, const float *out, float *grad_mean,
    float *grad_cov, float *grad_sh_coeffs, float *grad_alpha, const float *grad_out, const float *topleft,
    const float *c2w, int tile_size, int n_tiles_h, int n_tiles_w, float pixel_size_x, float pixel_size_y, int H,
    int W, int C, float thresh, const float *bg_rgb, void *stream);
int gsmi_tile_based_vol_rendering_start_end(int64_t V, int64_t P, const float *mean, const float *cov,
                                            const float *color, const float *alpha, const int32_t *start,
                                            const int32_t *end, const int32_t *gaussian_ids, float *out,
                                            const float *topleft, int tile_size, int n_tiles_h, int n_tiles_w,
                                            float pixel_size_x, float pixel_size_y, int H, int W, float thresh,
                                            void *stream);
int gsmi_tile_based_vol_rendering_backward_start_end(
    int64_t V, int64_t P, const float *mean, const float *cov, const float *color, const float *alpha,
    const int32_t *start, const int32_t *end, const int32_t *gaussian_ids, const float *out, float *grad_mean,
    float *grad_cov, float *grad_color, float *grad_alpha, const float *grad_out, const float *topleft,
    int tile_size, int n_tiles_h, int n_tiles_w, float pixel_size_x, float pixel_size_y, int H, int W,
    float thresh, void *stream);

/*
 * Extended SH forward (the entries above call it; this build's pipeline
 * calls it directly).  Same arithmetic as gsmi_tile_based_vol_rendering_sh
 * (bg == NULL) or _with_bg (bg = device float[3]), plus:
 * SH coefficient (g, ch, k) is read at
 * sh[sh_row(g) * sh_stride_g + ch * sh_stride_c + k], sh_row(g) =
 * sh_index ? sh_index[g] : g (the reference layout is sh_stride_g = 3*C*C,
 * sh_stride_c = C*C, sh_index = NULL), so the raw [N,3,C_max^2] tensor is
 * read through the compacted records' original index without a copy.
 * flags bit 0: also write pixels of empty tiles (zeros; the reference leaves
 * them untouched).  Optional per-pixel state for the backward pass (NULL to
 * skip): T_final [H*W], n_done [H*W] (list entries examined before the
 * transmittance break), pre_bg [H*W*3] (colour before the background blend;
 * bg mode only).  exact != 0 evaluates colours with the oracle's exact
 * operation sequence (bit-reproducible output); exact == 0 uses fused
 * multiply-adds for the SH dot product and the hardware exponential for the
 * colour sigmoid.  All compositing decisions are exact in both modes.
 */
int gsmi_render_sh_forward(int64_t V, int64_t P, const float *mean, const float *cov, const float *sh,
                           int64_t sh_stride_g, int64_t sh_stride_c, const int32_t *sh_index, const float *alpha,
                           const int32_t *start, const int32_t *end, const int32_t *gaussian_ids, float *out,
                           const float *topleft, const float *c2w, int tile_size, int n_tiles_h, int n_tiles_w,
                           float pixel_size_x, float pixel_size_y, int H, int W, int C, float thresh,
                           const float *bg, float *T_final, int32_t *n_done, float *pre_bg, int flags, int exact,
                           void *stream);

/* flags bit of the SH backward entries: reduce the SH gradients (orders 3
 * and 4) on bf16 matrix cores as a three-term split (products with relative
 * error up to 2^-15, biased toward zero) instead of fp32.  Off by default;
 * not fp32 arithmetic. */
#define GSMI_BWD_SPLIT_BF16 1
/* flags bit of the SH backward entries (saved state, not with
 * GSMI_BWD_SPLIT_BF16): for SH orders 3 and 4 the colours (SH dot products)
 * of the entries are formed on fp32 matrix cores, five list entries at a
 * time, instead of per entry on the vector units; fp32 products, summed in
 * another order.  Orders 1 and 2 run on the vector units either way.  Off by
 * default. */
#define GSMI_BWD_COLOUR_MFMA 2

/*
 * Extended SH backward (reverse-order adjoint).  With the forward's T_final
 * and n_done it recomputes the transmittance backwards from them; with
 * T_final = n_done = NULL it recomputes the forward state front to back
 * inside the launch first (the reference-shaped entries above).  Per-Gaussian
 * gradients are reduced within each wavefront, then accumulated (float
 * atomics) into grad_mean [V,2], grad_cov [V,4], grad_sh (same addressing as
 * sh) and grad_alpha [V].  out is the forward's output (read with bg only:
 * the blended colour) and pre_bg the colour before blending (bg with saved
 * state only).  flags: 0, GSMI_BWD_SPLIT_BF16 or GSMI_BWD_COLOUR_MFMA (saved
 * state only).
 * workspace (saved state only; NULL for none) of
 * gsmi_backward_order_workspace_size(n_tiles_h * n_tiles_w) bytes: the tiles
 * are launched most expensive first (cost from n_done), which shortens the
 * launch's tail; the gradients are the same.  No counterpart in the
 * reference beyond the entries above.
 */
size_t gsmi_backward_order_workspace_size(int n_tiles_total);
int gsmi_render_sh_backward(int64_t V, int64_t P, const float *mean, const float *cov, const float *sh,
                            int64_t sh_stride_g, int64_t sh_stride_c, const int32_t *sh_index, const float *alpha,
                            const int32_t *start, const int32_t *end, const int32_t *gaussian_ids, const float *out,
                            const float *T_final, const int32_t *n_done, const float *pre_bg,
                            const float *grad_out, float *grad_mean, float *grad_cov, float *grad_sh,
                            float *grad_alpha, const float *topleft, const float *c2w, int tile_size,
                            int n_tiles_h, int n_tiles_w, float pixel_size_x, float pixel_size_y, int H, int W,
                            int C, float thresh, const float *bg, int flags, void *workspace,
                            size_t workspace_bytes, void *stream);

/*
 * gsmi_render_sh_backward with the launch-order workspace taken from a
 * file-scope device array instead of an argument: allocates nothing (no
 * hipMalloc, no allocating library call).  No counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid allocation).  Requires the saved state
 * (T_final and n_done; GSMI_ERR_ARG otherwise) and at most 262,144 tiles
 * (n_tiles_h * n_tiles_w; GSMI_ERR_WORKSPACE beyond).  The array (2 MiB + 256 bytes per
 * GPU, held until the process ends) is shared by every call on a GPU: calls
 * that overlap on one GPU race.
 */
int gsmi_render_sh_backward_noalloc(int64_t V, int64_t P, const float *mean, const float *cov, const float *sh,
                                    int64_t sh_stride_g, int64_t sh_stride_c, const int32_t *sh_index,
                                    const float *alpha, const int32_t *start, const int32_t *end,
                                    const int32_t *gaussian_ids, const float *out, const float *T_final,
                                    const int32_t *n_done, const float *pre_bg, const float *grad_out,
                                    float *grad_mean, float *grad_cov, float *grad_sh, float *grad_alpha,
                                    const float *topleft, const float *c2w, int tile_size, int n_tiles_h,
                                    int n_tiles_w, float pixel_size_x, float pixel_size_y, int H, int W, int C,
                                    float thresh, const float *bg, int flags, void *stream);

/*
 * Extended RGB compositing (render_start_end), called by the reference-shaped
 * entries above: the forward with optional T_final / n_done outputs and flags
 * bit 0 (write empty tiles' pixels); the backward from saved T_final / n_done
 * or, with both NULL, recomputing them inside the launch.
 */
int gsmi_render_rgb_forward(int64_t V, int64_t P, const float *mean, const float *cov, const float *color,
                            const float *alpha, const int32_t *start, const int32_t *end,
                            const int32_t *gaussian_ids, float *out, const float *topleft, int tile_size,
                            int n_tiles_h, int n_tiles_w, float pixel_size_x, float pixel_size_y, int H, int W,
                            float thresh, float *T_final, int32_t *n_done, int flags, void *stream);
int gsmi_render_rgb_backward(int64_t V, int64_t P, const float *mean, const float *cov, const float *color,
                             const float *alpha, const int32_t *start, const int32_t *end,
                             const int32_t *gaussian_ids, const float *out, const float *T_final,
                             const int32_t *n_done, const float *grad_out, float *grad_mean, float *grad_cov,
                             float *grad_color, float *grad_alpha, const float *topleft, int tile_size,
                             int n_tiles_h, int n_tiles_w, float pixel_size_x, float pixel_size_y, int H, int W,
                             float thresh, void *stream);

/*
 * Two-level lists (no counterpart in the reference; single view, no
 * background).  The pairs are binned per block of 2^shift x 2^shift tiles
 * (gsmi_preprocess_count with tile_size 16 << shift gives their count, then
 * gsmi_tile_culling_aabb_start_end on the block grid with the shift in its
 * flags, or over bounds shifted by gsmi_shift_rects),
 * and the forward filters each tile's list from its block's depth-ordered
 * list on the fly: the filtered list equals the tile's own list, so images,
 * T_final and n_done are those of gsmi_render_sh_forward.  The
 * forward records each tile's filtered entries (up to where its pixels
 * saturated) in fids[foff[t] .. fend[t]), so the backward takes (foff, fend,
 * fids) as its start / end / gaussian_ids.  foff: exclusive prefix of the
 * tiles' block-list lengths (gsmi_two_level_offsets); fids holds their sum,
 * at most (2^shift)^2 x the block pair count.
 */
int gsmi_shift_rects(int64_t V, const int32_t *tl, const int32_t *br, int shift, int32_t *tl_s, int32_t *br_s,
                     void *stream);
size_t gsmi_two_level_offsets_workspace_size(int n_tiles);
int gsmi_two_level_offsets(int n_tiles_h, int n_tiles_w, int shift, const int32_t *sstart, const int32_t *send,
                           int32_t *foff, void *workspace, size_t workspace_bytes, void *stream);
int gsmi_render_sh_two_level(int64_t V, const float *mean, const float *cov, const float *sh, int64_t sh_stride_g,
                             int64_t sh_stride_c, const int32_t *sh_index, const float *alpha, const int32_t *tl,
                             const int32_t *br, const int32_t *sstart, const int32_t *send, const int32_t *sids,
                             int shift, const int32_t *foff, int32_t *fids, int32_t *fend, float *out,
                             const float *topleft, const float *c2w, int n_tiles_h, int n_tiles_w,
                             float pixel_size_x, float pixel_size_y, int H, int W, int C, float thresh,
                             float *T_final, int32_t *n_done, int exact, void *stream);

/* ------------------------------------------------------------------------ */
/* Projection adjoint                                                        */
/* ------------------------------------------------------------------------ */

/*
 * Replaces the torch autograd backward through project_gaussians and the
 * activations (gs/renderer.py:390-419, utils/transforms.py:31-45,
 * utils/activations.py:13-19, sh_renderer.py:205-209 masked gathers):
 * keeps the reference's detaches (Jacobian and depth are constants, the
 * mean gradient flows only through xy / depth).  For i < N: culled
 * Gaussians (vis[i] == 0) get zero gradients; visible ones read the
 * compositing gradients at v = vis_off[i] (grad_mean2d [V,2], grad_cov [V,4],
 * grad_alpha [V]) and write grad_mean [N,3], grad_qvec [N,4], grad_log_s
 * [N,3], grad_logit [N] (overwrite, not accumulate).  c2w_host is a host
 * array [3,4].
 */
int gsmi_project_backward(int64_t N, const float *mean, const float *qvec, const float *log_s, const float *logit,
                          const float *c2w_host, const uint8_t *vis, const int32_t *vis_off, const float *grad_mean2d,
                          const float *grad_cov, const float *grad_alpha, float *grad_mean, float *grad_qvec,
                          float *grad_log_s, float *grad_logit, void *stream);

/* ------------------------------------------------------------------------ */
/* Multi-view batches                                                        */
/* ------------------------------------------------------------------------ */
/*
 * No single counterpart in the reference, which renders one camera per call
 * (SHRenderer.forward, gs/sh_renderer.py:188-316; the viewer loops over
 * clients, utils/viewer/viser_viewer.py:118-158).  These run that same
 * pipeline for B cameras in one pass per stage: `views` is a device array of
 * B gsmi_view records; per-entry arrays are [B, N] (entry b * N + i =
 * Gaussian i in view b); visible records are numbered view-major.  Each
 * view's masks, records, tile lists and image are bit-identical to the
 * single-view calls on that camera.
 */
size_t gsmi_preprocess_views_workspace_size(int64_t N, int B);

/* Visibility and pair counts of every (view, Gaussian); vis [B*N],
 * vis_off [B*N] (visible record index, valid where vis), bases (device,
 * int64 [2*(B+1)]): bases[2b], bases[2b+1] = first record / first pair of
 * view b; bases[2B], bases[2B+1] = totals. */
int gsmi_preprocess_views_count(int64_t N, int B, const gsmi_view *views, const float *mean, const float *qvec,
                                const float *log_s, float D, float frustum_radius, int skip_culling, uint8_t *vis,
                                int32_t *vis_off, int64_t *bases, void *workspace, size_t workspace_bytes,
                                void *stream);

/* As gsmi_preprocess_views_count with the pairs counted per block of
 * 2^shift x 2^shift tiles (shift 0: the call above).  No counterpart in the
 * reference: the count of the two-level binning of a batch. */
int gsmi_preprocess_views_count_blocks(int64_t N, int B, const gsmi_view *views, const float *mean,
                                       const float *qvec, const float *log_s, float D, float frustum_radius,
                                       int skip_culling, int shift, uint8_t *vis, int32_t *vis_off, int64_t *bases,
                                       void *workspace, size_t workspace_bytes, void *stream);

/* Compacted records of every visible (view, Gaussian), as
 * gsmi_preprocess_emit; tl/br are in the view's own tile coordinates. */
int gsmi_preprocess_views_emit(int64_t N, int B, const gsmi_view *views, const float *mean, const float *qvec,
                               const float *log_s, const float *logit, float D, const uint8_t *vis,
                               const int32_t *vis_off, float *mean2d, float *cov2d, float *depth, float *alpha,
                               int32_t *tl, int32_t *br, int32_t *orig, void *stream);

/* Depth-first binning of all views' records into one tile numbering (view
 * b's tile (x, y) is tile_base + y * n_tiles_w + x); start/end [n_tiles_total].
 * Workspace: gsmi_binning_workspace_size(V, P, n_tiles_total). */
int gsmi_bin_views(int64_t V, int64_t P, int B, const gsmi_view *views, const int64_t *bases, int n_tiles_total,
                   const int32_t *tl, const int32_t *br, const float *depth, int32_t *gaussian_ids,
                   int32_t *start, int32_t *end, void *workspace, size_t workspace_bytes, void *stream);

/* SH compositing of every tile of every view in one launch; out is the flat
 * [sum(H*W), 3] image (view b at pix_base), T_final / n_done per pixel. */
int gsmi_render_views_sh(int64_t V, int64_t P, int B, const gsmi_view *views, int n_tiles_total,
                         const float *mean, const float *cov, const float *sh, int64_t sh_stride_g,
                         int64_t sh_stride_c, const int32_t *sh_index, const float *alpha, const int32_t *start,
                         const int32_t *end, const int32_t *gaussian_ids, float *out, int C, float thresh,
                         const float *bg, float *T_final, int32_t *n_done, float *pre_bg, int flags, int exact,
                         void *stream);

/* Two-level form of gsmi_render_views_sh (no counterpart in the reference;
 * forward only, no background): block_views[b] is view b with its block grid
 * (n_tiles_w/h = ceil(tiles / 2^shift), tile_base numbering the blocks of
 * all views); bstart/bend/bids are gsmi_bin_views' lists over block_views
 * (records' tl/br shifted by gsmi_shift_rects), tl/br the records' tile
 * bounds.  Each tile filters its block's list as it composites, so the
 * images (and T_final / n_done when given) equal gsmi_render_views_sh's. */
int gsmi_render_views_sh_two_level(int64_t V, int B, const gsmi_view *views, const gsmi_view *block_views,
                                   int n_tiles_total, const float *mean, const float *cov, const float *sh,
                                   int64_t sh_stride_g, int64_t sh_stride_c, const int32_t *sh_index,
                                   const float *alpha, const int32_t *tl, const int32_t *br, const int32_t *bstart,
                                   const int32_t *bend, const int32_t *bids, int shift, float *out, int C,
                                   float thresh, float *T_final, int32_t *n_done, int exact, void *stream);

/* Two-level lists of a batch kept per view (no counterpart in the
 * reference; the batched training path): a 16-px tile's filtered list is
 * recorded in its view's own array, so that no offset exceeds the int32
 * range however many views the batch holds.
 * gsmi_two_level_offsets_views: foff[t] for every tile t of every view (the
 * views' tiles in order, view_tiles [B] host array of tile counts) = the
 * exclusive prefix, within the tile's view, of the lengths of the block
 * lists the tiles fall in (block_views / bstart / bend as in
 * gsmi_render_views_sh_two_level); so view b's array needs at most
 * 4^shift x (block pairs of view b) entries.  Workspace:
 * gsmi_two_level_offsets_workspace_size(max view_tiles).
 * gsmi_render_views_sh_two_level_record: gsmi_render_views_sh_two_level that
 * also writes each tile's filtered list (up to where its pixels saturate) to
 * view_fids[b] (device array of B pointers) from foff[t] on and its end to
 * fend[t]; T_final / n_done as gsmi_render_views_sh.
 * gsmi_render_views_backward_sh_lists: gsmi_render_views_backward_sh over
 * those lists (start = foff, end = fend, entries at view_ids[b]), saved
 * state required, no background. */
int gsmi_two_level_offsets_views(int B, const gsmi_view *views, const gsmi_view *block_views,
                                 const int32_t *view_tiles, int shift, const int32_t *bstart, const int32_t *bend,
                                 int32_t *foff, void *workspace, size_t workspace_bytes, void *stream);
int gsmi_render_views_sh_two_level_record(int64_t V, int B, const gsmi_view *views, const gsmi_view *block_views,
                                          int n_tiles_total, const float *mean, const float *cov, const float *sh,
                                          int64_t sh_stride_g, int64_t sh_stride_c, const int32_t *sh_index,
                                          const float *alpha, const int32_t *tl, const int32_t *br,
                                          const int32_t *bstart, const int32_t *bend, const int32_t *bids, int shift,
                                          const int32_t *foff, int32_t *const *view_fids, int32_t *fend, float *out,
                                          int C, float thresh, float *T_final, int32_t *n_done, int exact,
                                          void *stream);
int gsmi_render_views_backward_sh_lists(int64_t V, int64_t P, int B, const gsmi_view *views, int n_tiles_total,
                                        const float *mean, const float *cov, const float *sh, int64_t sh_stride_g,
                                        int64_t sh_stride_c, const int32_t *sh_index, const float *alpha,
                                        const int32_t *start, const int32_t *end, const int32_t *const *view_ids,
                                        const float *T_final, const int32_t *n_done, const float *grad_out,
                                        float *grad_mean, float *grad_cov, float *grad_sh, float *grad_alpha, int C,
                                        float thresh, int flags, void *workspace, size_t workspace_bytes,
                                        void *stream);

/* Reverse-order adjoint of gsmi_render_views_sh; gradients of records seen
 * in several views accumulate (SH rows through sh_index).  T_final, n_done,
 * pre_bg, flags and workspace as for gsmi_render_sh_backward (workspace of
 * gsmi_backward_order_workspace_size(n_tiles_total) bytes: the tiles most
 * expensive first). */
int gsmi_render_views_backward_sh(int64_t V, int64_t P, int B, const gsmi_view *views, int n_tiles_total,
                                  const float *mean, const float *cov, const float *sh, int64_t sh_stride_g,
                                  int64_t sh_stride_c, const int32_t *sh_index, const float *alpha,
                                  const int32_t *start, const int32_t *end, const int32_t *gaussian_ids,
                                  const float *out, const float *T_final, const int32_t *n_done,
                                  const float *pre_bg, const float *grad_out, float *grad_mean, float *grad_cov,
                                  float *grad_sh, float *grad_alpha, int C, float thresh, const float *bg,
                                  int flags, void *workspace, size_t workspace_bytes, void *stream);

/* Projection adjoint over all views: per Gaussian the gradients of every
 * view that sees it are summed in view order, then chained to the raw
 * parameters (overwrite; zeros where no view sees the Gaussian). */
int gsmi_project_backward_views(int64_t N, int B, const gsmi_view *views, const float *mean, const float *qvec,
                                const float *log_s, const float *logit, const uint8_t *vis, const int32_t *vis_off,
                                const float *grad_mean2d, const float *grad_cov, const float *grad_alpha,
                                float *grad_mean, float *grad_qvec, float *grad_log_s, float *grad_logit,
                                void *stream);

/* ------------------------------------------------------------------------ */
/* Training-step stages after the adjoint                                    */
/* ------------------------------------------------------------------------ */

size_t gsmi_loss_workspace_size(int64_t n);

/* Base photometric term of get_loss_fn (utils/loss.py:5-22, l1_loss /
 * mse_loss with reduction "mean") over n image values: *loss (device float)
 * = mean |out - gt| (l2 = 0) or mean (out - gt)^2 (l2 = 1), and grad (may be
 * NULL) = d loss / d out.  Sums in double, fixed order. */
int gsmi_photometric_loss(int64_t n, const float *out, const float *gt, int l2, float *grad, float *loss,
                          void *workspace, size_t workspace_bytes, void *stream);

size_t gsmi_ssim_loss_workspace_size(int H, int W, int CH);

/* The whole loss of get_loss_fn (utils/loss.py:14-21) for images [H, W, CH]
 * (HWC, CH <= 4): *loss = ssim_mult * ssim_loss(out, gt, win) +
 * (1 - ssim_mult) * base (base as above), grad = d loss / d out.  ssim_loss
 * is kornia's: mean of (1 - SSIM) / 2 with a win x win Gaussian window of
 * sigma 1.5 (win odd, win / 2 < min(H, W), win <= 21), reflect padding,
 * C1 = 1e-4, C2 = 9e-4, 1e-12 added to the denominator. */
int gsmi_photometric_ssim_loss(int H, int W, int CH, const float *out, const float *gt, int win, float ssim_mult,
                               int l2, float *grad, float *loss, void *workspace, size_t workspace_bytes,
                               void *stream);

/* One torch.optim.Adam step (gs/sh_renderer.py:720-729: per-group lr, betas
 * (0.9, 0.99), eps) over up to 8 parameter groups in one launch; arrays of
 * n_groups device pointers (host arrays).  fresh = 1: moments start at zero
 * and step must be 1, exp_avg / exp_avg_sq are neither read nor written —
 * the step the reference takes, as it rebuilds its optimizer every
 * iteration (main_sh.py:238). */
int gsmi_adam_step(int n_groups, float *const *params, const float *const *grads, float *const *exp_avg,
                   float *const *exp_avg_sq, const int64_t *numel, const double *lr, int step, double beta1,
                   double beta2, double eps, int fresh, void *stream);

/* ------------------------------------------------------------------------ */
/* Adaptive density control (gs/sh_renderer.py:423-661)                      */
/* ------------------------------------------------------------------------ */

/* update_grads (sh_renderer.py:602-623): acc[index[r] or r] = max(acc, |g_r|)
 * (use_max) or acc + |g_r| for the n rows of grad [n, dim], dim 3 (mean
 * gradient) or 2 (mean2d gradient of the visible rows, index = original). */
int gsmi_grad_stats(int64_t n, const float *grad, int dim, const int32_t *index, int use_max, float *acc,
                    void *stream);

/* cnt[i] += vis[i] (sh_renderer.py:215-216). */
int gsmi_count_visible(int64_t N, const uint8_t *vis, int32_t *cnt, void *stream);

size_t gsmi_densify_workspace_size(int64_t N);

/* split_gaussians, phase 1 (sh_renderer.py:426-460): classify, scan and
 * count; synchronises and writes the host array counts[3] = (unsplit,
 * clone, split).  The new N is unsplit + clone + 2 * split. */
int gsmi_split_clone_count(int64_t N, const float *acc, const int32_t *cnt, int use_mean, float grad_thresh,
                           const float *log_s, float scale_thresh, int64_t *counts, void *workspace,
                           size_t workspace_bytes, void *stream);

/* split_gaussians, phase 2 (:461-534): writes the new raw parameter arrays
 * (unsplit rows, clones, then the split rows twice) from the same workspace;
 * noise [2 * split, 3] standard normal; sh_row = floats per sh row. */
int gsmi_split_clone_write(int64_t N, const float *mean, const float *qvec, const float *log_s, const float *logit,
                           const float *sh, int sh_row, const float *noise, float shrink, float *out_mean,
                           float *out_qvec, float *out_log_s, float *out_logit, float *out_sh, void *workspace,
                           size_t workspace_bytes, void *stream);

/* remove_low_alpha_gaussians (:536-558): keep sigmoid(logit) >= alpha_thresh,
 * order preserved; phase 1 synchronises and writes the host *kept. */
int gsmi_prune_alpha_count(int64_t N, const float *logit, float alpha_thresh, int64_t *kept, void *workspace,
                           size_t workspace_bytes, void *stream);
int gsmi_prune_alpha_write(int64_t N, const float *mean, const float *qvec, const float *log_s, const float *logit,
                           const float *sh, int sh_row, float *out_mean, float *out_qvec, float *out_log_s,
                           float *out_logit, float *out_sh, void *workspace, size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------------ */
/* Render server (continuous batching)                                       */
/* ------------------------------------------------------------------------ */
/*
 * Replaces the viewer's serial per-client render loop
 * (utils/viewer/viser_viewer.py:118-158, vis.py:18-32: one
 * SHRenderer.forward under no_grad per client, one at a time).  Scenes stay
 * resident (caller-owned device tensors, registered once); camera requests
 * are queued, and each of n_slots worker threads (one HIP stream each)
 * takes the oldest pending request plus, in arrival order, pending requests
 * of any scene with the same SH layout, up to max_views views and the pixel
 * budget of gsmi_server_configure, and runs them as one multi-view pass
 * (scenes counted and emitted group by group, binned and composited
 * together; each image written straight into its request's buffer) while
 * other workers' batches are still running.  All per-batch buffers come
 * from one device arena of arena_bytes allocated by gsmi_server_create; a
 * batch keeps the longest prefix of its views whose key/sort buffers fit
 * (the rest wait in the queue).  Each result equals the single-view forward
 * of that camera bit for bit.  The only entry points that allocate device
 * memory.
 */
typedef struct gsmi_server gsmi_server;

/* Camera record as CameraInfo + SHRenderer.forward derive it (frustum
 * planes utils/camera.py:249-283, tile grid and pixel mapping
 * gs/sh_renderer.py:238-248); host only, no GPU call.  tile_base and
 * pix_base are left 0. */
int gsmi_make_view(const float *c2w, double fx, double fy, double cx, double cy, int W, int H, double near_plane,
                   double far_plane, int tile_size, gsmi_view *out);

int gsmi_server_create(int device, int64_t arena_bytes, int max_views, int n_slots, float D, float frustum_radius,
                       float thresh, int exact, gsmi_server **out);

/* Batch shaping: at most max_batch_pixels pixels per batch (a single larger
 * request still runs alone; default unlimited), and queue_grid > 0 composites
 * each batch with that many workgroups taking tiles from per-XCD work queues
 * (0, the default: one workgroup per tile).  Applies to batches formed
 * afterwards. */
int gsmi_server_configure(gsmi_server *s, int64_t max_batch_pixels, int queue_grid);

/* Registers a resident scene (raw SHRenderer parameters on the device; the
 * caller keeps them alive and unchanged while requests use them). */
int gsmi_server_add_scene(gsmi_server *s, int64_t N, int C, int64_t sh_stride_c, const float *mean,
                          const float *qvec, const float *log_s, const float *logit, const float *sh,
                          int *scene_id);

/* Queues one camera (c2w host [3,4]); the image [H,W,3] is written to the
 * caller's device buffer `out`, which must stay valid until the request is
 * reported complete.  Thread-safe; returns at once. */
int gsmi_server_submit(gsmi_server *s, int scene_id, const float *c2w, double fx, double fy, double cx, double cy,
                       int W, int H, double near_plane, double far_plane, float *out, int64_t *request_id);

/* Up to max completed requests, in completion order (arrays may be NULL):
 * id, submit and completion time in ms since the server started, status,
 * number of views in the batch that served it. */
int gsmi_server_poll(gsmi_server *s, int64_t *ids, double *t_submit, double *t_done, int32_t *status,
                     int32_t *batch_views, int max, int *n_out);

/* Blocks until nothing is queued or in flight (GSMI_ERR_LAUNCH on timeout). */
int gsmi_server_drain(gsmi_server *s, double timeout_ms);

/* stats[0..5]: batches, views served, largest batch, admission splits,
 * failed requests, arena bytes. */
int gsmi_server_stats(gsmi_server *s, int64_t *stats);

/* The first n (<= 7) of: the six values above, then the batches that mixed
 * scenes. */
int gsmi_server_stats_ex(gsmi_server *s, int64_t *stats, int n);

/* Drains, stops the threads, frees the arena. */
int gsmi_server_destroy(gsmi_server *s);

/* ------------------------------------------------------------------------ */
/* Test hooks                                                                */
/* ------------------------------------------------------------------------ */

/* y[i] = the specified exponential of x[i] (binary32), n elements; used by
 * the tests to check it bit for bit against the oracle. */
int gsmi_exp_spec(int64_t n, const float *x, float *y, void *stream);

/* The same exponential in the branch-free form the compositing kernels use
 * for a Gaussian's exponent (x <= 0 or NaN); positive x[i] take the general
 * form.  Used by the tests to check the two agree bit for bit. */
int gsmi_exp_spec_nonpos(int64_t n, const float *x, float *y, void *stream);

/* The forward's quotient q[i] / det[i] from the staged reciprocal of det
 * (two FMA residual steps; the IEEE division where they give NaN or
 * |q| < 2^-100) and, when gauss is not null, its Gaussian value
 * exp(-radial / 2) (radial < 0 taken as 1000); no counterpart in the
 * reference.  Used by the tests to check both bit for bit. */
int gsmi_div_staged(int64_t n, const float *q, const float *det, float *out, float *gauss, void *stream);

/* The backward's launch order for given tile costs cost[0..n): the tiles
 * most expensive first by cost level, the order the backward derives from
 * its tile costs (workspace of gsmi_backward_order_workspace_size(n)
 * bytes); no counterpart in the reference.  Used by the tests to check the
 * order is a permutation sorted by level, stable within a level. */
int gsmi_tile_order_from_costs(int n, const uint32_t *cost, int32_t *order, void *workspace, size_t workspace_bytes,
                               void *stream);

/* The render server's scheduling policy on the host alone (no device use);
 * no counterpart in the reference.  gsmi_sched_arena_run: an arena of `size`
 * bytes driven by n_ops operations (op, x, y) — op 0 allocates x bytes (out:
 * offset, or -1), op 1 frees (offset x, bytes y), op 2 reads the largest
 * free range.  gsmi_sched_select_batch: requests 0..n-1 in arrival order
 * (scene[i], pixels[i]; layout[s] is scene s's SH layout key): the batch the
 * server takes (request indices, in batch order) and the queue left behind.
 * gsmi_sched_admit: the prefix of a counted batch (records vV, pairs vP,
 * blocks n_blocks per view) admitted against `largest` free bytes, whether
 * its totals fit int32 and its phase-B bytes. */
int gsmi_sched_arena_run(int64_t size, int n_ops, const int64_t *ops, int64_t *out);
int gsmi_sched_select_batch(int n, const int32_t *scene, const int64_t *pixels, int n_scenes,
                            const int32_t *layout, int max_views, int64_t max_pixels, int32_t *batch,
                            int32_t *n_batch, int32_t *rest);
int gsmi_sched_admit(int B0, const int64_t *vV, const int64_t *vP, const int64_t *n_blocks, uint64_t largest,
                     int32_t *B, int32_t *fits_int, uint64_t *bytes);

#ifdef __cplusplus
}
#endif

#endif /* GSMI_H */
