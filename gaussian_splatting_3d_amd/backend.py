"""``_backend``: the reference extension's function surface on the gfx950 C ABI.

Same names, positional arguments and in-place semantics as the pybind module
``_gs`` (gs/src/bindings.cpp:5-66, gs/src/render.h) for the functions on the
SHRenderer path, so code written against ``_backend.<name>(...)`` runs
unchanged.  Tensor checks raise RuntimeError with the reference's wording
(gs/src/include/common.h:29-54).  Every call goes to libgsmi.so on the
current torch stream; there is no CPU path.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check, ptr, stream_of

TILE = 16
GSMI_BWD_SPLIT_BF16 = 1  # include/gsmi.h
GSMI_BWD_COLOUR_MFMA = 2  # include/gsmi.h


def _check_cuda(name, t):
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise RuntimeError(f"{name} must be a CUDA tensor")


def _check_contig(name, t):
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be a contiguous tensor")


def check_float(name, t):
    _check_cuda(name, t)
    _check_contig(name, t)
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be a floating tensor")


def check_int(name, t):
    _check_cuda(name, t)
    _check_contig(name, t)
    if t.dtype != torch.int32:
        raise RuntimeError(f"{name} must be an int tensor")


def check_bool(name, t):
    _check_cuda(name, t)
    _check_contig(name, t)
    if t.dtype != torch.bool:
        raise RuntimeError(f"{name} must be an bool tensor")


def _workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


class _Backend:
    """Reference-named entry points (in place, caller-owned outputs)."""

    # gs/src/render.cu:15-43
    def culling_gaussian_bsphere(self, mean, qvec, svec, normal, pts, mask, thresh):
        for n, t in (("mean", mean), ("qvec", qvec), ("svec", svec), ("normal", normal), ("pts", pts)):
            check_float(n, t)
        check_bool("mask", mask)
        L = _lib.load()
        N = mean.shape[0]
        check(L.gsmi_culling_gaussian_bsphere(N, ptr(mean), ptr(qvec), ptr(svec), ptr(normal), ptr(pts),
                                              ptr(mask), float(thresh), stream_of(mean.device)),
              "culling_gaussian_bsphere")

    # gs/src/render.cu:380-397
    def tile_culling_aabb_start_end(self, aabb_topleft, aabb_bottomright, gaussian_ids, start, end, depth,
                                    n_tiles_h, n_tiles_w, sorted_keys=None):
        check_int("aabb_topleft", aabb_topleft)
        check_int("aabb_bottomright", aabb_bottomright)
        check_int("gaussian_ids", gaussian_ids)
        check_int("start", start)
        check_int("end", end)
        check_float("depth", depth)
        L = _lib.load()
        V = aabb_topleft.shape[0]
        P = gaussian_ids.shape[0]
        n_tiles = int(n_tiles_h) * int(n_tiles_w)
        ws = _workspace(L.gsmi_binning_workspace_size(V, P, n_tiles), depth.device)
        check(L.gsmi_tile_culling_aabb_start_end(V, P, ptr(aabb_topleft), ptr(aabb_bottomright), ptr(depth),
                                                 int(n_tiles_h), int(n_tiles_w), ptr(gaussian_ids), ptr(start),
                                                 ptr(end), ptr(sorted_keys), ptr(ws), ws.numel(), 1,
                                                 stream_of(depth.device)),
              "tile_culling_aabb_start_end")

    # ------------------------------------------------------------------ SH
    @staticmethod
    def _sh_checks(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w):
        check_float("mean", mean)
        check_float("cov", cov)
        check_float("alpha", alpha)
        check_float("sh_coeffs", sh_coeffs)
        check_int("start", start)
        check_int("end", end)
        check_int("gaussian_ids", gaussian_ids)
        check_float("out", out)
        check_float("topleft", topleft)
        check_float("c2w", c2w)

    def sh_forward(self, mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w, tile_size,
                   n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh, bg_rgb=None, T_final=None,
                   n_done=None, pre_bg=None, write_empty=False, exact=False, sh_stride_g=None, sh_stride_c=None,
                   sh_index=None):
        """Forward SH compositing with optional per-pixel state outputs."""
        L = _lib.load()
        C = int(C)
        sg = 3 * C * C if sh_stride_g is None else int(sh_stride_g)
        sc = C * C if sh_stride_c is None else int(sh_stride_c)
        check(L.gsmi_render_sh_forward(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(sh_coeffs), sg, sc, ptr(sh_index),
            ptr(alpha), ptr(start), ptr(end), ptr(gaussian_ids), ptr(out), ptr(topleft), ptr(c2w), int(tile_size),
            int(n_tiles_h), int(n_tiles_w), float(pixel_size_x), float(pixel_size_y), int(H), int(W), C,
            float(thresh), ptr(bg_rgb), ptr(T_final), ptr(n_done), ptr(pre_bg), 1 if write_empty else 0,
            1 if exact else 0, stream_of(out.device)), "render_sh_forward")

    def sh_forward_two_level(self, mean, cov, sh_coeffs, alpha, tl, br, block_start, block_end, block_ids, shift,
                             tile_off, tile_ids, tile_end, out, topleft, c2w, n_tiles_h, n_tiles_w, pixel_size_x,
                             pixel_size_y, H, W, C, thresh, T_final, n_done, exact=False, sh_stride_g=None,
                             sh_stride_c=None, sh_index=None):
        """Forward SH compositing from lists per block of 2^shift x 2^shift
        tiles; records each tile's filtered list in tile_ids[tile_off[t] ..
        tile_end[t]) for the backward.  No counterpart in the reference."""
        L = _lib.load()
        C = int(C)
        sg = 3 * C * C if sh_stride_g is None else int(sh_stride_g)
        sc = C * C if sh_stride_c is None else int(sh_stride_c)
        check(L.gsmi_render_sh_two_level(
            mean.shape[0], ptr(mean), ptr(cov), ptr(sh_coeffs), sg, sc, ptr(sh_index), ptr(alpha), ptr(tl), ptr(br),
            ptr(block_start), ptr(block_end), ptr(block_ids), int(shift), ptr(tile_off), ptr(tile_ids),
            ptr(tile_end), ptr(out), ptr(topleft), ptr(c2w), int(n_tiles_h), int(n_tiles_w), float(pixel_size_x),
            float(pixel_size_y), int(H), int(W), C, float(thresh), ptr(T_final), ptr(n_done), 1 if exact else 0,
            stream_of(out.device)), "render_sh_two_level")

    def sh_backward(self, mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, T_final, n_done, pre_bg,
                    grad_out, grad_mean, grad_cov, grad_sh_coeffs, grad_alpha, topleft, c2w, tile_size, n_tiles_h,
                    n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh, bg_rgb=None, sh_stride_g=None,
                    sh_stride_c=None, sh_index=None, split_bf16=False, colour_mfma=False):
        """Reverse-order adjoint (accumulates).  From the saved per-pixel
        state, with the tiles launched most expensive first (a small
        workspace holds the order); with T_final = n_done = None the launch
        recomputes that state itself.  split_bf16: the SH-gradient products
        as a three-term bf16 split (opt-in, not fp32).  colour_mfma: the
        colours formed on fp32 matrix cores (opt-in)."""
        L = _lib.load()
        C = int(C)
        sg = 3 * C * C if sh_stride_g is None else int(sh_stride_g)
        sc = C * C if sh_stride_c is None else int(sh_stride_c)
        ws, nws = None, 0
        if n_done is not None:
            nws = int(L.gsmi_backward_order_workspace_size(int(n_tiles_h) * int(n_tiles_w)))
            ws = torch.empty(max(nws, 1), dtype=torch.uint8, device=grad_out.device)
        check(L.gsmi_render_sh_backward(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(sh_coeffs), sg, sc, ptr(sh_index),
            ptr(alpha), ptr(start), ptr(end), ptr(gaussian_ids), ptr(out), ptr(T_final), ptr(n_done), ptr(pre_bg),
            ptr(grad_out), ptr(grad_mean), ptr(grad_cov), ptr(grad_sh_coeffs), ptr(grad_alpha), ptr(topleft),
            ptr(c2w), int(tile_size), int(n_tiles_h), int(n_tiles_w), float(pixel_size_x), float(pixel_size_y),
            int(H), int(W), C, float(thresh), ptr(bg_rgb),
            (GSMI_BWD_SPLIT_BF16 if split_bf16 else 0) | (GSMI_BWD_COLOUR_MFMA if colour_mfma else 0), ptr(ws), nws,
            stream_of(grad_out.device)), "render_sh_backward")

    # gs/src/render.cu:483-544
    def tile_based_vol_rendering_sh(self, mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft,
                                    c2w, tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C,
                                    thresh):
        self._sh_checks(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w)
        if not 1 <= int(C) <= 4:
            return  # the reference's switch has no default case (render.cu:506-543)
        check(_lib.load().gsmi_tile_based_vol_rendering_sh(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(sh_coeffs), ptr(alpha), ptr(start),
            ptr(end), ptr(gaussian_ids), ptr(out), ptr(topleft), ptr(c2w), int(tile_size), int(n_tiles_h),
            int(n_tiles_w), float(pixel_size_x), float(pixel_size_y), int(H), int(W), int(C), float(thresh),
            stream_of(out.device)), "tile_based_vol_rendering_sh")

    def _grad_checks(self, grad_mean, grad_cov, grad_sh_coeffs, grad_alpha, grad_out, sh_name="grad_sh_coeffs"):
        for n, t in (("grad_mean", grad_mean), ("grad_cov", grad_cov), (sh_name, grad_sh_coeffs),
                     ("grad_alpha", grad_alpha), ("grad_out", grad_out)):
            check_float(n, t)

    # gs/src/render.cu:546-624
    def tile_based_vol_rendering_backward_sh(self, mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out,
                                             grad_mean, grad_cov, grad_sh_coeffs, grad_alpha, grad_out, topleft,
                                             c2w, tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H,
                                             W, C, thresh):
        self._sh_checks(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w)
        self._grad_checks(grad_mean, grad_cov, grad_sh_coeffs, grad_alpha, grad_out)
        if not 1 <= int(C) <= 4:
            return
        check(_lib.load().gsmi_tile_based_vol_rendering_backward_sh(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(sh_coeffs), ptr(alpha), ptr(start),
            ptr(end), ptr(gaussian_ids), ptr(out), ptr(grad_mean), ptr(grad_cov), ptr(grad_sh_coeffs),
            ptr(grad_alpha), ptr(grad_out), ptr(topleft), ptr(c2w), int(tile_size), int(n_tiles_h), int(n_tiles_w),
            float(pixel_size_x), float(pixel_size_y), int(H), int(W), int(C), float(thresh),
            stream_of(grad_out.device)), "tile_based_vol_rendering_backward_sh")

    # gs/src/render.cu:780-842 (vol_render_bg.h:12-100)
    def tile_based_vol_rendering_sh_with_bg(self, mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out,
                                            topleft, c2w, tile_size, n_tiles_h, n_tiles_w, pixel_size_x,
                                            pixel_size_y, H, W, C, thresh, bg_rgb):
        self._sh_checks(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w)
        check_float("bg_rgb", bg_rgb)
        if not 1 <= int(C) <= 4:
            return
        check(_lib.load().gsmi_tile_based_vol_rendering_sh_with_bg(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(sh_coeffs), ptr(alpha), ptr(start),
            ptr(end), ptr(gaussian_ids), ptr(out), ptr(topleft), ptr(c2w), int(tile_size), int(n_tiles_h),
            int(n_tiles_w), float(pixel_size_x), float(pixel_size_y), int(H), int(W), int(C), float(thresh),
            ptr(bg_rgb), stream_of(out.device)), "tile_based_vol_rendering_sh_with_bg")

    # gs/src/render.cu:844-925 (vol_render_bg.h:121-234)
    def tile_based_vol_rendering_backward_sh_with_bg(self, mean, cov, sh_coeffs, alpha, start, end, gaussian_ids,
                                                     out, grad_mean, grad_cov, grad_sh_coeffs, grad_alpha, grad_out,
                                                     topleft, c2w, tile_size, n_tiles_h, n_tiles_w, pixel_size_x,
                                                     pixel_size_y, H, W, C, thresh, bg_rgb):
        self._sh_checks(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w)
        self._grad_checks(grad_mean, grad_cov, grad_sh_coeffs, grad_alpha, grad_out)
        check_float("bg_rgb", bg_rgb)
        if not 1 <= int(C) <= 4:
            return
        check(_lib.load().gsmi_tile_based_vol_rendering_backward_sh_with_bg(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(sh_coeffs), ptr(alpha), ptr(start),
            ptr(end), ptr(gaussian_ids), ptr(out), ptr(grad_mean), ptr(grad_cov), ptr(grad_sh_coeffs),
            ptr(grad_alpha), ptr(grad_out), ptr(topleft), ptr(c2w), int(tile_size), int(n_tiles_h), int(n_tiles_w),
            float(pixel_size_x), float(pixel_size_y), int(H), int(W), int(C), float(thresh), ptr(bg_rgb),
            stream_of(grad_out.device)), "tile_based_vol_rendering_backward_sh_with_bg")

    # ------------------------------------------------------------------ RGB
    def rgb_forward(self, mean, cov, color, alpha, start, end, gaussian_ids, out, topleft, tile_size, n_tiles_h,
                    n_tiles_w, pixel_size_x, pixel_size_y, H, W, thresh, T_final=None, n_done=None,
                    write_empty=False):
        L = _lib.load()
        check(L.gsmi_render_rgb_forward(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(color), ptr(alpha), ptr(start),
            ptr(end), ptr(gaussian_ids), ptr(out), ptr(topleft), int(tile_size), int(n_tiles_h), int(n_tiles_w),
            float(pixel_size_x), float(pixel_size_y), int(H), int(W), float(thresh), ptr(T_final), ptr(n_done),
            1 if write_empty else 0, stream_of(out.device)), "render_rgb_forward")

    def rgb_backward(self, mean, cov, color, alpha, start, end, gaussian_ids, out, T_final, n_done, grad_out,
                     grad_mean, grad_cov, grad_color, grad_alpha, topleft, tile_size, n_tiles_h, n_tiles_w,
                     pixel_size_x, pixel_size_y, H, W, thresh):
        """RGB adjoint from saved T_final / n_done, or (both None) with the
        forward state recomputed inside the launch."""
        L = _lib.load()
        check(L.gsmi_render_rgb_backward(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(color), ptr(alpha), ptr(start),
            ptr(end), ptr(gaussian_ids), ptr(out), ptr(T_final), ptr(n_done), ptr(grad_out), ptr(grad_mean),
            ptr(grad_cov), ptr(grad_color), ptr(grad_alpha), ptr(topleft), int(tile_size), int(n_tiles_h),
            int(n_tiles_w), float(pixel_size_x), float(pixel_size_y), int(H), int(W), float(thresh),
            stream_of(grad_out.device)), "render_rgb_backward")

    # gs/src/render.cu:399-441
    def tile_based_vol_rendering_start_end(self, mean, cov, color, alpha, start, end, gaussian_ids, out, topleft,
                                           tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W,
                                           thresh):
        for n, t in (("mean", mean), ("cov", cov), ("color", color), ("alpha", alpha), ("out", out),
                     ("topleft", topleft)):
            check_float(n, t)
        for n, t in (("start", start), ("end", end), ("gaussian_ids", gaussian_ids)):
            check_int(n, t)
        check(_lib.load().gsmi_tile_based_vol_rendering_start_end(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(color), ptr(alpha), ptr(start),
            ptr(end), ptr(gaussian_ids), ptr(out), ptr(topleft), int(tile_size), int(n_tiles_h), int(n_tiles_w),
            float(pixel_size_x), float(pixel_size_y), int(H), int(W), float(thresh), stream_of(out.device)),
            "tile_based_vol_rendering_start_end")

    # gs/src/render.cu:443-481
    def tile_based_vol_rendering_backward_start_end(self, mean, cov, color, alpha, start, end, gaussian_ids, out,
                                                    grad_mean, grad_cov, grad_color, grad_alpha, grad_out, topleft,
                                                    tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H,
                                                    W, thresh):
        for n, t in (("mean", mean), ("cov", cov), ("color", color), ("alpha", alpha), ("out", out),
                     ("grad_mean", grad_mean), ("grad_cov", grad_cov), ("grad_color", grad_color),
                     ("grad_alpha", grad_alpha), ("grad_out", grad_out), ("topleft", topleft)):
            check_float(n, t)
        for n, t in (("start", start), ("end", end), ("gaussian_ids", gaussian_ids)):
            check_int(n, t)
        check(_lib.load().gsmi_tile_based_vol_rendering_backward_start_end(
            mean.shape[0], gaussian_ids.shape[0], ptr(mean), ptr(cov), ptr(color), ptr(alpha), ptr(start),
            ptr(end), ptr(gaussian_ids), ptr(out), ptr(grad_mean), ptr(grad_cov), ptr(grad_color), ptr(grad_alpha),
            ptr(grad_out), ptr(topleft), int(tile_size), int(n_tiles_h), int(n_tiles_w), float(pixel_size_x),
            float(pixel_size_y), int(H), int(W), float(thresh), stream_of(grad_out.device)),
            "tile_based_vol_rendering_backward_start_end")


_backend = _Backend()
