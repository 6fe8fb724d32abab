"""Fused gfx950 pipeline for one camera: raw Gaussians -> image, and back.

Forward (all on the current torch stream):
  1. gsmi_preprocess_count  — activation, frustum cull, projection, tile
     bounds and pair count per raw Gaussian; prefix sums give the compacted
     index and the pair offset of every visible Gaussian.
  2. one host read of (V, P) to size the outputs (the reference reads P the
     same way, gs/culling.py:33-35);
  3. gsmi_preprocess_emit   — compacted SoA records (mean2d, cov2d, depth,
     alpha, tile bounds, original index);
  4-5. gsmi_tile_culling_aabb_start_end — depth-first binning: stable sort of
     the visible Gaussians by depth, pair expansion in that order, stable sort
     of the pairs by tile, start/end per tile (the order of the reference's
     (tile<<32 | depth) key sort);
  6. gsmi_render_sh_forward — compositing; SH rows are read from the
     raw [N,3,C_max^2] tensor through the original index (no gather copy).
  Two-level form (cfg.two_level, the default without background): steps 1-5
  count, sort and range the pairs per 64-px block of 4 x 4 tiles (about a
  sixth of the tile pairs at 3M Gaussians / 1080p), and step 6 is
  gsmi_render_sh_two_level, which filters each tile's list out of its block's
  depth-ordered list as it composites (the same list, so the same image, T and
  n_done) and records it for the backward.
Backward:
  7. gsmi_render_sh_backward — reverse-order adjoint; SH
     gradients land directly in the raw-layout [N,3,C_max^2] buffer;
  8. gsmi_project_backward  — projection/activation adjoint with the
     reference's detaches, zero for culled Gaussians.
This reproduces SHRenderer.forward (gs/sh_renderer.py:188-316) and the
gradients its autograd graph produces.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, host_f32, ptr, stream_of
from .backend import _backend
from .camera import CameraInfo

TILE = 16


def _bits_for(n):
    b = 0
    while (1 << b) < n:
        b += 1
    return b


@dataclass
class Binning:
    """Per-view intermediates kept for the backward pass and for inspection."""

    N: int
    V: int
    P: int
    vis: torch.Tensor
    vis_off: torch.Tensor
    pair_off: torch.Tensor
    mean2d: torch.Tensor
    cov2d: torch.Tensor
    depth: torch.Tensor
    alpha: torch.Tensor
    tl: torch.Tensor
    br: torch.Tensor
    orig: torch.Tensor
    keys: torch.Tensor | None  # int64 storage of the sorted uint64 keys (when asked for)
    ids: torch.Tensor
    start: torch.Tensor
    end: torch.Tensor
    c2w_host: np.ndarray
    consts: dict
    # two-level binning: lists per block of 2^shift x 2^shift tiles (ids,
    # start and end above are then the per-tile lists the forward recorded)
    shift: int = 0
    block_ids: torch.Tensor | None = None
    block_start: torch.Tensor | None = None
    block_end: torch.Tensor | None = None
    # camera constants on the device: c2w [3,4], topleft [2] (and bg [3])
    c2w_d: torch.Tensor | None = None
    topleft_d: torch.Tensor | None = None
    bg_d: torch.Tensor | None = None


def preprocess_and_bin(mean, qvec, log_s, logit, c2w, camera: CameraInfo, *, tile_size=TILE, D=6.0,
                       frustum_culling_radius=1.0, skip_frustum_culling=False, want_keys=False,
                       block_shift=0, bg_rgb=None, on_sync=None) -> Binning:
    """Steps 1-5 of the module docstring.  With block_shift > 0 the pairs are
    binned per block of 2^shift x 2^shift tiles instead (P counts block
    pairs); tl / br stay the tile bounds.  on_sync(V), when given, runs right
    after the pair count is read back: device work it enqueues runs while the
    host prepares the emission (the GPU would idle there otherwise)."""
    L = _lib.load()
    dev = mean.device
    st = stream_of(dev)
    N = int(mean.shape[0])
    c2w_h = np.ascontiguousarray(
        (c2w.detach().cpu().numpy() if isinstance(c2w, torch.Tensor) else np.asarray(c2w)), np.float32
    ).reshape(3, 4)
    normals, pts = camera.frustum_planes_np(c2w_h)
    ic = camera.image_constants(tile_size)
    f32 = np.float32
    fx, fy, cx, cy = float(f32(camera.fx)), float(f32(camera.fy)), float(f32(camera.cx)), float(f32(camera.cy))
    vis = torch.empty(N, dtype=torch.uint8, device=dev)
    vis_off = torch.empty(N, dtype=torch.int32, device=dev)
    pair_off = torch.empty(N, dtype=torch.int32, device=dev)
    totals = torch.empty(2, dtype=torch.int64, device=dev)
    ws = torch.empty(max(int(L.gsmi_preprocess_workspace_size(N)), 1), dtype=torch.uint8, device=dev)
    c2w_c = host_f32(c2w_h)
    if block_shift and tile_size != TILE:
        raise ValueError("two-level binning needs 16-pixel tiles")
    check(L.gsmi_preprocess_count(N, ptr(mean), ptr(qvec), ptr(log_s), c2w_c, host_f32(normals), host_f32(pts),
                                  fx, fy, cx, cy, ic["W"], ic["H"], float(D), float(frustum_culling_radius),
                                  1 if skip_frustum_culling else 0, tile_size << block_shift, ptr(vis), ptr(vis_off),
                                  ptr(pair_off), ptr(totals), ptr(ws), ws.numel(), st), "preprocess_count")
    V, P = (int(x) for x in totals.cpu().tolist())
    if P >= 2 ** 31:
        raise RuntimeError(f"{P} (Gaussian, tile) pairs exceed the int32 range of the tile lists")
    # The camera constants go to the device now, while the stream is empty: a
    # copy from pageable memory returns only when the stream has drained, so
    # made after the binning launches it would idle the GPU until the host
    # caught up again.
    consts = [c2w_h.reshape(-1), np.asarray(ic["topleft"], np.float32)]
    if bg_rgb is not None:
        consts.append(np.asarray(bg_rgb, np.float32))
    consts_d = torch.from_numpy(np.concatenate(consts).astype(np.float32)).to(dev)
    if on_sync is not None:
        on_sync(V)
    mean2d = torch.empty(V, 2, dtype=torch.float32, device=dev)
    cov2d = torch.empty(V, 2, 2, dtype=torch.float32, device=dev)
    depth = torch.empty(V, 1, dtype=torch.float32, device=dev)
    alpha = torch.empty(V, dtype=torch.float32, device=dev)
    tl = torch.empty(V, 2, dtype=torch.int32, device=dev)
    br = torch.empty(V, 2, dtype=torch.int32, device=dev)
    orig = torch.empty(V, dtype=torch.int32, device=dev)
    ids = torch.empty(P, dtype=torch.int32, device=dev)
    check(L.gsmi_preprocess_emit(N, ptr(mean), ptr(qvec), ptr(log_s), ptr(logit), c2w_c, fx, fy, cx, cy, ic["W"],
                                 ic["H"], float(D), tile_size, ptr(vis), ptr(vis_off), ptr(mean2d), ptr(cov2d),
                                 ptr(depth), ptr(alpha), ptr(tl), ptr(br), ptr(orig), st), "preprocess_emit")
    nth, ntw = ic["n_tiles_h"], ic["n_tiles_w"]
    if block_shift:  # the binning shifts the bounds itself (flags bits 8-10)
        nth = (nth + (1 << block_shift) - 1) >> block_shift
        ntw = (ntw + (1 << block_shift) - 1) >> block_shift
    n_tiles = nth * ntw
    start = torch.empty(n_tiles, dtype=torch.int32, device=dev)
    end = torch.empty(n_tiles, dtype=torch.int32, device=dev)
    keys = torch.empty(P, dtype=torch.int64, device=dev) if want_keys else None
    bws = torch.empty(max(int(L.gsmi_binning_workspace_size(V, P, n_tiles)), 1), dtype=torch.uint8, device=dev)
    check(L.gsmi_tile_culling_aabb_start_end(V, P, ptr(tl), ptr(br), ptr(depth), nth, ntw, ptr(ids), ptr(start),
                                             ptr(end), ptr(keys), ptr(bws), bws.numel(), block_shift << 8, st),
          "binning")
    b = Binning(N, V, P, vis, vis_off, pair_off, mean2d, cov2d, depth, alpha, tl, br, orig, keys, ids, start, end,
                c2w_h, ic)
    b.c2w_d, b.topleft_d = consts_d[:12], consts_d[12:14]
    b.bg_d = consts_d[14:17] if bg_rgb is not None else None
    if block_shift:
        b.shift, b.block_ids, b.block_start, b.block_end = block_shift, ids, start, end
    return b


BLOCK_SHIFT = 2  # two-level binning: blocks of 4 x 4 tiles (64 px)
# The recorded per-tile lists of the two-level forward are indexed with int32
# offsets and sized for (tiles per block) x (block pairs) entries; a view
# whose bound reaches this limit is binned per tile instead.
TWO_LEVEL_ENTRY_LIMIT = 2 ** 31


def _forward_two_level(b: Binning, sh_coeffs, out, topleft, c2w_d, C, thresh, T, nd, exact):
    """Two-level forward: composite from the block lists, recording each
    tile's filtered list; afterwards b.ids / b.start / b.end are those
    per-tile lists (start = offset, end = offset + recorded length), which
    the backward reads as ordinary tile lists."""
    L = _lib.load()
    ic = b.consts
    dev = out.device
    st = stream_of(dev)
    nth, ntw = ic["n_tiles_h"], ic["n_tiles_w"]
    n_tiles = nth * ntw
    cap = (1 << (2 * b.shift)) * b.P  # >= sum over tiles of their block's list length
    if cap >= TWO_LEVEL_ENTRY_LIMIT:  # forward_view bins such views per tile
        raise RuntimeError(f"two-level lists need {cap} entries, beyond the int32 range")
    foff = torch.empty(n_tiles, dtype=torch.int32, device=dev)
    fend = torch.empty(n_tiles, dtype=torch.int32, device=dev)
    fids = torch.empty(cap, dtype=torch.int32, device=dev)  # empty when nothing is visible (P = 0)
    ows = torch.empty(max(int(L.gsmi_two_level_offsets_workspace_size(n_tiles)), 1), dtype=torch.uint8, device=dev)
    check(L.gsmi_two_level_offsets(nth, ntw, b.shift, ptr(b.block_start), ptr(b.block_end), ptr(foff), ptr(ows),
                                   ows.numel(), st), "two_level_offsets")
    cmax2 = sh_coeffs.shape[-1]
    _backend.sh_forward_two_level(b.mean2d, b.cov2d, sh_coeffs, b.alpha, b.tl, b.br, b.block_start, b.block_end,
                                  b.block_ids, b.shift, foff, fids, fend, out, topleft, c2w_d, nth, ntw,
                                  ic["pixel_size_x"], ic["pixel_size_y"], ic["H"], ic["W"], C, thresh, T, nd,
                                  exact=exact, sh_stride_g=3 * cmax2, sh_stride_c=cmax2, sh_index=b.orig)
    b.ids, b.start, b.end = fids, foff, fend


@dataclass
class ViewState:
    """Everything the backward pass of one view needs."""

    b: Binning
    C: int
    cfg: object
    topleft: torch.Tensor
    c2w_d: torch.Tensor
    bg: torch.Tensor | None
    out: torch.Tensor
    T: torch.Tensor
    n_done: torch.Tensor
    pre_bg: torch.Tensor | None
    # zeroed gradient buffers made during the forward (prepare_backward);
    # taken by the first backward_view
    g_sh: torch.Tensor | None = None
    bnd: torch.Tensor | None = None


def forward_view(mean, qvec, log_s, logit, sh_coeffs, c2w, camera, C, cfg, exact=False, prepare_backward=False):
    """Steps 1-6: returns (image [H*W*3], ViewState).  prepare_backward: also
    make the backward's zeroed gradient buffers, at the point of the forward
    where the GPU would otherwise wait for the host."""
    two = bool(getattr(cfg, "two_level", False)) and not cfg.bg and cfg.tile_size == TILE
    early = {}

    def on_sync(V):
        early["g_sh"] = torch.zeros_like(sh_coeffs)
        early["bnd"] = torch.zeros(V * 7, dtype=torch.float32, device=mean.device)

    def bin_view(shift):
        return preprocess_and_bin(mean, qvec, log_s, logit, c2w, camera, tile_size=cfg.tile_size,
                                  D=cfg.tile_culling_radius, frustum_culling_radius=cfg.frustum_culling_radius,
                                  skip_frustum_culling=cfg.skip_frustum_culling, block_shift=shift,
                                  bg_rgb=cfg.bg_rgb if cfg.bg else None,
                                  on_sync=on_sync if prepare_backward else None)

    b = bin_view(BLOCK_SHIFT if two else 0)
    if two and (1 << (2 * BLOCK_SHIFT)) * b.P >= TWO_LEVEL_ENTRY_LIMIT:
        two = False  # the recorded lists would overflow int32 offsets: per-tile lists (same image)
        b = bin_view(0)
    ic = b.consts
    H, W = ic["H"], ic["W"]
    dev = mean.device
    topleft, c2w_d, bg = b.topleft_d, b.c2w_d, b.bg_d
    out = torch.empty(H * W * 3, dtype=torch.float32, device=dev)
    T = torch.empty(H * W, dtype=torch.float32, device=dev)
    nd = torch.empty(H * W, dtype=torch.int32, device=dev)
    pre = torch.empty(H * W * 3, dtype=torch.float32, device=dev) if cfg.bg else None
    cmax2 = sh_coeffs.shape[-1]
    if two:
        _forward_two_level(b, sh_coeffs, out, topleft, c2w_d, C, cfg.T_thresh, T, nd, exact)
    else:
        _backend.sh_forward(b.mean2d, b.cov2d, sh_coeffs, b.alpha, b.start, b.end, b.ids, out, topleft, c2w_d,
                            cfg.tile_size, ic["n_tiles_h"], ic["n_tiles_w"], ic["pixel_size_x"], ic["pixel_size_y"],
                            H, W, C, cfg.T_thresh, bg_rgb=bg, T_final=T, n_done=nd, pre_bg=pre, write_empty=True,
                            exact=exact, sh_stride_g=3 * cmax2, sh_stride_c=cmax2, sh_index=b.orig)
    return out, ViewState(b, int(C), cfg, topleft, c2w_d, bg, out, T, nd, pre, early.get("g_sh"), early.get("bnd"))


def backward_view(s: ViewState, mean, qvec, log_s, logit, sh_coeffs, grad_out, g_sh=None, split_bf16=False,
                  colour_mfma=False):
    """Steps 7-8: gradients w.r.t. the raw parameters.  SH gradients are
    accumulated into g_sh when given (several views), else a new buffer.
    Returns (g_mean, g_qvec, g_log_s, g_logit, g_sh, boundary) where boundary
    holds the compositing gradients (mean2d, cov2d, alpha) of visible
    Gaussians.  split_bf16: SH-gradient products on bf16 matrix cores as a
    three-term split (opt-in; the default forms them in fp32); colour_mfma:
    the colours formed on fp32 matrix cores (opt-in)."""
    b, cfg, ic = s.b, s.cfg, s.b.consts
    dev = mean.device
    L = _lib.load()
    # one fill for the three boundary gradients; cov first so every part
    # keeps the alignment of a separate allocation (16 / 16 / 8 bytes)
    bnd, s.bnd = s.bnd, None
    if bnd is None:
        bnd = torch.zeros(b.V * 7, dtype=torch.float32, device=dev)
    g_cov = bnd[: 4 * b.V].view(b.V, 2, 2)
    g_m2d = bnd[4 * b.V: 6 * b.V].view(b.V, 2)
    g_al = bnd[6 * b.V:]
    pre_sh, s.g_sh = s.g_sh, None
    if g_sh is None:
        g_sh = pre_sh if pre_sh is not None else torch.zeros_like(sh_coeffs)
    if b.V == 0:  # nothing visible: every gradient is zero (no launch takes empty buffers)
        return (torch.zeros_like(mean), torch.zeros_like(qvec), torch.zeros_like(log_s), torch.zeros_like(logit),
                g_sh, dict(mean2d=g_m2d, cov=g_cov, alpha=g_al))
    cmax2 = sh_coeffs.shape[-1]
    _backend.sh_backward(b.mean2d, b.cov2d, sh_coeffs, b.alpha, b.start, b.end, b.ids, s.out, s.T, s.n_done,
                         s.pre_bg, grad_out.contiguous().view(-1), g_m2d, g_cov, g_sh, g_al, s.topleft, s.c2w_d,
                         cfg.tile_size, ic["n_tiles_h"], ic["n_tiles_w"], ic["pixel_size_x"], ic["pixel_size_y"],
                         ic["H"], ic["W"], s.C, cfg.T_thresh, bg_rgb=s.bg, sh_stride_g=3 * cmax2, sh_stride_c=cmax2,
                         sh_index=b.orig, split_bf16=split_bf16, colour_mfma=colour_mfma)
    g_mean = torch.empty_like(mean)
    g_q = torch.empty_like(qvec)
    g_ls = torch.empty_like(log_s)
    g_lg = torch.empty_like(logit)
    check(L.gsmi_project_backward(b.N, ptr(mean), ptr(qvec), ptr(log_s), ptr(logit), host_f32(b.c2w_host),
                                  ptr(b.vis), ptr(b.vis_off), ptr(g_m2d), ptr(g_cov), ptr(g_al), ptr(g_mean),
                                  ptr(g_q), ptr(g_ls), ptr(g_lg), stream_of(dev)), "project_backward")
    return g_mean, g_q, g_ls, g_lg, g_sh, dict(mean2d=g_m2d, cov=g_cov, alpha=g_al)


class _RasterizeSH(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mean, qvec, log_s, logit, sh_coeffs, c2w, camera, C, cfg, prepare_backward):
        mean, qvec, log_s, logit, sh_coeffs = (t.contiguous() for t in (mean, qvec, log_s, logit, sh_coeffs))
        out, state = forward_view(mean, qvec, log_s, logit, sh_coeffs, c2w, camera, C, cfg,
                                  prepare_backward=prepare_backward)
        ctx.state = state
        ctx.save_for_backward(mean, qvec, log_s, logit, sh_coeffs)
        ic = state.b.consts
        return out.view(ic["H"], ic["W"], 3)

    @staticmethod
    def backward(ctx, grad):
        mean, qvec, log_s, logit, sh_coeffs = ctx.saved_tensors
        g_mean, g_q, g_ls, g_lg, g_sh, _ = backward_view(ctx.state, mean, qvec, log_s, logit, sh_coeffs, grad)
        return g_mean, g_q, g_ls, g_lg, g_sh, None, None, None, None, None


def rasterize_sh(mean, qvec, svec_before_activation, alpha_before_activation, sh_coeffs, c2w, camera, C, cfg):
    """Image [H,W,3] of one camera; differentiable w.r.t. the five raw
    parameter tensors (SHRenderer storage convention)."""
    params = (mean, qvec, svec_before_activation, alpha_before_activation, sh_coeffs)
    # a backward will follow only with grad mode on and a parameter that needs it
    prepare = torch.is_grad_enabled() and any(t.requires_grad for t in params)
    return _RasterizeSH.apply(*params, c2w, camera, int(C), cfg, prepare)
