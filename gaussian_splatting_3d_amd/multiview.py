"""Multi-view batches: B cameras over one set of Gaussians in one pass per stage.

The reference renders one camera per ``SHRenderer.forward`` call
(gs/sh_renderer.py:188-316) and its viewer loops over clients
(utils/viewer/viser_viewer.py:118-158).  Here the B cameras of a batch share
every launch:

  1. gsmi_preprocess_views_count — cull/project/count every (view, Gaussian),
     one scan numbers the visible records view-major;
  2. one host read of the per-view bases (B+1 record and pair offsets);
  3. gsmi_preprocess_views_emit — compacted records of all views;
  4. gsmi_bin_views — one depth-first binning over all views, tiles numbered
     per view from its tile_base;
  5. gsmi_render_views_sh — one compositing launch over every tile of every
     view (mixed resolutions), writing a flat [sum(H*W), 3] image;
backward:
  6. gsmi_render_views_backward_sh — one adjoint launch; SH gradients of a
     Gaussian seen in several views accumulate in the raw-layout buffer;
With two-level lists (the default without background) steps 1 and 4 count
and bin per 64-px block, and step 5 is gsmi_render_views_sh_two_level_record:
every 16-px tile filters its block's list while compositing and records what
it examined in its view's own array (view-relative int32 offsets from
gsmi_two_level_offsets_views), which step 6 reads
(gsmi_render_views_backward_sh_lists).
  7. gsmi_project_backward_views — per Gaussian, the views' gradients are
     summed in view order and chained to the raw parameters once.

Per view, every mask, record, tile list and pixel equals the single-view
pipeline's (rasterize.forward_view) bit for bit; gradients are the sum of the
per-view gradients (up to float summation order).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream_of
from .camera import CameraInfo

TILE = 16
VIEW_WORDS = 64  # gsmi_view is 64 4-byte words (include/gsmi.h)


def make_views(cams, tile_size=TILE):
    """gsmi_view records for [(c2w [3,4], CameraInfo), ...] as a [B, 64]
    float32 array (integer fields stored bit-wise), plus per-view constants.

    Every float is rounded to binary32 exactly as the single-view path passes
    it (camera.image_constants, frustum_planes_np), so the batch reproduces
    single-view results bit for bit."""
    B = len(cams)
    rec = np.zeros((B, VIEW_WORDS), np.float32)
    ints = rec.view(np.int32)
    meta = []
    tile_base = 0
    pix_base = 0
    for b, (c2w, cam) in enumerate(cams):
        c = np.ascontiguousarray(
            (c2w.detach().cpu().numpy() if isinstance(c2w, torch.Tensor) else np.asarray(c2w)), np.float32
        ).reshape(3, 4)
        normals, pts = cam.frustum_planes_np(c)
        ic = cam.image_constants(tile_size)
        rec[b, 0:12] = c.reshape(-1)
        rec[b, 12:30] = normals.reshape(-1)
        rec[b, 30:48] = pts.reshape(-1)
        rec[b, 48:52] = np.array([cam.fx, cam.fy, cam.cx, cam.cy], np.float64).astype(np.float32)
        rec[b, 52:56] = np.array([ic["topleft"][0], ic["topleft"][1], ic["pixel_size_x"], ic["pixel_size_y"]],
                                 np.float64).astype(np.float32)
        ints[b, 56:60] = [ic["W"], ic["H"], ic["n_tiles_w"], ic["n_tiles_h"]]
        ints[b, 60] = tile_base
        ints[b, 61] = 0
        rec[b, 62:64] = np.array([pix_base], np.int64).view(np.float32)
        n_tiles = ic["n_tiles_w"] * ic["n_tiles_h"]
        meta.append(dict(ic, c2w=c, tile_base=tile_base, pix_base=pix_base, n_tiles=n_tiles))
        tile_base += n_tiles
        pix_base += ic["H"] * ic["W"]
    return rec, meta, tile_base, pix_base


@dataclass
class ViewsBinning:
    N: int
    B: int
    V: int
    P: int
    bases: np.ndarray  # int64 [B+1, 2]: (first record, first pair) per view; last row = totals
    views_d: torch.Tensor
    meta: list
    n_tiles: int
    n_pix: int
    vis: torch.Tensor
    vis_off: torch.Tensor
    mean2d: torch.Tensor
    cov2d: torch.Tensor
    depth: torch.Tensor
    alpha: torch.Tensor
    tl: torch.Tensor
    br: torch.Tensor
    orig: torch.Tensor
    ids: torch.Tensor
    start: torch.Tensor
    end: torch.Tensor
    # two-level lists: pairs (P, bases) are per 64-px block; block_views_d,
    # n_blocks describe the block grids; after the forward ids / start / end
    # are the recorded per-tile lists, view b's at view_ids_d[b]
    shift: int = 0
    block_views_d: torch.Tensor | None = None
    n_blocks: int = 0
    view_ids_d: torch.Tensor | None = None
    view_fids: list | None = None


def block_view_records(rec, shift):
    """gsmi_view records of the views' block grids (2^shift tiles a side),
    blocks numbered across the batch."""
    brec = rec.copy()
    ints = brec.view(np.int32)
    base = 0
    for b in range(rec.shape[0]):
        ntw, nth = int(ints[b, 58]), int(ints[b, 59])
        bw, bh = (ntw + (1 << shift) - 1) >> shift, (nth + (1 << shift) - 1) >> shift
        ints[b, 58], ints[b, 59], ints[b, 60] = bw, bh, base
        base += bw * bh
    return brec, base


def preprocess_and_bin_views(mean, qvec, log_s, logit, cams, cfg, block_shift=0) -> ViewsBinning:
    """Steps 1-4 of the module docstring (per 64-px block with block_shift)."""
    L = _lib.load()
    dev = mean.device
    st = stream_of(dev)
    N = int(mean.shape[0])
    rec, meta, n_tiles, n_pix = make_views(cams, cfg.tile_size)
    B = len(cams)
    views_d = torch.from_numpy(rec).to(dev)
    vis = torch.empty(B * N, dtype=torch.uint8, device=dev)
    vis_off = torch.empty(B * N, dtype=torch.int32, device=dev)
    bases_d = torch.empty(2 * (B + 1), dtype=torch.int64, device=dev)
    ws = torch.empty(max(int(L.gsmi_preprocess_views_workspace_size(N, B)), 1), dtype=torch.uint8, device=dev)
    D = float(cfg.tile_culling_radius)
    check(L.gsmi_preprocess_views_count_blocks(N, B, ptr(views_d), ptr(mean), ptr(qvec), ptr(log_s), D,
                                               float(cfg.frustum_culling_radius),
                                               1 if cfg.skip_frustum_culling else 0, int(block_shift), ptr(vis),
                                               ptr(vis_off), ptr(bases_d), ptr(ws), ws.numel(), st),
          "preprocess_views_count")
    bases = bases_d.cpu().numpy().reshape(B + 1, 2)
    V, P = int(bases[B, 0]), int(bases[B, 1])
    if P >= 2 ** 31:
        raise RuntimeError(f"{P} (Gaussian, tile) pairs exceed the int32 range of the tile lists")
    f = dict(dtype=torch.float32, device=dev)
    mean2d = torch.empty(V, 2, **f)
    cov2d = torch.empty(V, 2, 2, **f)
    depth = torch.empty(V, 1, **f)
    alpha = torch.empty(V, **f)
    tl = torch.empty(V, 2, dtype=torch.int32, device=dev)
    br = torch.empty(V, 2, dtype=torch.int32, device=dev)
    orig = torch.empty(V, dtype=torch.int32, device=dev)
    check(L.gsmi_preprocess_views_emit(N, B, ptr(views_d), ptr(mean), ptr(qvec), ptr(log_s), ptr(logit), D,
                                       ptr(vis), ptr(vis_off), ptr(mean2d), ptr(cov2d), ptr(depth), ptr(alpha),
                                       ptr(tl), ptr(br), ptr(orig), st), "preprocess_views_emit")
    ids = torch.empty(P, dtype=torch.int32, device=dev)
    bin_views_d, n_bin, btl, bbr = views_d, n_tiles, tl, br
    bviews_d = None
    if block_shift:
        brec, n_bin = block_view_records(rec, block_shift)
        bviews_d = bin_views_d = torch.from_numpy(brec).to(dev)
        btl = torch.empty_like(tl)
        bbr = torch.empty_like(br)
        check(L.gsmi_shift_rects(V, ptr(tl), ptr(br), int(block_shift), ptr(btl), ptr(bbr), st), "shift_rects")
    start = torch.empty(n_bin, dtype=torch.int32, device=dev)
    end = torch.empty(n_bin, dtype=torch.int32, device=dev)
    bws = torch.empty(max(int(L.gsmi_binning_workspace_size(V, P, n_bin)), 1), dtype=torch.uint8, device=dev)
    check(L.gsmi_bin_views(V, P, B, ptr(bin_views_d), ptr(bases_d), n_bin, ptr(btl), ptr(bbr), ptr(depth), ptr(ids),
                           ptr(start), ptr(end), ptr(bws), bws.numel(), st), "bin_views")
    return ViewsBinning(N, B, V, P, bases, views_d, meta, n_tiles, n_pix, vis, vis_off, mean2d, cov2d, depth, alpha,
                        tl, br, orig, ids, start, end, int(block_shift), bviews_d, n_bin if block_shift else 0)


@dataclass
class ViewsState:
    b: ViewsBinning
    C: int
    cfg: object
    bg: torch.Tensor | None
    out: torch.Tensor
    T: torch.Tensor
    n_done: torch.Tensor
    pre_bg: torch.Tensor | None

    def image(self, v, flat=None):
        """[H, W, 3] view of camera v inside the flat image."""
        m = self.b.meta[v]
        src = self.out if flat is None else flat
        return src.view(-1, 3)[m["pix_base"]:m["pix_base"] + m["H"] * m["W"]].view(m["H"], m["W"], 3)


BLOCK_SHIFT = 2  # two-level lists: blocks of 4 x 4 tiles


def _two_level_views(b: ViewsBinning, sh_coeffs, C, cfg, out, T, nd, exact):
    """Two-level forward of a batch: per-tile offsets within each view, one
    list array per view (sized for the view's block pairs), the recording
    forward; afterwards b.ids / b.start / b.end / b.view_ids_d are the
    recorded per-tile lists."""
    L = _lib.load()
    dev = out.device
    st = stream_of(dev)
    view_tiles = np.array([m["n_tiles"] for m in b.meta], np.int32)
    foff = torch.empty(b.n_tiles, dtype=torch.int32, device=dev)
    fend = torch.empty(b.n_tiles, dtype=torch.int32, device=dev)
    ows = torch.empty(max(int(L.gsmi_two_level_offsets_workspace_size(int(view_tiles.max()))), 1), dtype=torch.uint8,
                      device=dev)
    check(L.gsmi_two_level_offsets_views(b.B, ptr(b.views_d), ptr(b.block_views_d),
                                         view_tiles.ctypes.data_as(ctypes.c_void_p), b.shift, ptr(b.start),
                                         ptr(b.end), ptr(foff), ptr(ows), ows.numel(), st), "two_level_offsets_views")
    per_view = np.diff(b.bases[:, 1])  # block pairs of each view
    caps = [int((1 << (2 * b.shift)) * p) for p in per_view]  # >= the view's sum of tile caps
    if max(caps) >= 2 ** 31:
        raise RuntimeError(f"a view needs {max(caps)} recorded list entries, beyond the int32 range")
    fids = [torch.empty(max(c, 1), dtype=torch.int32, device=dev) for c in caps]
    view_ids_d = torch.tensor([f.data_ptr() for f in fids], dtype=torch.int64, device=dev)
    cmax2 = sh_coeffs.shape[-1]
    check(L.gsmi_render_views_sh_two_level_record(
        b.V, b.B, ptr(b.views_d), ptr(b.block_views_d), b.n_tiles, ptr(b.mean2d), ptr(b.cov2d), ptr(sh_coeffs),
        3 * cmax2, cmax2, ptr(b.orig), ptr(b.alpha), ptr(b.tl), ptr(b.br), ptr(b.start), ptr(b.end), ptr(b.ids),
        b.shift, ptr(foff), ptr(view_ids_d), ptr(fend), ptr(out), int(C), float(cfg.T_thresh), ptr(T), ptr(nd),
        1 if exact else 0, st), "render_views_sh_two_level_record")
    b.start, b.end, b.view_fids, b.view_ids_d = foff, fend, fids, view_ids_d
    b.ids = None


def forward_views(mean, qvec, log_s, logit, sh_coeffs, cams, C, cfg, exact=False):
    """Steps 1-5: returns (flat image [sum(H*W)*3], ViewsState)."""
    L = _lib.load()
    two = bool(getattr(cfg, "two_level", False)) and not cfg.bg and cfg.tile_size == TILE
    b = preprocess_and_bin_views(mean, qvec, log_s, logit, cams, cfg, block_shift=BLOCK_SHIFT if two else 0)
    dev = mean.device
    bg = torch.tensor(cfg.bg_rgb, dtype=torch.float32, device=dev) if cfg.bg else None
    out = torch.empty(b.n_pix * 3, dtype=torch.float32, device=dev)
    T = torch.empty(b.n_pix, dtype=torch.float32, device=dev)
    nd = torch.empty(b.n_pix, dtype=torch.int32, device=dev)
    pre = torch.empty(b.n_pix * 3, dtype=torch.float32, device=dev) if cfg.bg else None
    cmax2 = sh_coeffs.shape[-1]
    if not 1 <= C <= 4 or C * C > cmax2:
        raise RuntimeError(f"SH order C={C} does not fit sh_coeffs with {cmax2} coefficients per channel")
    if two:
        if b.V > 0:
            _two_level_views(b, sh_coeffs, C, cfg, out, T, nd, exact)
        else:  # nothing visible: every pixel of every view is empty
            out.zero_()
            T.fill_(1.0)
            nd.zero_()
        return out, ViewsState(b, int(C), cfg, bg, out, T, nd, pre)
    check(L.gsmi_render_views_sh(b.V, b.P, b.B, ptr(b.views_d), b.n_tiles, ptr(b.mean2d), ptr(b.cov2d),
                                 ptr(sh_coeffs), 3 * cmax2, cmax2, ptr(b.orig), ptr(b.alpha), ptr(b.start),
                                 ptr(b.end), ptr(b.ids), ptr(out), int(C), float(cfg.T_thresh), ptr(bg), ptr(T),
                                 ptr(nd), ptr(pre), 1, 1 if exact else 0, stream_of(dev)), "render_views_sh")
    return out, ViewsState(b, int(C), cfg, bg, out, T, nd, pre)


def backward_views(s: ViewsState, mean, qvec, log_s, logit, sh_coeffs, grad_out, boundary=False):
    """Steps 6-7: gradients w.r.t. the raw parameters summed over the views.
    Returns (g_mean, g_qvec, g_log_s, g_logit, g_sh); with boundary=True also
    a dict of the compositing gradients of the records (mean2d [V,2], cov
    [V,2,2], alpha [V], view-major as the records)."""
    L = _lib.load()
    b = s.b
    dev = mean.device
    st = stream_of(dev)
    g_m2d = torch.zeros(b.V, 2, dtype=torch.float32, device=dev)
    g_cov = torch.zeros(b.V, 2, 2, dtype=torch.float32, device=dev)
    g_al = torch.zeros(b.V, dtype=torch.float32, device=dev)
    g_sh = torch.zeros_like(sh_coeffs)
    cmax2 = sh_coeffs.shape[-1]
    go = grad_out.contiguous().view(-1)
    if go.numel() != b.n_pix * 3:
        raise RuntimeError(f"grad_out has {go.numel()} values, expected {b.n_pix * 3}")
    bd = dict(mean2d=g_m2d, cov=g_cov, alpha=g_al)
    if b.V == 0:  # no view sees a Gaussian: every gradient is zero (no launch takes empty buffers)
        g = (torch.zeros_like(mean), torch.zeros_like(qvec), torch.zeros_like(log_s), torch.zeros_like(logit), g_sh)
        return (*g, bd) if boundary else g
    nws = int(L.gsmi_backward_order_workspace_size(b.n_tiles))  # launch order, most expensive tiles first
    ows = torch.empty(max(nws, 1), dtype=torch.uint8, device=dev)
    if b.view_ids_d is not None:  # two-level: the recorded per-view lists
        check(L.gsmi_render_views_backward_sh_lists(
            b.V, b.P, b.B, ptr(b.views_d), b.n_tiles, ptr(b.mean2d), ptr(b.cov2d), ptr(sh_coeffs), 3 * cmax2, cmax2,
            ptr(b.orig), ptr(b.alpha), ptr(b.start), ptr(b.end), ptr(b.view_ids_d), ptr(s.T), ptr(s.n_done), ptr(go),
            ptr(g_m2d), ptr(g_cov), ptr(g_sh), ptr(g_al), s.C, float(s.cfg.T_thresh), 0, ptr(ows), nws, st),
            "render_views_backward_sh_lists")
    else:
        check(L.gsmi_render_views_backward_sh(b.V, b.P, b.B, ptr(b.views_d), b.n_tiles, ptr(b.mean2d), ptr(b.cov2d),
                                          ptr(sh_coeffs), 3 * cmax2, cmax2, ptr(b.orig), ptr(b.alpha), ptr(b.start),
                                          ptr(b.end), ptr(b.ids), ptr(s.out), ptr(s.T), ptr(s.n_done),
                                          ptr(s.pre_bg), ptr(go), ptr(g_m2d), ptr(g_cov), ptr(g_sh), ptr(g_al), s.C,
                                          float(s.cfg.T_thresh), ptr(s.bg), 0, ptr(ows), nws, st),
          "render_views_backward_sh")
    g_mean = torch.empty_like(mean)
    g_q = torch.empty_like(qvec)
    g_ls = torch.empty_like(log_s)
    g_lg = torch.empty_like(logit)
    check(L.gsmi_project_backward_views(b.N, b.B, ptr(b.views_d), ptr(mean), ptr(qvec), ptr(log_s), ptr(logit),
                                        ptr(b.vis), ptr(b.vis_off), ptr(g_m2d), ptr(g_cov), ptr(g_al), ptr(g_mean),
                                        ptr(g_q), ptr(g_ls), ptr(g_lg), st), "project_backward_views")
    g = (g_mean, g_q, g_ls, g_lg, g_sh)
    return (*g, bd) if boundary else g


class _RasterizeSHViews(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mean, qvec, log_s, logit, sh_coeffs, cams, C, cfg):
        mean, qvec, log_s, logit, sh_coeffs = (t.contiguous() for t in (mean, qvec, log_s, logit, sh_coeffs))
        out, state = forward_views(mean, qvec, log_s, logit, sh_coeffs, cams, C, cfg)
        ctx.state = state
        ctx.save_for_backward(mean, qvec, log_s, logit, sh_coeffs)
        return out

    @staticmethod
    def backward(ctx, grad):
        mean, qvec, log_s, logit, sh_coeffs = ctx.saved_tensors
        g = backward_views(ctx.state, mean, qvec, log_s, logit, sh_coeffs, grad)
        return (*g, None, None, None)


def rasterize_sh_views(mean, qvec, svec_before_activation, alpha_before_activation, sh_coeffs, cams, C, cfg):
    """Images of B cameras ([(c2w, CameraInfo), ...], mixed resolutions
    allowed) as a list of [H_b, W_b, 3] views of one flat buffer;
    differentiable w.r.t. the five raw parameter tensors."""
    flat = _RasterizeSHViews.apply(mean, qvec, svec_before_activation, alpha_before_activation, sh_coeffs, cams,
                                   int(C), cfg)
    out = []
    base = 0
    for _, cam in cams:
        n = cam.h * cam.w
        out.append(flat.view(-1, 3)[base:base + n].view(cam.h, cam.w, 3))
        base += n
    return out
