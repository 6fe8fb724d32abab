"""Training-step stages around the rasteriser: loss seeding and the optimizer.

* ``photometric_loss`` — the reference's loss (utils/loss.py:5-22): the base
  term (``l1_loss`` / ``mse_loss``, reduction "mean") and its gradient w.r.t.
  the image in one HBM pass (gsmi_photometric_loss), and with ``ssim_mult``
  the SSIM term that the reference takes from kornia (``ssim_loss``, window
  ``ssim_win``), fused with the base term and its gradient
  (gsmi_photometric_ssim_loss).
* ``FusedAdam`` — ``torch.optim.Adam`` over the five SHRenderer parameter
  groups (gs/sh_renderer.py:710-729, betas (0.9, 0.99)) as one launch
  (gsmi_adam_step).  ``fresh=True`` reproduces what the reference actually
  runs: it rebuilds the optimizer every iteration (main_sh.py:238), so every
  update is a first Adam step from zero moments.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream_of

GROUPS = ("mean", "qvec", "svec", "sh_coeffs", "alpha")


def photometric_loss(out, gt, kind="l1", want_grad=True, ssim_mult=0.0, ssim_win=11):
    """(loss scalar tensor, d loss / d out or None).

    get_loss_fn (utils/loss.py:5-22): ssim_mult * ssim_loss(out, gt, ssim_win)
    + (1 - ssim_mult) * base, base = l1_loss or mse_loss (mean).  With
    ssim_mult > 0 the images are [H, W, C] (the rasteriser's HWC layout) and
    the gradient is always computed."""
    L = _lib.load()
    if out.shape != gt.shape:
        raise RuntimeError(f"image shapes differ: {tuple(out.shape)} vs {tuple(gt.shape)}")
    if kind not in ("l1", "l2"):
        raise NotImplementedError(kind)
    o = out.contiguous().float()
    g_t = gt.contiguous().float()
    loss = torch.empty((), dtype=torch.float32, device=o.device)
    if ssim_mult:
        if o.dim() != 3:
            raise RuntimeError("the SSIM term needs images of shape [H, W, C]")
        H, W, CH = (int(x) for x in o.shape)
        grad = torch.empty_like(o)
        ws = torch.empty(int(L.gsmi_ssim_loss_workspace_size(H, W, CH)), dtype=torch.uint8, device=o.device)
        check(L.gsmi_photometric_ssim_loss(H, W, CH, ptr(o), ptr(g_t), int(ssim_win), float(ssim_mult),
                                           1 if kind == "l2" else 0, ptr(grad), ptr(loss), ptr(ws), ws.numel(),
                                           stream_of(o.device)), "photometric_ssim_loss")
        return loss, grad
    n = o.numel()
    grad = torch.empty_like(o) if want_grad else None
    ws = torch.empty(int(L.gsmi_loss_workspace_size(n)), dtype=torch.uint8, device=o.device)
    check(L.gsmi_photometric_loss(n, ptr(o), ptr(g_t), 1 if kind == "l2" else 0, ptr(grad), ptr(loss), ptr(ws),
                                  ws.numel(), stream_of(o.device)), "photometric_loss")
    return loss, grad


class FusedAdam:
    """Adam over named parameter groups in one launch per step.

    groups: list of (tensor, lr); tensors are updated in place from their
    ``.grad`` (or from grads passed to ``step``)."""

    def __init__(self, groups, betas=(0.9, 0.99), eps=1e-8, fresh=False):
        if not 1 <= len(groups) <= 8:
            raise ValueError("between 1 and 8 parameter groups")
        self.params = [p for p, _ in groups]
        self.lr = [float(lr) for _, lr in groups]
        self.betas = (float(betas[0]), float(betas[1]))
        self.eps = float(eps)
        self.fresh = bool(fresh)
        self.t = 0
        for p in self.params:
            if p.dtype != torch.float32 or not p.is_contiguous():
                raise RuntimeError("FusedAdam needs contiguous float32 parameters")
        self.m = None if fresh else [torch.zeros_like(p) for p in self.params]
        self.v = None if fresh else [torch.zeros_like(p) for p in self.params]

    @torch.no_grad()
    def step(self, grads=None):
        L = _lib.load()
        grads = grads if grads is not None else [p.grad for p in self.params]
        grads = [g.contiguous().float() for g in grads]
        k = len(self.params)
        self.t = 1 if self.fresh else self.t + 1
        arr = lambda ts: (ctypes.c_void_p * k)(*[t.data_ptr() for t in ts])
        numel = (ctypes.c_int64 * k)(*[p.numel() for p in self.params])
        lrs = (ctypes.c_double * k)(*self.lr)
        check(L.gsmi_adam_step(k, arr(self.params), arr(grads), None if self.fresh else arr(self.m),
                               None if self.fresh else arr(self.v), numel, lrs, self.t, self.betas[0], self.betas[1],
                               self.eps, 1 if self.fresh else 0, stream_of(self.params[0].device)), "adam_step")
