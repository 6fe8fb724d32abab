"""Adaptive density control of SHRenderer on the GPU (gs/sh_renderer.py:423-661).

``DensityControl`` keeps the per-Gaussian statistics the reference keeps
(``grad_mean`` accumulator and the visibility count ``cnt``) and performs
split/clone, low-opacity pruning and opacity reset as stream compactions in
the native library (csrc/densify.hip).  The parameters are the five raw
tensors in the reference's storage convention; operations return new
tensors (N changes), like the reference which rebuilds its Parameters.

Semantics kept from the reference, including its quirks: the "mean"
reduction divides by ``cnt + 1e-5``; split Gaussians are sampled with the
transposed rotation (``einsum("bij,bj->bi", R^T, noise * s)``, :469-474),
and the two copies are laid out as ``.repeat(2, 1)`` does (all split rows,
then all again); clone sources stay in place and the clones follow the
unsplit block.  The split noise is drawn with ``torch.randn`` (on the
device, optionally from a given generator) and handed to the kernel.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import check, ptr, stream_of


@dataclass
class DensifyConfig:
    """Defaults: conf/garden_sh.yaml:52-58 (split type and reduction from the
    SHRenderer config keys of the same names)."""

    pos_grad_thresh: float = 2e-4
    split_scale_thresh: float = 0.02
    scale_shrink_factor: float = 1.6
    split_type: str = "mean_grad"      # or "2d_mean_grad"
    split_reduction: str = "mean"      # or "max"
    alpha_thresh: float = 0.05
    alpha_reset_val: float = 0.1


class DensityControl:
    def __init__(self, N, device="cuda:0", cfg: DensifyConfig | None = None):
        self.cfg = cfg or DensifyConfig()
        if self.cfg.split_type not in ("mean_grad", "2d_mean_grad"):
            raise NotImplementedError(self.cfg.split_type)
        if self.cfg.split_reduction not in ("mean", "max"):
            raise NotImplementedError(self.cfg.split_reduction)
        self.device = torch.device(device)
        self.reset_stats(N)

    def reset_stats(self, N):
        self.grad_mean = torch.zeros(N, dtype=torch.float32, device=self.device)
        self.cnt = torch.zeros(N, dtype=torch.int32, device=self.device)

    def count_visible(self, vis):
        """cnt[vis] += 1 (sh_renderer.py:215-216); vis: uint8/bool [N]."""
        v = vis.to(torch.uint8).contiguous()
        check(_lib.load().gsmi_count_visible(v.numel(), ptr(v), ptr(self.cnt), stream_of(self.device)),
              "count_visible")

    def update_grads(self, grad, index=None):
        """grad: dL/dmean [N,3] ("mean_grad") or the compacted dL/dmean2d
        [V,2] with index = original Gaussian of each row ("2d_mean_grad")."""
        g = grad.contiguous().float()
        dim = 3 if self.cfg.split_type == "mean_grad" else 2
        if g.shape[-1] != dim:
            raise RuntimeError(f"expected a [*, {dim}] gradient for split_type {self.cfg.split_type}")
        idx = None if index is None else index.to(torch.int32).contiguous()
        check(_lib.load().gsmi_grad_stats(g.shape[0], ptr(g), dim, ptr(idx),
                                          1 if self.cfg.split_reduction == "max" else 0, ptr(self.grad_mean),
                                          stream_of(self.device)), "grad_stats")

    @torch.no_grad()
    def split_gaussians(self, mean, qvec, log_s, logit, sh, generator=None, noise=None):
        """Returns the new five tensors and (n_unsplit, n_clone, n_split)."""
        L = _lib.load()
        N = int(mean.shape[0])
        st = stream_of(self.device)
        ws = torch.empty(max(int(L.gsmi_densify_workspace_size(N)), 1), dtype=torch.uint8, device=self.device)
        counts = (ctypes.c_int64 * 3)()
        check(L.gsmi_split_clone_count(N, ptr(self.grad_mean), ptr(self.cnt),
                                       1 if self.cfg.split_reduction == "mean" else 0,
                                       float(self.cfg.pos_grad_thresh), ptr(log_s),
                                       float(self.cfg.split_scale_thresh), counts, ptr(ws), ws.numel(), st),
              "split_clone_count")
        n_unsplit, n_clone, k = (int(x) for x in counts)
        if noise is None:
            noise = torch.randn(2 * k, 3, device=self.device, generator=generator)
        noise = noise.contiguous().float()
        if noise.shape != (2 * k, 3):
            raise RuntimeError(f"noise must be [{2 * k}, 3]")
        M = n_unsplit + n_clone + 2 * k
        sh_row = int(sh[0].numel()) if N else int(math.prod(sh.shape[1:]))
        out = [torch.empty((M,) + tuple(t.shape[1:]), dtype=torch.float32, device=self.device)
               for t in (mean, qvec, log_s, logit, sh)]
        check(L.gsmi_split_clone_write(N, ptr(mean), ptr(qvec), ptr(log_s), ptr(logit), ptr(sh), sh_row, ptr(noise),
                                       float(self.cfg.scale_shrink_factor), *[ptr(t) for t in out], ptr(ws),
                                       ws.numel(), st), "split_clone_write")
        self.reset_stats(M)
        return out, (n_unsplit, n_clone, k)

    @torch.no_grad()
    def remove_low_alpha(self, mean, qvec, log_s, logit, sh):
        L = _lib.load()
        N = int(mean.shape[0])
        st = stream_of(self.device)
        ws = torch.empty(max(int(L.gsmi_densify_workspace_size(N)), 1), dtype=torch.uint8, device=self.device)
        kept = ctypes.c_int64()
        check(L.gsmi_prune_alpha_count(N, ptr(logit), float(self.cfg.alpha_thresh), ctypes.byref(kept), ptr(ws),
                                       ws.numel(), st), "prune_alpha_count")
        M = int(kept.value)
        sh_row = int(math.prod(sh.shape[1:]))
        out = [torch.empty((M,) + tuple(t.shape[1:]), dtype=torch.float32, device=self.device)
               for t in (mean, qvec, log_s, logit, sh)]
        check(L.gsmi_prune_alpha_write(N, ptr(mean), ptr(qvec), ptr(log_s), ptr(logit), ptr(sh), sh_row,
                                       *[ptr(t) for t in out], ptr(ws), ws.numel(), st), "prune_alpha_write")
        self.reset_stats(M)
        return out

    @torch.no_grad()
    def reset_alpha(self, logit):
        """alpha_before_activation <- logit(alpha_reset_val) (sh_renderer.py:580-585)."""
        v = self.cfg.alpha_reset_val
        logit.fill_(math.log(v / (1.0 - v)))
