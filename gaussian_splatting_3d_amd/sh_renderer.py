"""SHRenderer: the reference model's rasterisation path on gfx950.

Mirrors gs/sh_renderer.py:43-324 for what the hot path needs: the same
parameters (``mean``, ``qvec``, ``svec_before_activation``, ``sh_coeffs``,
``alpha_before_activation``), the same activations (exp scales, sigmoid
opacity), the same config keys, ``now_C`` (active SH order) and
``forward(c2w, camera_info) -> [H, W, 3]``.  ``forward`` runs the fused
pipeline of ``rasterize.py``; gradients reach the five parameters exactly as
through the reference's autograd graph.  Training-side machinery
(densification, logging, schedulers, optimiser) is not part of this module.
"""
from __future__ import annotations

import numpy as np
import torch

from .rasterize import rasterize_sh
from .scenes import RasterConfig

sh_base = 0.28209479177387814


def _get(cfg, key, default=None):
    if cfg is None:
        return default
    if isinstance(cfg, dict):
        return cfg.get(key, default)
    if hasattr(cfg, "get"):
        try:
            return cfg.get(key, default)
        except TypeError:
            pass
    return getattr(cfg, key, default)


class SHRenderer(torch.nn.Module):
    def __init__(self, cfg, pts=None, rgb=None):
        super().__init__()
        self.cfg = cfg
        self.device = _get(cfg, "device", "cuda")
        self.max_C = int(_get(cfg, "sh_order", 1))
        if pts is not None and rgb is not None:
            self.initialize(cfg, pts, rgb)
        self.now_C = 1
        self.set_cfg(cfg)

    def initialize(self, cfg, pts, rgb):
        """gs/sh_renderer.py:79-102 with svec_init_method == "fixed"."""
        self.N = pts.shape[0]
        self.mean = torch.nn.Parameter(pts)
        q = torch.zeros([self.N, 4])
        q[..., 0] = 1.0
        self.qvec = torch.nn.Parameter(q)
        sh = torch.zeros((self.N, 3, self.max_C * self.max_C), device=rgb.device)
        sh[:, :, 0] = torch.logit(rgb) / sh_base
        self.sh_coeffs = torch.nn.Parameter(sh)
        self.svec_before_activation = torch.nn.Parameter(
            torch.full([self.N, 3], float(np.log(_get(cfg, "svec_init", 0.01)))))
        self.alpha_before_activation = torch.nn.Parameter(
            torch.full([self.N], float(torch.logit(torch.tensor(_get(cfg, "alpha_init", 0.1))))))

    @classmethod
    def from_params(cls, mean, qvec, svec_before_activation, alpha_before_activation, sh_coeffs, cfg=None):
        cfg = cfg if cfg is not None else {}
        m = cls.__new__(cls)
        torch.nn.Module.__init__(m)
        m.cfg = cfg
        m.device = mean.device
        m.N = mean.shape[0]
        m.mean = torch.nn.Parameter(mean)
        m.qvec = torch.nn.Parameter(qvec)
        m.svec_before_activation = torch.nn.Parameter(svec_before_activation)
        m.alpha_before_activation = torch.nn.Parameter(alpha_before_activation)
        m.sh_coeffs = torch.nn.Parameter(sh_coeffs)
        C = int(round(sh_coeffs.shape[-1] ** 0.5))
        m.max_C = C
        m.now_C = C
        m.set_cfg(cfg)
        return m

    def set_cfg(self, cfg):
        """gs/sh_renderer.py:104-151 (rendering keys)."""
        d = RasterConfig()
        self.tile_size = int(_get(cfg, "tile_size", d.tile_size))
        self.frustum_culling_radius = float(_get(cfg, "frustum_culling_radius", d.frustum_culling_radius))
        self.tile_culling_radius = float(_get(cfg, "tile_culling_radius", d.tile_culling_radius))
        self.T_thresh = float(_get(cfg, "T_thresh", d.T_thresh))
        self.depth_detach = bool(_get(cfg, "depth_detach", True))
        if not self.depth_detach:
            raise NotImplementedError("depth_detach=False: the projection adjoint implements the detached depth")
        self.bg = bool(_get(cfg, "bg", False))
        self.bg_rgb = tuple(_get(cfg, "bg_rgb", [1.0, 1.0, 1.0])) if self.bg else (1.0, 1.0, 1.0)
        self.skip_frustum_culling = bool(_get(cfg, "skip_frustum_culling", False))
        self.raster_cfg = RasterConfig(self.tile_size, self.frustum_culling_radius, self.tile_culling_radius,
                                       self.T_thresh, self.skip_frustum_culling, self.bg, self.bg_rgb)

    @property
    def svec(self):
        return torch.exp(self.svec_before_activation)

    @property
    def alpha(self):
        return torch.sigmoid(self.alpha_before_activation)

    def forward(self, c2w, camera_info):
        """gs/sh_renderer.py:188-316: image [H, W, 3]."""
        return rasterize_sh(self.mean, self.qvec, self.svec_before_activation, self.alpha_before_activation,
                            self.sh_coeffs, c2w, camera_info, self.now_C, self.raster_cfg)
