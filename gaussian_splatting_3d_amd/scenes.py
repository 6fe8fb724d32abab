"""Seeded synthetic scenes and cameras for the benchmark workloads.

Parameters follow the reference model's storage convention
(gs/sh_renderer.py:79-102): ``mean`` [N,3], ``qvec`` [N,4] (w,x,y,z),
``svec_before_activation`` [N,3] (log scale), ``alpha_before_activation`` [N]
(opacity logit), ``sh_coeffs`` [N,3,C*C].  Everything is drawn on the CPU from
``torch.Generator().manual_seed(seed)`` so the same seed gives the same scene
on every host.  Knobs not fixed by the workload definitions come from the
reference's conf/lego.yaml: tile 16, frustum radius 1.0, tile-culling radius
D = 6.0, T_thresh 1e-4.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from .camera import CameraInfo, get_c2w_from_up_and_look_at


@dataclass
class RasterConfig:
    """Renderer knobs, named as the reference's config keys."""

    tile_size: int = 16
    frustum_culling_radius: float = 1.0
    tile_culling_radius: float = 6.0
    T_thresh: float = 1e-4
    skip_frustum_culling: bool = False
    bg: bool = False
    bg_rgb: tuple = (1.0, 1.0, 1.0)
    # bin per 64-px block, filter to 16-px tiles inside the forward (same
    # image, T and n_done); no counterpart in the reference, ignored with bg
    two_level: bool = True


@dataclass
class Scene:
    mean: torch.Tensor
    qvec: torch.Tensor
    svec_before_activation: torch.Tensor
    alpha_before_activation: torch.Tensor
    sh_coeffs: torch.Tensor
    C: int
    cameras: list = field(default_factory=list)  # list of (c2w [3,4] float32, CameraInfo)
    name: str = ""

    @property
    def N(self):
        return self.mean.shape[0]

    def to(self, device):
        for k in ("mean", "qvec", "svec_before_activation", "alpha_before_activation", "sh_coeffs"):
            setattr(self, k, getattr(self, k).to(device))
        return self


def _normalised_quats(g, n):
    q = torch.randn(n, 4, generator=g)
    return q / q.norm(dim=-1, keepdim=True)


def orbit_camera(radius, theta, fx, fy, W, H, near=0.01, far=1000.0, height=0.0):
    pos = np.array([radius * math.cos(theta), radius * math.sin(theta), height])
    c2w = get_c2w_from_up_and_look_at(np.array([0.0, 0.0, 1.0]), np.zeros(3), pos)
    cam = CameraInfo(fx, fy, W / 2, H / 2, W, H, near, far)
    return torch.from_numpy(c2w), cam


def plain_forward_10k(seed=1, n=10_000):
    """10k Gaussians, means U[-1,1]^3, log-scales N(-4,0.5), SH degree 0,
    camera at (0,0,-3) looking at the origin, fx=fy=300, 256x256."""
    g = torch.Generator().manual_seed(seed)
    mean = torch.rand(n, 3, generator=g) * 2 - 1
    log_s = torch.randn(n, 3, generator=g) * 0.5 - 4.0
    qvec = _normalised_quats(g, n)
    logit = torch.randn(n, generator=g)
    sh = torch.randn(n, 3, 1, generator=g) * 0.3
    c2w = torch.from_numpy(get_c2w_from_up_and_look_at(np.array([0.0, 1.0, 0.0]), np.zeros(3),
                                                       np.array([0.0, 0.0, -3.0])))
    cam = CameraInfo(300.0, 300.0, 128.0, 128.0, 256, 256, 0.01, 5.0)
    return Scene(mean, qvec, log_s, logit, sh, 1, [(c2w, cam)], "plain_forward_10k_256x256_sh0")


def _sh3(g, n):
    sh = torch.randn(n, 3, 16, generator=g) * 0.05
    sh[:, :, 0] = torch.randn(n, 3, generator=g) * 0.5
    return sh


def fwd_bwd_500k(seed=2, n=500_000, W=1297, H=840):
    """500k Gaussians, means N(0, 2^2 I), log-scales N(-3.5,0.7), SH degree 3,
    one camera on a radius-6 orbit, fx=fy=1100, 1297x840."""
    g = torch.Generator().manual_seed(seed)
    mean = torch.randn(n, 3, generator=g) * 2.0
    log_s = torch.randn(n, 3, generator=g) * 0.7 - 3.5
    qvec = _normalised_quats(g, n)
    logit = torch.randn(n, generator=g) - 1.0
    sh = _sh3(g, n)
    theta = float(torch.rand(1, generator=g)) * 2 * math.pi
    return Scene(mean, qvec, log_s, logit, sh, 4, [orbit_camera(6.0, theta, 1100.0, 1100.0, W, H)],
                 f"fwd_bwd_{n // 1000}k_{W}x{H}_sh3")


def dense_3m(seed=3, n=3_000_000, W=1920, H=1080):
    """3M Gaussians in a 10x10x10 box, 20% in 50 blobs (sigma 0.3), log-scales
    N(-3,1) clipped at -0.5, SH degree 3, 1920x1080, radius-15 orbit camera,
    fx=fy=1600."""
    g = torch.Generator().manual_seed(seed)
    n_blob = n // 5
    n_uni = n - n_blob
    uni = torch.rand(n_uni, 3, generator=g) * 10 - 5
    centres = torch.rand(50, 3, generator=g) * 10 - 5
    which = torch.randint(0, 50, (n_blob,), generator=g)
    blob = centres[which] + torch.randn(n_blob, 3, generator=g) * 0.3
    mean = torch.cat([uni, blob], 0)
    log_s = (torch.randn(n, 3, generator=g) - 3.0).clamp(max=-0.5)
    qvec = _normalised_quats(g, n)
    logit = torch.randn(n, generator=g) - 1.0
    sh = _sh3(g, n)
    theta = float(torch.rand(1, generator=g)) * 2 * math.pi
    return Scene(mean, qvec, log_s, logit, sh, 4, [orbit_camera(15.0, theta, 1600.0, 1600.0, W, H)],
                 f"dense_{n // 1000}k_{W}x{H}_sh3")


def serving_scene(n, seed=4):
    """Resident serving scene built as the 500k forward+backward scene."""
    s = fwd_bwd_500k(seed=seed, n=n)
    s.cameras = []
    s.name = f"serve_{n // 1000}k_sh3"
    return s


def multiview_6m(seed=5, n=6_000_000, W=3840, H=2160, views=8):
    """6M Gaussians, means N(0, 4^2 I), log-scales N(-3.5,0.8), SH degree 3,
    8 cameras equally spaced on a radius-8 orbit at 3840x2160, fx=fy=3300."""
    g = torch.Generator().manual_seed(seed)
    mean = torch.randn(n, 3, generator=g) * 4.0
    log_s = torch.randn(n, 3, generator=g) * 0.8 - 3.5
    qvec = _normalised_quats(g, n)
    logit = torch.randn(n, generator=g) - 1.0
    sh = _sh3(g, n)
    cams = [orbit_camera(8.0, 2 * math.pi * k / views, 3300.0, 3300.0, W, H) for k in range(views)]
    return Scene(mean, qvec, log_s, logit, sh, 4, cams, f"multiview_{n // 1000}k_{views}x{W}x{H}_sh3")


def target_image(H, W, seed):
    """Seeded uniform[0,1] target image [H,W,3] for the L1 loss."""
    g = torch.Generator().manual_seed(seed + 100)
    return torch.rand(H, W, 3, generator=g)
