"""Tile bounds and pair count: drop-in for gs/culling.py:8-37.

``tile_culling_aabb_count(mean, cov, tile_size, camera_info, D)`` returns
``(N_with_dub, aabb_topleft, aabb_bottomright)`` exactly as the reference:
a Python int and two int32 [N,2] tensors of tile coordinates, computed on the
GPU by gsmi_tile_culling_aabb_count with the reference's float32 operation
order (half extent sqrt(D * cov_ii), pixel = int32(x * f + c) with truncation,
clamp to the image, floor-divide by the tile size).  Reading the count back
is the one host synchronisation, as in the reference (.item()).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream_of
from .backend import check_float


@torch.no_grad()
def tile_culling_aabb_count(mean, cov, tile_size, camera_info, D):
    check_float("mean", mean)
    cov = cov.contiguous()
    check_float("cov", cov)
    L = _lib.load()
    V = mean.shape[0]
    dev = mean.device
    tl = torch.empty(V, 2, dtype=torch.int32, device=dev)
    br = torch.empty(V, 2, dtype=torch.int32, device=dev)
    cnt = torch.empty(1, dtype=torch.int64, device=dev)
    ws = torch.empty(max(int(L.gsmi_aabb_count_workspace_size(V)), 1), dtype=torch.uint8, device=dev)
    f = np.float32
    check(L.gsmi_tile_culling_aabb_count(V, ptr(mean), ptr(cov), float(f(camera_info.fx)), float(f(camera_info.fy)),
                                         float(f(camera_info.cx)), float(f(camera_info.cy)), int(camera_info.w),
                                         int(camera_info.h), float(f(D)), int(tile_size), ptr(tl), ptr(br), ptr(cnt),
                                         ptr(ws), ws.numel(), stream_of(dev)), "tile_culling_aabb_count")
    return int(cnt.item()), tl, br
