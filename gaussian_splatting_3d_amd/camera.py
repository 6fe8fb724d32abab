"""Pinhole camera description and frustum planes.

Mirrors ``CameraInfo`` of utils/camera.py:219-314 (same constructor arguments,
attributes and ``get_frustum`` / ``camera_space_to_pixel_space`` /
``from_fov_camera`` behaviour) and ``get_c2w_from_up_and_look_at`` of
gs/debug.py:23-37.  OpenCV axes: x right, y down, z forward.
"""
from __future__ import annotations

import numpy as np
import torch

_f32 = np.float32


def _cross3(a, b):
    return np.array(
        [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], _f32
    )


class CameraInfo:
    """utils/camera.py:219-314."""

    def __init__(self, fx, fy, cx, cy, w, h, near_plane, far_plane) -> None:
        self.fx, self.fy = fx, fy
        self.cx, self.cy = cx, cy
        self.w, self.h = w, h
        self.yfov = 2 * np.arctan(self.h / (2 * self.fy))
        self.aspect = w / h
        self.near_plane = near_plane
        self.far_plane = far_plane

    def downsample(self, scale):
        self.fx /= scale
        self.fy /= scale
        self.cx /= scale
        self.cy /= scale
        self.w //= scale
        self.h //= scale
        self.yfov = 2 * np.arctan(self.h / (2 * self.fy))
        self.aspect = self.w / self.h

    def upsample(self, scale):
        self.fx *= scale
        self.fy *= scale
        self.cx *= scale
        self.cy *= scale
        self.w *= int(scale)
        self.h *= int(scale)

    def frustum_planes_np(self, c2w):
        """Six plane normals and points as float32 numpy arrays [6,3].

        utils/camera.py:249-283: right = c2w[:,0], up = -c2w[:,1], lookat =
        c2w[:,2]; near/far planes along lookat, four side planes through the
        eye; normals normalised with max(|n|, 1e-12).  Evaluated in float32,
        one rounding per operation.
        """
        c = np.ascontiguousarray(
            (c2w.detach().cpu().numpy() if isinstance(c2w, torch.Tensor) else np.asarray(c2w)), _f32
        ).reshape(3, 4)
        up = -c[:, 1]
        right = c[:, 0]
        lookat = c[:, 2]
        t = c[:, 3]
        half_vside = self.far_plane * np.tan(self.yfov * 0.5)
        half_hside = half_vside * self.aspect
        hv, hh = _f32(half_vside), _f32(half_hside)
        near_point = _f32(self.near_plane) * lookat
        far_point = _f32(self.far_plane) * lookat
        normals = np.stack(
            [
                lookat,
                -lookat,
                _cross3(far_point - hh * right, up),
                _cross3(up, far_point + hh * right),
                _cross3(far_point + hv * up, right),
                _cross3(right, far_point - hv * up),
            ]
        ).astype(_f32)
        pts = np.stack([near_point + t, far_point + t, t, t, t, t]).astype(_f32)
        nrm = np.sqrt(normals[:, 0] * normals[:, 0] + normals[:, 1] * normals[:, 1] + normals[:, 2] * normals[:, 2])
        normals = (normals / np.maximum(nrm, _f32(1e-12))[:, None]).astype(_f32)
        return normals, pts

    def get_frustum(self, c2w):
        """Returns (normals [6,3], pts [6,3]) as float32 tensors on c2w's device."""
        n, p = self.frustum_planes_np(c2w)
        dev = c2w.device if isinstance(c2w, torch.Tensor) else "cpu"
        return torch.from_numpy(n).to(dev), torch.from_numpy(p).to(dev)

    def camera_space_to_pixel_space(self, pts):
        """utils/camera.py:290-303 (float32 multiply then add, int32 cast)."""
        if pts.shape[1] == 3:
            pts = pts[:, :2] / pts[:, 2:]
        assert pts.shape[1] == 2
        pts[:, 0] = pts[:, 0] * self.fx + self.cx
        pts[:, 1] = pts[:, 1] * self.fy + self.cy
        if isinstance(pts, np.ndarray):
            pts = pts.astype(np.int32)
        elif isinstance(pts, torch.Tensor):
            pts = pts.to(torch.int32)
        return pts

    @classmethod
    def from_fov_camera(cls, fov, aspect, resolution, near_plane, far_plane):
        """utils/camera.py:305-314."""
        W = resolution
        H = int(resolution / aspect)
        cx = W / 2
        cy = H / 2
        fx = cx / np.tan(fov / 2)
        fy = cy / np.tan(fov / 2)
        return cls(fx, fy, cx, cy, W, H, near_plane, far_plane)

    def image_constants(self, tile_size=16):
        """Tile grid and pixel mapping used by SHRenderer.forward
        (gs/sh_renderer.py:238-248)."""
        H, W = self.h, self.w
        n_tiles_h = H // tile_size + (H % tile_size > 0)
        n_tiles_w = W // tile_size + (W % tile_size > 0)
        topleft = (float(_f32(-self.cx / self.fx)), float(_f32(-self.cy / self.fy)))
        return dict(H=H, W=W, n_tiles_h=int(n_tiles_h), n_tiles_w=int(n_tiles_w), topleft=topleft,
                    pixel_size_x=1.0 / self.fx, pixel_size_y=1.0 / self.fy)


def get_c2w_from_up_and_look_at(up, look_at, pos):
    """gs/debug.py:23-37: camera-to-world [3,4] float32, OpenCV axes (y = -up)."""
    up = np.asarray(up, np.float64)
    look_at = np.asarray(look_at, np.float64)
    pos = np.asarray(pos, np.float64)
    up = up / np.linalg.norm(up)
    z = look_at - pos
    z = z / np.linalg.norm(z)
    y = -up
    x = np.cross(y, z)
    y = np.cross(z, x)
    c2w = np.zeros([3, 4], dtype=np.float32)
    c2w[:3, 0] = x
    c2w[:3, 1] = y
    c2w[:3, 2] = z
    c2w[:3, 3] = pos
    return c2w
