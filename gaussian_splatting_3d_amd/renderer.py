"""Rasterise entry points with the reference's autograd signatures.

Drop-in for gs/renderer.py:365-999: ``render_sh``, ``render_sh_bg``,
``render_start_end`` and ``render`` take the same positional arguments and
return the same flat ``[H*W*3]`` image; their backward returns gradients for
the first four inputs and ``None`` for every other input.  ``project_gaussians``
(with ``jacobian`` and ``project_pts``) is the reference's PyTorch projection,
kept for callers that build their own pipeline; the fused path in
``rasterize.py`` replaces it with gfx950 kernels.

The forward keeps, besides the image, the per-pixel final transmittance and
the number of list entries examined before the transmittance break; the
backward is the reverse-order adjoint that starts from them.
"""
from __future__ import annotations

import torch

from .backend import _backend


@torch.no_grad()
def jacobian(u):
    """gs/renderer.py:365-377 (no gradient)."""
    l = torch.norm(u, dim=-1)
    J = torch.zeros(u.size(0), 3, 3).to(u)
    J[..., 0, 0] = 1.0 / u[..., 2]
    J[..., 2, 0] = u[..., 0] / l
    J[..., 1, 1] = 1.0 / u[..., 2]
    J[..., 2, 1] = u[..., 1] / l
    J[..., 0, 2] = -u[..., 0] / u[..., 2] / u[..., 2]
    J[..., 1, 2] = -u[..., 1] / u[..., 2] / u[..., 2]
    J[..., 2, 2] = u[..., 2] / l
    return J


def project_pts(pts, c2w):
    """gs/renderer.py:380-387: world -> camera, R^T (p - t)."""
    d = -c2w[..., :3, 3]
    W = torch.transpose(c2w[..., :3, :3], -1, -2)
    return torch.einsum("ij,bj->bi", W, pts + d)


def quat_to_rotmat(qvec):
    """Unit-quaternion (w,x,y,z) rotation matrices after normalisation
    (utils/transforms.py:31-45 via kornia, WXYZ order)."""
    q = torch.nn.functional.normalize(qvec, p=2.0, dim=-1, eps=1e-12)
    w, x, y, z = q.unbind(-1)
    tx, ty, tz = 2.0 * x, 2.0 * y, 2.0 * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    one = torch.ones_like(w)
    return torch.stack([one - (tyy + tzz), txy - twz, txz + twy,
                        txy + twz, one - (txx + tzz), tyz - twx,
                        txz - twy, tyz + twx, one - (txx + tyy)], -1).view(-1, 3, 3)


def qsvec2rotmat_batched(qvec, svec):
    """utils/transforms.py:31-45: R(q) with columns scaled by s."""
    return svec.unsqueeze(-2) * quat_to_rotmat(qvec)


def project_gaussians(mean, qvec, svec, c2w, detach_depth: bool = True):
    """gs/renderer.py:390-419: (mean2d [N,2], cov2d [N,2,2], JW [N,3,3], depth [N,1])."""
    projected_mean = project_pts(mean, c2w)
    rotmat = qsvec2rotmat_batched(qvec, svec)
    sigma = rotmat @ torch.transpose(rotmat, -1, -2)
    W = torch.transpose(c2w[:3, :3], -1, -2)
    J = jacobian(projected_mean)
    JW = torch.einsum("bij,jk->bik", J, W)
    projected_cov = torch.bmm(torch.bmm(JW, sigma), torch.transpose(JW, -1, -2).contiguous())[..., :2, :2].contiguous()
    if detach_depth:
        depth = projected_mean[..., 2:].clone().contiguous().detach()
    else:
        depth = projected_mean[..., 2:].clone().contiguous()
    projected_mean = projected_mean[..., :2].contiguous() / depth
    return projected_mean, projected_cov, JW, depth


def _pixel_state(H, W, device, bg=False):
    T = torch.empty(H * W, dtype=torch.float32, device=device)
    nd = torch.empty(H * W, dtype=torch.int32, device=device)
    pre = torch.empty(H * W * 3, dtype=torch.float32, device=device) if bg else None
    return T, nd, pre


class _render_sh(torch.autograd.Function):
    """gs/renderer.py:672-828."""

    @staticmethod
    def forward(ctx, mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, topleft, c2w, tile_size, n_tiles_h,
                n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh):
        out = torch.zeros([H * W * 3], dtype=torch.float32, device=mean.device)
        T, nd, _ = _pixel_state(H, W, mean.device)
        _backend._sh_checks(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w)
        if not 1 <= int(C) <= 4:
            raise RuntimeError(f"render_sh: SH order C={C} is not supported (1..4)")
        _backend.sh_forward(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w, tile_size,
                            n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh, T_final=T,
                            n_done=nd, write_empty=True)
        ctx.save_for_backward(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w, T, nd)
        ctx.const = [tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh]
        return out

    @staticmethod
    def backward(ctx, grad):
        mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w, T, nd = ctx.saved_tensors
        grad_mean = torch.zeros_like(mean)
        grad_cov = torch.zeros_like(cov)
        grad_sh_coeffs = torch.zeros_like(sh_coeffs)
        grad_alpha = torch.zeros_like(alpha)
        tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh = ctx.const
        _backend.sh_backward(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, T, nd, None,
                             grad.contiguous(), grad_mean, grad_cov, grad_sh_coeffs, grad_alpha, topleft, c2w,
                             tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh)
        return (grad_mean, grad_cov, grad_sh_coeffs, grad_alpha) + (None,) * 14


class _render_sh_bg(torch.autograd.Function):
    """gs/renderer.py:831-993.  Pixels of empty tiles receive bg (the intent
    of vol_render_bg.h:28-37); elsewhere out = out*T + bg*(1-T) when
    T > thresh (vol_render_bg.h:90-94).  The backward uses the reference's
    formula with the blended output in place of the final colour."""

    @staticmethod
    def forward(ctx, mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, topleft, c2w, tile_size, n_tiles_h,
                n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh, bg_rgb):
        out = torch.zeros([H * W * 3], dtype=torch.float32, device=mean.device)
        T, nd, pre = _pixel_state(H, W, mean.device, bg=True)
        _backend._sh_checks(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w)
        if not 1 <= int(C) <= 4:
            raise RuntimeError(f"render_sh_bg: SH order C={C} is not supported (1..4)")
        _backend.sh_forward(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w, tile_size,
                            n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh, bg_rgb=bg_rgb,
                            T_final=T, n_done=nd, pre_bg=pre, write_empty=True)
        ctx.save_for_backward(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w, bg_rgb, T,
                              nd, pre)
        ctx.const = [tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh]
        return out

    @staticmethod
    def backward(ctx, grad):
        (mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, topleft, c2w, bg_rgb, T, nd,
         pre) = ctx.saved_tensors
        grad_mean = torch.zeros_like(mean)
        grad_cov = torch.zeros_like(cov)
        grad_sh_coeffs = torch.zeros_like(sh_coeffs)
        grad_alpha = torch.zeros_like(alpha)
        tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh = ctx.const
        _backend.sh_backward(mean, cov, sh_coeffs, alpha, start, end, gaussian_ids, out, T, nd, pre,
                             grad.contiguous(), grad_mean, grad_cov, grad_sh_coeffs, grad_alpha, topleft, c2w,
                             tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, C, thresh,
                             bg_rgb=bg_rgb)
        return (grad_mean, grad_cov, grad_sh_coeffs, grad_alpha) + (None,) * 15


class _render_start_end(torch.autograd.Function):
    """gs/renderer.py:539-669: RGB colours, start/end tile ranges."""

    @staticmethod
    def forward(ctx, mean, cov, color, alpha, start, end, gaussian_ids, topleft, tile_size, n_tiles_h, n_tiles_w,
                pixel_size_x, pixel_size_y, H, W, thresh):
        out = torch.zeros([H * W * 3], dtype=torch.float32, device=mean.device)
        T, nd, _ = _pixel_state(H, W, mean.device)
        _backend.rgb_forward(mean, cov, color, alpha, start, end, gaussian_ids, out, topleft, tile_size, n_tiles_h,
                             n_tiles_w, pixel_size_x, pixel_size_y, H, W, thresh, T_final=T, n_done=nd,
                             write_empty=True)
        ctx.save_for_backward(mean, cov, color, alpha, start, end, gaussian_ids, out, topleft, T, nd)
        ctx.const = [tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, thresh]
        return out

    @staticmethod
    def backward(ctx, grad):
        mean, cov, color, alpha, start, end, gaussian_ids, out, topleft, T, nd = ctx.saved_tensors
        grad_mean = torch.zeros_like(mean)
        grad_cov = torch.zeros_like(cov)
        grad_color = torch.zeros_like(color)
        grad_alpha = torch.zeros_like(alpha)
        tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, thresh = ctx.const
        _backend.rgb_backward(mean, cov, color, alpha, start, end, gaussian_ids, out, T, nd, grad.contiguous(),
                              grad_mean, grad_cov, grad_color, grad_alpha, topleft, tile_size, n_tiles_h, n_tiles_w,
                              pixel_size_x, pixel_size_y, H, W, thresh)
        return (grad_mean, grad_cov, grad_color, grad_alpha) + (None,) * 12


class _render(torch.autograd.Function):
    """gs/renderer.py:422-536: RGB colours with an offset array
    (tile t owns sorted entries [offset[t], offset[t+1]))."""

    @staticmethod
    def forward(ctx, mean, cov, color, alpha, offset, gaussian_ids, topleft, tile_size, n_tiles_h, n_tiles_w,
                pixel_size_x, pixel_size_y, H, W, thresh):
        start = offset[:-1].contiguous()
        end = offset[1:].contiguous()
        out = torch.zeros([H * W * 3], dtype=torch.float32, device=mean.device)
        T, nd, _ = _pixel_state(H, W, mean.device)
        _backend.rgb_forward(mean, cov, color, alpha, start, end, gaussian_ids, out, topleft, tile_size, n_tiles_h,
                             n_tiles_w, pixel_size_x, pixel_size_y, H, W, thresh, T_final=T, n_done=nd,
                             write_empty=True)
        ctx.save_for_backward(mean, cov, color, alpha, start, end, gaussian_ids, out, topleft, T, nd)
        ctx.const = [tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, thresh]
        return out

    @staticmethod
    def backward(ctx, grad):
        mean, cov, color, alpha, start, end, gaussian_ids, out, topleft, T, nd = ctx.saved_tensors
        grad_mean = torch.zeros_like(mean)
        grad_cov = torch.zeros_like(cov)
        grad_color = torch.zeros_like(color)
        grad_alpha = torch.zeros_like(alpha)
        tile_size, n_tiles_h, n_tiles_w, pixel_size_x, pixel_size_y, H, W, thresh = ctx.const
        _backend.rgb_backward(mean, cov, color, alpha, start, end, gaussian_ids, out, T, nd, grad.contiguous(),
                              grad_mean, grad_cov, grad_color, grad_alpha, topleft, tile_size, n_tiles_h, n_tiles_w,
                              pixel_size_x, pixel_size_y, H, W, thresh)
        return (grad_mean, grad_cov, grad_color, grad_alpha) + (None,) * 11


render = _render.apply
render_start_end = _render_start_end.apply
render_sh = _render_sh.apply
render_sh_bg = _render_sh_bg.apply
