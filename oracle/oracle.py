"""CPU oracle for the rasteriser path — TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check in __graft_entry__.py and the CPU-baseline leg of
bench.py may import this module, and only as the checker.  The product package
never imports it.

What lives here:
  * ctypes bindings to liboracle.so (gs_oracle.c): the float arithmetic of the
    reference path restated in C (compositing forward/backward, cull,
    projection, tile AABB), see the header of gs_oracle.c for the arithmetic
    convention;
  * numpy restatements of the integer work: key emission, stable sort, tile
    ranges (gs/src/include/aabb_culling.h:15-41, 70-103, 192-256);
  * the frustum planes (utils/camera.py:249-283) in float32 numpy;
  * the reference's autograd graph over the projection (gs/renderer.py:365-419,
    utils/transforms.py:31-45, utils/activations.py) in float64 torch, with the
    reference's detaches, used as the projection-adjoint oracle;
  * an FP64 brute-force compositor for tiny scenes and a central-difference
    harness, used to pin the C restatement.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

f32 = np.float32


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i32, i64, fl = ctypes.c_int, ctypes.c_int64, ctypes.c_float
        L.orc_expf.restype = fl
        L.orc_expf.argtypes = [fl]
        L.orc_f2i_sat.restype = ctypes.c_int32
        L.orc_f2i_sat.argtypes = [fl]
        L.orc_render_sh_fwd.restype = i32
        L.orc_render_sh_fwd.argtypes = [i32] + [P] * 10 + [i32, i32, i32, fl, fl, i32, i32, fl, i32, P, P, P, P]
        L.orc_render_sh_bwd.restype = i32
        L.orc_render_sh_bwd.argtypes = [i32] + [P] * 15 + [i32, i32, i32, fl, fl, i32, i32, fl, i32, P, i32, P, P,
                                                           P, P]
        L.orc_render_sh_bwd_f64.restype = i32
        L.orc_render_sh_bwd_f64.argtypes = [i32] + [P] * 14 + [i32, i32, i32, fl, fl, i32, i32, fl, i32, P, i32]
        L.orc_set_threads.argtypes = [i32]
        L.orc_get_threads.restype = i32
        L.orc_render_rgb_fwd.restype = i32
        L.orc_render_rgb_fwd.argtypes = [P] * 9 + [i32, i32, i32, fl, fl, i32, i32, fl, P, P]
        L.orc_render_rgb_bwd.restype = i32
        L.orc_render_rgb_bwd.argtypes = [P] * 14 + [i32, i32, i32, fl, fl, i32, i32, fl, i32]
        L.orc_cull.restype = None
        L.orc_cull.argtypes = [i32, P, P, P, P, P, fl]
        L.orc_activate.restype = None
        L.orc_activate.argtypes = [i32, P, P, P, P]
        L.orc_quat_to_rot.restype = None
        L.orc_quat_to_rot.argtypes = [P, P]
        L.orc_project.restype = None
        L.orc_project.argtypes = [i32, P, P, P, P, P, P, P]
        L.orc_aabb.restype = i64
        L.orc_aabb.argtypes = [i32, P, P, fl, fl, fl, fl, i32, i32, fl, i32, P, P]
        L.orc_expf_vec.restype = None
        L.orc_expf_vec.argtypes = [i64, P, P]
        L.orc_sh_basis_vec.restype = None
        L.orc_sh_basis_vec.argtypes = [i64, P, P, i32]
        L.orc_gauss2d_f_vec.restype = None
        L.orc_gauss2d_f_vec.argtypes = [i64, P, P, P, P]
        L.orc_direction.restype = None
        L.orc_direction.argtypes = [P, fl, fl, P]
        _lib = L
    return _lib


def _c(a, dtype):
    """Contiguous numpy array of `dtype` (torch tensors are moved to CPU)."""
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# --------------------------------------------------------------------------
# scalar helpers
# --------------------------------------------------------------------------

def expf(x):
    x = _c(x, f32).reshape(-1)
    y = np.empty_like(x)
    lib().orc_expf_vec(x.size, _p(x), _p(y))
    return y


def sh_basis(dirs, C):
    d = _c(dirs, f32).reshape(-1, 3)
    out = np.zeros((d.shape[0], C * C), f32)
    lib().orc_sh_basis_vec(d.shape[0], _p(d), _p(out), C)
    return out


def gauss2d(mean, cov, pos):
    m = _c(mean, f32).reshape(-1, 2)
    c = _c(cov, f32).reshape(-1, 4)
    p = _c(pos, f32).reshape(-1, 2)
    out = np.empty(m.shape[0], f32)
    lib().orc_gauss2d_f_vec(m.shape[0], _p(m), _p(c), _p(p), _p(out))
    return out


def direction(c2w, px, py):
    c = _c(c2w, f32).reshape(-1)
    d = np.empty(3, f32)
    lib().orc_direction(_p(c), f32(px), f32(py), _p(d))
    return d


def f2i_sat(v):
    """CUDA float->int32 conversion semantics, elementwise (numpy)."""
    v = np.asarray(v, f32)
    out = np.zeros(v.shape, np.int64)
    big = v >= f32(2147483648.0)
    small = v <= f32(-2147483648.0)
    ok = ~(np.isnan(v) | big | small)
    out[ok] = np.trunc(v[ok]).astype(np.int64)
    out[big] = 2147483647
    out[small] = -2147483648
    return out.astype(np.int32)


# --------------------------------------------------------------------------
# camera (utils/camera.py:219-314)
# --------------------------------------------------------------------------

@dataclass
class Camera:
    fx: float
    fy: float
    cx: float
    cy: float
    w: int
    h: int
    near_plane: float
    far_plane: float

    @property
    def yfov(self):
        return 2 * np.arctan(self.h / (2 * self.fy))

    @property
    def aspect(self):
        return self.w / self.h


def _cross(a, b):
    return np.array(
        [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], f32
    )


def frustum(c2w, cam: Camera):
    """utils/camera.py:249-283 in float32: (normals[6,3], pts[6,3])."""
    c = _c(c2w, f32).reshape(3, 4)
    up = -c[:, 1]
    right = c[:, 0]
    lookat = c[:, 2]
    t = c[:, 3]
    half_vside = cam.far_plane * np.tan(cam.yfov * 0.5)
    half_hside = half_vside * cam.aspect
    hv, hh = f32(half_vside), f32(half_hside)
    near_point = f32(cam.near_plane) * lookat
    far_point = f32(cam.far_plane) * lookat
    normals = [
        lookat,
        -lookat,
        _cross(far_point - hh * right, up),
        _cross(up, far_point + hh * right),
        _cross(far_point + hv * up, right),
        _cross(right, far_point - hv * up),
    ]
    pts = [near_point + t, far_point + t, t, t, t, t]
    normals = np.stack(normals).astype(f32)
    nrm = np.sqrt(normals[:, 0] * normals[:, 0] + normals[:, 1] * normals[:, 1] + normals[:, 2] * normals[:, 2])
    normals = normals / np.maximum(nrm, f32(1e-12))[:, None]
    return normals.astype(f32), np.stack(pts).astype(f32)


# --------------------------------------------------------------------------
# preprocessing (C restatement) and binning (numpy)
# --------------------------------------------------------------------------

def activate(log_s, logit):
    ls = _c(log_s, f32).reshape(-1, 3)
    lg = _c(logit, f32).reshape(-1)
    s = np.empty_like(ls)
    a = np.empty_like(lg)
    lib().orc_activate(lg.size, _p(ls), _p(lg), _p(s), _p(a))
    return s, a


def cull(mean, svec, normals, pts, thresh):
    m = _c(mean, f32).reshape(-1, 3)
    s = _c(svec, f32).reshape(-1, 3)
    n = _c(normals, f32).reshape(6, 3)
    p = _c(pts, f32).reshape(6, 3)
    mask = np.zeros(m.shape[0], np.uint8)
    lib().orc_cull(m.shape[0], _p(m), _p(s), _p(n), _p(p), _p(mask), f32(thresh))
    return mask.astype(bool)


def quat_to_rot(q):
    q = _c(q, f32).reshape(4)
    R = np.empty(9, f32)
    lib().orc_quat_to_rot(_p(q), _p(R))
    return R.reshape(3, 3)


def project(mean, qvec, svec, c2w):
    m = _c(mean, f32).reshape(-1, 3)
    q = _c(qvec, f32).reshape(-1, 4)
    s = _c(svec, f32).reshape(-1, 3)
    c = _c(c2w, f32).reshape(12)
    V = m.shape[0]
    mean2d = np.empty((V, 2), f32)
    cov2d = np.empty((V, 4), f32)
    depth = np.empty((V,), f32)
    lib().orc_project(V, _p(m), _p(q), _p(s), _p(c), _p(mean2d), _p(cov2d), _p(depth))
    return mean2d, cov2d, depth


def aabb(mean2d, cov2d, cam: Camera, D, tile_size=16):
    m = _c(mean2d, f32).reshape(-1, 2)
    c = _c(cov2d, f32).reshape(-1, 4)
    V = m.shape[0]
    tl = np.zeros((V, 2), np.int32)
    br = np.zeros((V, 2), np.int32)
    P = lib().orc_aabb(V, _p(m), _p(c), f32(cam.fx), f32(cam.fy), f32(cam.cx), f32(cam.cy),
                       cam.w, cam.h, f32(D), tile_size, _p(tl), _p(br))
    return int(P), tl, br


def emit_keys(tl, br, depth, n_tiles_w):
    """aabb_culling.h:15-41: per Gaussian i, x outer, y inner; key =
    (y*n_tiles_w + x) << 32 | float_bits(depth_i), value = i.  Emission is in
    index order (the reference reserves slots with an atomic counter, so its
    pre-sort order is unspecified)."""
    tl = np.asarray(tl, np.int64)
    br = np.asarray(br, np.int64)
    nx = br[:, 0] - tl[:, 0] + 1
    ny = br[:, 1] - tl[:, 1] + 1
    cnt = nx * ny
    P = int(cnt.sum())
    gid = np.repeat(np.arange(tl.shape[0], dtype=np.int64), cnt)
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]]) if tl.shape[0] else np.zeros(0, np.int64)
    local = np.arange(P, dtype=np.int64) - np.repeat(first, cnt)
    nyg = np.repeat(ny, cnt)
    x = np.repeat(tl[:, 0], cnt) + local // nyg
    y = np.repeat(tl[:, 1], cnt) + local % nyg
    tile = y * n_tiles_w + x
    dbits = np.asarray(depth, f32).reshape(-1).view(np.uint32).astype(np.uint64)
    keys = (tile.astype(np.uint64) << np.uint64(32)) | dbits[gid]
    return keys, gid.astype(np.int32)


def sort_pairs(keys, values):
    """Ascending stable sort on the 64-bit key (aabb_culling.h:235-241)."""
    order = np.argsort(keys, kind="stable")
    return keys[order], values[order]


def tile_ranges(sorted_keys, n_tiles):
    """aabb_culling.h:70-103, 248-249: start/end per tile, -1 when empty."""
    start = -np.ones(n_tiles, np.int32)
    end = -np.ones(n_tiles, np.int32)
    P = sorted_keys.size
    if P == 0:
        return start, end
    tile = (sorted_keys >> np.uint64(32)).astype(np.int64)
    first = np.ones(P, bool)
    first[1:] = tile[1:] != tile[:-1]
    last = np.ones(P, bool)
    last[:-1] = tile[1:] != tile[:-1]
    idx = np.arange(P, dtype=np.int64)
    start[tile[first]] = idx[first]
    end[tile[last]] = idx[last] + 1
    return start, end


def bin_gaussians(tl, br, depth, n_tiles_h, n_tiles_w):
    keys, vals = emit_keys(tl, br, depth, n_tiles_w)
    skeys, svals = sort_pairs(keys, vals)
    start, end = tile_ranges(skeys, n_tiles_h * n_tiles_w)
    return dict(keys=keys, values=vals, sorted_keys=skeys, ids=svals, start=start, end=end)


# --------------------------------------------------------------------------
# compositing (C restatement)
# --------------------------------------------------------------------------

def render_sh_fwd(mean, cov, sh, alpha, start, end, ids, topleft, c2w, tile_size, n_tiles_h,
                  n_tiles_w, psx, psy, H, W, C, thresh, mode=0, bg=None, out=None):
    """render_sh / render_sh_bg forward.  Returns (out[H*W*3], T[H*W], n_done[H*W], pre_bg)."""
    m = _c(mean, f32)
    c = _c(cov, f32)
    s = _c(sh, f32)
    a = _c(alpha, f32)
    st = _c(start, np.int32)
    en = _c(end, np.int32)
    gi = _c(ids, np.int32)
    tlp = _c(topleft, f32)
    cw = _c(c2w, f32)
    o = np.zeros(H * W * 3, f32) if out is None else _c(out, f32).copy()
    T = np.ones(H * W, f32)
    nd = np.zeros(H * W, np.int32)
    pre = np.zeros(H * W * 3, f32)
    b = _c(bg if bg is not None else [0, 0, 0], f32)
    rc = lib().orc_render_sh_fwd(C, _p(m), _p(c), _p(s), _p(a), _p(st), _p(en), _p(gi), _p(o), _p(tlp),
                                 _p(cw), tile_size, n_tiles_h, n_tiles_w, f32(psx), f32(psy), H, W,
                                 f32(thresh), mode, _p(b), _p(T), _p(nd), _p(pre))
    if rc != 0:
        raise RuntimeError(f"oracle render_sh_fwd failed ({rc})")
    return o, T, nd, pre


def render_sh_bwd(mean, cov, sh, alpha, start, end, ids, out, grad_out, topleft, c2w, tile_size,
                  n_tiles_h, n_tiles_w, psx, psy, H, W, C, thresh, mode=0, bg=None, want_abs=False):
    """Reference backward (forward-order recompute).  Returns (grads dict,
    mismatch flag); with want_abs, (grads, mismatch, abs) where abs holds,
    per gradient component, the sum of its per-pixel terms' rounding scales
    (float64): |term| for SH, and for mean, cov and alpha the term with the
    factor P = sum g (y T - (F - o)/(1 - aG)) replaced by the sum of the
    magnitudes of its parts, since F - o cancels (gs_oracle.c)."""
    m = _c(mean, f32)
    c = _c(cov, f32)
    s = _c(sh, f32)
    a = _c(alpha, f32)
    V = a.size
    gm = np.zeros(V * 2, f32)
    gc = np.zeros(V * 4, f32)
    gs = np.zeros(s.size, f32)
    ga = np.zeros(V, f32)
    b = _c(bg if bg is not None else [0, 0, 0], f32)
    am = np.zeros(V * 2) if want_abs else None
    ac = np.zeros(V * 4) if want_abs else None
    ash = np.zeros(s.size) if want_abs else None
    aa = np.zeros(V) if want_abs else None
    rc = lib().orc_render_sh_bwd(C, _p(m), _p(c), _p(s), _p(a), _p(_c(start, np.int32)), _p(_c(end, np.int32)),
                                 _p(_c(ids, np.int32)), _p(_c(out, f32)), _p(gm), _p(gc), _p(gs), _p(ga),
                                 _p(_c(grad_out, f32)), _p(_c(topleft, f32)), _p(_c(c2w, f32)), tile_size,
                                 n_tiles_h, n_tiles_w, f32(psx), f32(psy), H, W, f32(thresh), mode, _p(b), V,
                                 _p(am), _p(ac), _p(ash), _p(aa))
    if rc < 0:
        raise RuntimeError("oracle render_sh_bwd: unsupported C")
    grads = dict(mean=gm.reshape(V, 2), cov=gc.reshape(V, 2, 2), sh=gs.reshape(np.shape(sh)), alpha=ga)
    if not want_abs:
        return grads, rc
    absd = dict(mean=am.reshape(V, 2), cov=ac.reshape(V, 2, 2), sh=ash.reshape(np.shape(sh)), alpha=aa)
    return grads, rc, absd


def render_sh_bwd_f64(mean, cov, sh, alpha, start, end, ids, grad_out, topleft, c2w, tile_size, n_tiles_h,
                      n_tiles_w, psx, psy, H, W, C, thresh, mode=0, bg=None):
    """The adjoint of render_sh_bwd evaluated in float64 along the float
    forward's decisions (gs_oracle.c orc_render_sh_bwd_f64): the exact value
    the float evaluations approximate.  Returns a dict of float64 arrays."""
    m = _c(mean, f32)
    c = _c(cov, f32)
    s = _c(sh, f32)
    a = _c(alpha, f32)
    V = a.size
    gm = np.zeros(V * 2)
    gc = np.zeros(V * 4)
    gs = np.zeros(s.size)
    ga = np.zeros(V)
    b = _c(bg if bg is not None else [0, 0, 0], f32)
    rc = lib().orc_render_sh_bwd_f64(C, _p(m), _p(c), _p(s), _p(a), _p(_c(start, np.int32)), _p(_c(end, np.int32)),
                                     _p(_c(ids, np.int32)), _p(gm), _p(gc), _p(gs), _p(ga), _p(_c(grad_out, f32)),
                                     _p(_c(topleft, f32)), _p(_c(c2w, f32)), tile_size, n_tiles_h, n_tiles_w,
                                     f32(psx), f32(psy), H, W, f32(thresh), mode, _p(b), V)
    if rc < 0:
        raise RuntimeError("oracle render_sh_bwd_f64: unsupported C")
    return dict(mean=gm.reshape(V, 2), cov=gc.reshape(V, 2, 2), sh=gs.reshape(np.shape(sh)), alpha=ga)


def set_threads(n):
    """Threads of the oracle's parallel compositing loops (OpenMP)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


def render_rgb_fwd(mean, cov, color, alpha, start, end, ids, topleft, tile_size, n_tiles_h, n_tiles_w,
                   psx, psy, H, W, thresh):
    o = np.zeros(H * W * 3, f32)
    T = np.ones(H * W, f32)
    nd = np.zeros(H * W, np.int32)
    lib().orc_render_rgb_fwd(_p(_c(mean, f32)), _p(_c(cov, f32)), _p(_c(color, f32)), _p(_c(alpha, f32)),
                             _p(_c(start, np.int32)), _p(_c(end, np.int32)), _p(_c(ids, np.int32)), _p(o),
                             _p(_c(topleft, f32)), tile_size, n_tiles_h, n_tiles_w, f32(psx), f32(psy), H, W,
                             f32(thresh), _p(T), _p(nd))
    return o, T, nd


def render_rgb_bwd(mean, cov, color, alpha, start, end, ids, out, grad_out, topleft, tile_size, n_tiles_h,
                   n_tiles_w, psx, psy, H, W, thresh):
    V = _c(alpha, f32).size
    gm = np.zeros(V * 2, f32)
    gc = np.zeros(V * 4, f32)
    gcol = np.zeros(V * 3, f32)
    ga = np.zeros(V, f32)
    rc = lib().orc_render_rgb_bwd(_p(_c(mean, f32)), _p(_c(cov, f32)), _p(_c(color, f32)), _p(_c(alpha, f32)),
                                  _p(_c(start, np.int32)), _p(_c(end, np.int32)), _p(_c(ids, np.int32)),
                                  _p(_c(out, f32)), _p(gm), _p(gc), _p(gcol), _p(ga), _p(_c(grad_out, f32)),
                                  _p(_c(topleft, f32)), tile_size, n_tiles_h, n_tiles_w, f32(psx), f32(psy), H,
                                  W, f32(thresh), V)
    return dict(mean=gm.reshape(V, 2), cov=gc.reshape(V, 2, 2), color=gcol.reshape(V, 3), alpha=ga), rc


# --------------------------------------------------------------------------
# full forward pipeline (gs/sh_renderer.py:188-316)
# --------------------------------------------------------------------------

def image_constants(cam: Camera, tile_size=16):
    """sh_renderer.py:238-248."""
    H, W = cam.h, cam.w
    nth = H // tile_size + (H % tile_size > 0)
    ntw = W // tile_size + (W % tile_size > 0)
    topleft = np.array([-cam.cx / cam.fx, -cam.cy / cam.fy], f32)
    return dict(H=H, W=W, n_tiles_h=nth, n_tiles_w=ntw, topleft=topleft,
                psx=f32(1.0 / cam.fx), psy=f32(1.0 / cam.fy))


def pipeline_forward(mean, qvec, log_s, logit, sh, c2w, cam: Camera, C, *, tile_size=16, D=6.0,
                     frustum_radius=1.0, thresh=1e-4, skip_frustum_culling=False, mode=0, bg=None):
    """SHRenderer.forward restated end to end: activations, cull, compaction,
    projection, AABB, binning, compositing.  Returns a dict of every stage."""
    mean = _c(mean, f32).reshape(-1, 3)
    qvec = _c(qvec, f32).reshape(-1, 4)
    s, a = activate(log_s, logit)
    sh = _c(sh, f32)
    normals, pts = frustum(c2w, cam)
    if skip_frustum_culling:
        mask = np.ones(mean.shape[0], bool)
    else:
        mask = cull(mean, s, normals, pts, frustum_radius)
    idx = np.nonzero(mask)[0]
    mean2d, cov2d, depth = project(mean[idx], qvec[idx], s[idx], c2w)
    P, tl, br = aabb(mean2d, cov2d, cam, D, tile_size)
    ic = image_constants(cam, tile_size)
    b = bin_gaussians(tl, br, depth, ic["n_tiles_h"], ic["n_tiles_w"])
    assert b["ids"].size == P
    sh_v = np.ascontiguousarray(sh[idx][..., : C * C])
    out, T, nd, pre = render_sh_fwd(mean2d, cov2d, sh_v, a[idx], b["start"], b["end"], b["ids"], ic["topleft"],
                                    c2w, tile_size, ic["n_tiles_h"], ic["n_tiles_w"], ic["psx"], ic["psy"],
                                    ic["H"], ic["W"], C, thresh, mode=mode, bg=bg)
    return dict(mask=mask, idx=idx, s=s, a=a, normals=normals, pts=pts, mean2d=mean2d, cov2d=cov2d,
                depth=depth, P=P, tl=tl, br=br, sh_v=sh_v, out=out, T=T, n_done=nd, pre_bg=pre, **b, **ic)


# --------------------------------------------------------------------------
# projection adjoint oracle: the reference's autograd graph in float64 torch
# --------------------------------------------------------------------------

def _torch():
    import torch

    return torch


def quat_to_rot_t(q):
    torch = _torch()
    qn = torch.nn.functional.normalize(q, p=2.0, dim=-1, eps=1e-12)
    w, x, y, z = qn.unbind(-1)
    tx, ty, tz = 2.0 * x, 2.0 * y, 2.0 * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    one = torch.ones_like(w)
    R = torch.stack([one - (tyy + tzz), txy - twz, txz + twy,
                     txy + twz, one - (txx + tzz), tyz - twx,
                     txz - twy, tyz + twx, one - (txx + tyy)], -1)
    return R.view(-1, 3, 3)


def project_t(mean, qvec, svec, c2w, detach_depth=True):
    """gs/renderer.py:365-419 as a torch graph (float64 recommended): J is
    computed under no_grad and depth is detached, exactly as the reference."""
    torch = _torch()
    d = -c2w[:3, 3]
    Wm = c2w[:3, :3].T
    u = torch.einsum("ij,bj->bi", Wm, mean + d)
    R = quat_to_rot_t(qvec)
    M = svec.unsqueeze(-2) * R
    sigma = M @ M.transpose(-1, -2)
    with torch.no_grad():
        J = torch.zeros(u.shape[0], 3, 3, dtype=u.dtype)
        l = torch.norm(u, dim=-1)
        J[:, 0, 0] = 1.0 / u[:, 2]
        J[:, 2, 0] = u[:, 0] / l
        J[:, 1, 1] = 1.0 / u[:, 2]
        J[:, 2, 1] = u[:, 1] / l
        J[:, 0, 2] = -u[:, 0] / u[:, 2] / u[:, 2]
        J[:, 1, 2] = -u[:, 1] / u[:, 2] / u[:, 2]
        J[:, 2, 2] = u[:, 2] / l
    JW = torch.einsum("bij,jk->bik", J, Wm)
    cov = (JW @ sigma @ JW.transpose(-1, -2).contiguous())[..., :2, :2]
    depth = u[..., 2:].detach() if detach_depth else u[..., 2:]
    mean2d = u[..., :2] / depth
    return mean2d, cov, depth


def projection_adjoint(mean, qvec, log_s, logit, sh, c2w, idx, grad_mean2d, grad_cov, grad_sh_v, grad_alpha_v,
                       C, detach_depth=True):
    """Chain the rasteriser-boundary gradients back to the raw parameters via
    the reference's graph (projection + activations + masked gather), float64.
    Returns grads for (mean, qvec, log_s, logit, sh) with zeros for culled
    Gaussians."""
    torch = _torch()
    dt = torch.float64
    P = [torch.tensor(_c(x, np.float64)).requires_grad_(True) for x in (mean, qvec, log_s, logit, sh)]
    m, q, ls, lg, shp = P
    cw = torch.tensor(_c(c2w, np.float64)).reshape(3, 4)
    ii = torch.tensor(np.asarray(idx, np.int64))
    s = torch.exp(ls)
    a = torch.sigmoid(lg)
    mean2d, cov, _ = project_t(m[ii], q[ii], s[ii], cw, detach_depth)
    sh_v = shp[ii][..., : C * C]
    loss = (mean2d * torch.tensor(_c(grad_mean2d, np.float64), dtype=dt)).sum()
    loss = loss + (cov * torch.tensor(_c(grad_cov, np.float64).reshape(-1, 2, 2), dtype=dt)).sum()
    loss = loss + (sh_v * torch.tensor(_c(grad_sh_v, np.float64), dtype=dt)).sum()
    loss = loss + (a[ii] * torch.tensor(_c(grad_alpha_v, np.float64), dtype=dt)).sum()
    gr = torch.autograd.grad(loss, P, allow_unused=True)
    return [np.zeros(p.shape) if g is None else g.numpy() for p, g in zip(P, gr)]


# --------------------------------------------------------------------------
# independent checks used to pin the C restatement
# --------------------------------------------------------------------------

def brute_force_render(mean2d, cov2d, sh, alpha, depth, tl, br, topleft, c2w, psx, psy, H, W, C, thresh,
                       tile_size=16):
    """FP64 per-pixel compositor written without tiles, batches or a sort
    kernel: for every pixel, the Gaussians whose AABB covers the pixel's tile,
    ordered by (depth bits, index); reference rules A.3 in float64 (sqrt,
    exp, sigmoid via numpy).  For tiny scenes only."""
    mean2d = np.asarray(mean2d, np.float64).reshape(-1, 2)
    cov2d = np.asarray(cov2d, np.float64).reshape(-1, 4)
    sh = np.asarray(sh, np.float64).reshape(len(mean2d), 3, -1)
    alpha = np.asarray(alpha, np.float64).reshape(-1)
    dbits = np.asarray(depth, f32).reshape(-1).view(np.uint32)
    c2 = np.asarray(c2w, np.float64).reshape(-1)[:9].reshape(3, 3)
    out = np.zeros((H, W, 3))
    for gy in range(H):
        for gx in range(W):
            tx, ty = gx // tile_size, gy // tile_size
            cand = [i for i in range(len(mean2d)) if tl[i, 0] <= tx <= br[i, 0] and tl[i, 1] <= ty <= br[i, 1]]
            cand.sort(key=lambda i: (int(dbits[i]), i))
            px = float(topleft[0]) + gx * float(psx)
            py = float(topleft[1]) + gy * float(psy)
            d = c2 @ np.array([px, py, 1.0])
            d = d / np.linalg.norm(d)
            Y = _sh_basis_f64(d, C)
            T = 1.0
            o = np.zeros(3)
            for i in cand:
                if T < thresh:
                    break
                a = min(alpha[i], 0.99)
                c0, c1, c2_, c3 = cov2d[i]
                det = c0 * c3 - c1 * c2_
                x, y = px - mean2d[i, 0], py - mean2d[i, 1]
                r = ((x * c3 - y * c2_) * x + (-x * c1 + y * c0) * y) / det
                if r < 0:
                    r = 1000.0
                G = np.exp(-0.5 * r)
                if a * G < 1.0 / 255.0:
                    continue
                w = a * T * G
                col = 1.0 / (1.0 + np.exp(-(sh[i] @ Y)))
                o += w * col
                T *= 1 - a * G
            out[gy, gx] = o
    return out


def _sh_basis_f64(d, C):
    """Real SH basis in float64 from the closed forms (degree <= 3)."""
    x, y, z = d
    pi = np.pi
    k0 = 1 / (2 * np.sqrt(pi))
    k1 = np.sqrt(3) / (2 * np.sqrt(pi))
    k2a = np.sqrt(15) / (2 * np.sqrt(pi))
    k2b = np.sqrt(5) / (4 * np.sqrt(pi))
    k2c = np.sqrt(15) / (4 * np.sqrt(pi))
    Y = [k0]
    if C > 1:
        Y += [-k1 * y, k1 * z, -k1 * x]
    if C > 2:
        Y += [k2a * x * y, -k2a * y * z, k2b * (3 * z * z - 1), -k2a * x * z, k2c * (x * x - y * y)]
    if C > 3:
        Y += [np.sqrt(70) / (8 * np.sqrt(pi)) * y * (-3 * x * x + y * y),
              np.sqrt(105) / (2 * np.sqrt(pi)) * x * y * z,
              np.sqrt(42) / (8 * np.sqrt(pi)) * y * (1 - 5 * z * z),
              np.sqrt(7) / (4 * np.sqrt(pi)) * z * (5 * z * z - 3),
              np.sqrt(42) / (8 * np.sqrt(pi)) * x * (1 - 5 * z * z),
              np.sqrt(105) / (4 * np.sqrt(pi)) * z * (x * x - y * y),
              np.sqrt(70) / (8 * np.sqrt(pi)) * x * (-x * x + 3 * y * y)]
    return np.array(Y[: C * C])


def composite_f64(mean2d, cov2d, sh, alpha, start, end, ids, topleft, c2w, psx, psy, H, W, C, thresh,
                  tile_size=16, decisions=None):
    """Float64 compositor over the binned lists, optionally with the skip /
    break decisions forced to a given set (so that central differences never
    cross a decision boundary).  Returns the image [H*W*3] in float64 and the
    decisions it used: dict pixel -> list of composited list positions."""
    mean2d = np.asarray(mean2d, np.float64).reshape(-1, 2)
    cov2d = np.asarray(cov2d, np.float64).reshape(-1, 4)
    V = len(mean2d)
    sh = np.asarray(sh, np.float64).reshape(V, 3, -1)
    alpha = np.asarray(alpha, np.float64).reshape(-1)
    c2 = np.asarray(c2w, np.float64).reshape(-1)[:9].reshape(3, 3)
    ntw = -(-W // tile_size)
    out = np.zeros(H * W * 3)
    used = {} if decisions is None else decisions
    for gy in range(H):
        for gx in range(W):
            tile = (gy // tile_size) * ntw + gx // tile_size
            s0, e0 = int(start[tile]), int(end[tile])
            if s0 < 0 or e0 - s0 == 0:
                continue
            px = float(f32(topleft[0]) + f32(gx) * f32(psx))
            py = float(f32(topleft[1]) + f32(gy) * f32(psy))
            d = c2 @ np.array([px, py, 1.0])
            Y = _sh_basis_f64(d / np.linalg.norm(d), C)
            pix = gy * W + gx
            T, o = 1.0, np.zeros(3)
            chosen = []
            for k in range(s0, e0):
                g = int(ids[k])
                a = min(alpha[g], 0.99)
                c0, c1, c2_, c3 = cov2d[g]
                det = c0 * c3 - c1 * c2_
                x, y = px - mean2d[g, 0], py - mean2d[g, 1]
                r = ((x * c3 - y * c2_) * x + (-x * c1 + y * c0) * y) / det
                G = np.exp(-0.5 * r) if r >= 0 else np.exp(-500.0)
                if decisions is None:
                    if T < thresh:
                        break
                    if a * G < 1.0 / 255.0:
                        continue
                    chosen.append(k)
                elif k not in used[pix]:
                    continue
                w = a * T * G
                col = 1.0 / (1.0 + np.exp(-(sh[g] @ Y)))
                o += w * col
                T *= 1 - a * G
            if decisions is None:
                used[pix] = set(chosen)
            out[3 * pix: 3 * pix + 3] = o
    return out, used


# --------------------------------------------------------------------------
# Adaptive density control (gs/sh_renderer.py:423-661), numpy float32
# --------------------------------------------------------------------------

def grad_stats(acc, grad, index=None, use_max=True):
    """update_grads (sh_renderer.py:602-623): acc[i] = max(acc[i], |g|) or
    acc[i] + |g|; index maps rows of grad to Gaussians (2D variant)."""
    acc = np.array(acc, np.float32, copy=True)
    g = np.asarray(grad, np.float32)
    nrm = np.sqrt((g * g).sum(-1, dtype=np.float32)).astype(np.float32)
    idx = np.arange(len(g)) if index is None else np.asarray(index)
    acc[idx] = np.maximum(acc[idx], nrm) if use_max else (acc[idx] + nrm).astype(np.float32)
    return acc


def split_clone(mean, qvec, log_s, logit, sh, acc, cnt, use_mean, grad_thresh, scale_thresh, shrink, noise):
    """split_gaussians (sh_renderer.py:426-534) with the Gaussian noise given:
    returns the five new arrays and (n_unsplit, n_clone, n_split)."""
    mean, qvec, log_s, logit, sh = (np.asarray(x, np.float32) for x in (mean, qvec, log_s, logit, sh))
    acc = np.asarray(acc, np.float32)
    grad = acc / (np.asarray(cnt).astype(np.float32) + np.float32(1e-5)) if use_mean else acc
    mask = grad > np.float32(grad_thresh)                                   # :435 / :437
    svec = expf(log_s).reshape(log_s.shape)                                 # self.svec (exp activation)
    split = mask & (svec > np.float32(scale_thresh)).any(-1)               # :450-452
    clone = mask & ~split                                                   # :454
    k = int(split.sum())
    s_rep = np.concatenate([svec[split]] * 2)                               # .repeat(2, 1)
    qs = np.concatenate([qvec[split]] * 2)
    R = np.stack([quat_to_rot(q) for q in qs]) if len(qs) else np.zeros((0, 3, 3), np.float32)
    gn = (np.asarray(noise, np.float32).reshape(2 * k, 3) * s_rep).astype(np.float32)
    # einsum("bij, bj -> bi", R^T, gn): (R^T gn)_a = sum_b R[b, a] gn[b]
    rt_gn = (R[:, 0, :] * gn[:, 0:1] + R[:, 1, :] * gn[:, 1:2] + R[:, 2, :] * gn[:, 2:3]).astype(np.float32)
    new_split_mean = (np.concatenate([mean[split]] * 2) + rt_gn).astype(np.float32)
    keep = ~split
    out = []
    for a, sp in ((mean, new_split_mean), (qvec, None), (log_s, None), (logit, None), (sh, None)):
        parts = [a[keep], a[clone]]
        if sp is not None:
            parts.append(sp)
        elif a is log_s:
            parts.append(np.log(s_rep / np.float32(shrink)).astype(np.float32))  # svec_inv_act(s / shrink)
        else:
            parts.append(np.concatenate([a[split]] * 2))
        out.append(np.concatenate(parts).astype(np.float32))
    return out, (int(keep.sum()), int(clone.sum()), k)


def prune_alpha(mean, qvec, log_s, logit, sh, alpha_thresh):
    """remove_low_alpha_gaussians (sh_renderer.py:536-558)."""
    lg = np.asarray(logit, np.float32)
    _, a = activate(np.zeros((lg.size, 3), np.float32), lg)
    m = a >= np.float32(alpha_thresh)
    return [np.asarray(x, np.float32)[m] for x in (mean, qvec, log_s, logit, sh)], int(m.sum())


# --------------------------------------------------------------------------
# Photometric loss with the SSIM term (utils/loss.py:14-21).  ssim_loss is
# kornia.losses.ssim_loss (third party, unpinned in requirements.txt, absent
# here), restated from kornia's published algorithm (kornia.metrics.ssim):
# 1-D Gaussian weights exp(-x^2 / (2 * 1.5^2)) for x = -win//2 .. win//2 in
# float32, normalised by their sum, 2-D window = outer product; filter2d =
# reflect padding by win//2 + correlation; C1 = (0.01)^2, C2 = (0.03)^2;
# SSIM = (2 mu1 mu2 + C1)(2 s12 + C2) / ((mu1^2 + mu2^2 + C1)(s11 + s22 + C2)
# + 1e-12); ssim_loss = mean((1 - SSIM) / 2).  Two independent restatements:
# torch float64 (pad + conv2d, autograd for the gradient) and numpy float64
# with scipy.ndimage.correlate1d in "mirror" mode (= torch's "reflect").
# --------------------------------------------------------------------------

SSIM_C1 = 0.01 ** 2
SSIM_C2 = 0.03 ** 2


def ssim_window1d(win, sigma=1.5):
    """kornia's get_gaussian_kernel1d(win, sigma), float32."""
    x = np.arange(win, dtype=np.float32) - np.float32(win // 2)
    g = np.exp(-(x * x) / np.float32(2.0 * sigma * sigma)).astype(np.float32)
    return (g / g.sum(dtype=np.float32)).astype(np.float32)


def ssim_map_scipy(img1, img2, win):
    """SSIM per pixel and channel of [H, W, C] images, float64, scipy filters."""
    from scipy import ndimage

    w = ssim_window1d(win).astype(np.float64)
    a = np.asarray(img1, np.float64)
    b = np.asarray(img2, np.float64)

    def filt(x):
        x = ndimage.correlate1d(x, w, axis=0, mode="mirror")
        return ndimage.correlate1d(x, w, axis=1, mode="mirror")

    mu1, mu2 = filt(a), filt(b)
    s11 = filt(a * a) - mu1 * mu1
    s22 = filt(b * b) - mu2 * mu2
    s12 = filt(a * b) - mu1 * mu2
    num = (2 * mu1 * mu2 + SSIM_C1) * (2 * s12 + SSIM_C2)
    den = (mu1 * mu1 + mu2 * mu2 + SSIM_C1) * (s11 + s22 + SSIM_C2)
    return num / (den + 1e-12)


def ssim_map_t(img1, img2, win):
    """The same in torch float64 (differentiable): [H, W, C] -> [H, W, C]."""
    import torch
    import torch.nn.functional as F

    R = win // 2
    w1 = torch.from_numpy(ssim_window1d(win).astype(np.float64))
    k2 = (w1[:, None] * w1[None, :])[None, None]  # [1, 1, win, win]
    a = img1.permute(2, 0, 1)[:, None]             # [C, 1, H, W]
    b = img2.permute(2, 0, 1)[:, None]

    def filt(x):
        return F.conv2d(F.pad(x, (R, R, R, R), mode="reflect"), k2)

    mu1, mu2 = filt(a), filt(b)
    s11 = filt(a * a) - mu1 * mu1
    s22 = filt(b * b) - mu2 * mu2
    s12 = filt(a * b) - mu1 * mu2
    num = (2 * mu1 * mu2 + SSIM_C1) * (2 * s12 + SSIM_C2)
    den = (mu1 * mu1 + mu2 * mu2 + SSIM_C1) * (s11 + s22 + SSIM_C2)
    return (num / (den + 1e-12))[:, 0].permute(1, 2, 0)


def photometric_ssim_loss(out, gt, win=11, ssim_mult=0.2, l2=False):
    """(loss, d loss / d out) of get_loss_fn in float64 (autograd over
    ssim_map_t); images [H, W, C]."""
    import torch

    o = torch.tensor(np.asarray(out, np.float64), requires_grad=True)
    g = torch.tensor(np.asarray(gt, np.float64))
    ssim_l = ((1.0 - ssim_map_t(o, g, win)) / 2.0).mean()
    base = ((o - g) ** 2).mean() if l2 else (o - g).abs().mean()
    loss = ssim_mult * ssim_l + (1.0 - ssim_mult) * base
    loss.backward()
    return float(loss.detach()), o.grad.numpy()


def reflect_filter_matrix(n, win):
    """Dense [n, n] matrix of one axis of "reflect pad by win//2, correlate
    with the 1-D window": row p = weights of input q in output p."""
    R = win // 2
    w = ssim_window1d(win).astype(np.float64)
    M = np.zeros((n, n))
    for p in range(n):
        for d in range(-R, R + 1):
            r = p + d
            q = -r if r < 0 else (2 * (n - 1) - r if r > n - 1 else r)
            M[p, q] += w[d + R]
    return M
