"""Render server: resident scenes, continuous batching of camera requests.

Python face of the native server in csrc/serve.hip (C ABI gsmi_server_*).
The reference's viewer renders each client's camera with one
``SHRenderer.forward`` call under ``no_grad``, one client after another
(utils/viewer/viser_viewer.py:118-158).  ``RenderServer`` keeps the scenes in
HBM, queues requests from any thread, and native worker threads (one HIP
stream each) batch pending requests, of several scenes when their SH layout
matches, into one multi-view launch sequence per batch under a view and
pixel budget, several batches in flight at once; per-batch memory comes from
a fixed device arena, and a batch admits as many views as their key/sort
buffers fit.  Every image equals ``rasterize.forward_view`` for that camera.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr


@dataclass
class Completion:
    request_id: int
    t_submit_ms: float
    t_done_ms: float
    status: int
    batch_views: int

    @property
    def latency_ms(self):
        return self.t_done_ms - self.t_submit_ms


def make_view_record(c2w, camera, tile_size=16):
    """gsmi_view bytes (256) for one camera, built by the native helper."""
    L = _lib.load()
    c = np.ascontiguousarray(
        (c2w.detach().cpu().numpy() if isinstance(c2w, torch.Tensor) else np.asarray(c2w)), np.float32
    ).reshape(12)
    rec = np.zeros(64, np.float32)
    check(L.gsmi_make_view(c.ctypes.data_as(ctypes.c_void_p), float(camera.fx), float(camera.fy),
                           float(camera.cx), float(camera.cy), int(camera.w), int(camera.h),
                           float(camera.near_plane), float(camera.far_plane), tile_size,
                           rec.ctypes.data_as(ctypes.c_void_p)), "make_view")
    return rec


class RenderServer:
    """Continuous-batching renderer on one GPU.

    arena_bytes: device memory for all in-flight batches (allocated once);
    max_views: views per batch; n_slots: worker threads, i.e. batches in
    flight at once; max_batch_pixels: pixel budget of a batch (None:
    unlimited); queue_grid: workgroups of the work-queue compositing launch
    (0: one workgroup per tile)."""

    def __init__(self, device="cuda:0", arena_bytes=16 << 30, max_views=32, n_slots=2, cfg=None, exact=False,
                 max_batch_pixels=None, queue_grid=0):
        from .scenes import RasterConfig

        self._L = _lib.load()
        self.device = torch.device(device)
        cfg = cfg or RasterConfig()
        if cfg.bg or cfg.skip_frustum_culling or cfg.tile_size != 16:
            raise RuntimeError("the render server supports the default tile size, culling on, black background")
        torch.cuda.init()
        h = ctypes.c_void_p()
        check(self._L.gsmi_server_create(self.device.index or 0, int(arena_bytes), int(max_views), int(n_slots),
                                          float(cfg.tile_culling_radius), float(cfg.frustum_culling_radius),
                                          float(cfg.T_thresh), 1 if exact else 0, ctypes.byref(h)), "server_create")
        self._h = h
        check(self._L.gsmi_server_configure(h, int(max_batch_pixels) if max_batch_pixels else 1 << 40,
                                            int(queue_grid)), "server_configure")
        self._scenes = []  # keep the registered tensors alive
        self._outs = {}    # request id -> output tensor (alive until completed)

    def add_scene(self, mean, qvec, svec_before_activation, alpha_before_activation, sh_coeffs, C):
        t = [x.to(self.device).contiguous().float() for x in
             (mean, qvec, svec_before_activation, alpha_before_activation, sh_coeffs)]
        sid = ctypes.c_int()
        check(self._L.gsmi_server_add_scene(self._h, int(t[0].shape[0]), int(C), int(t[4].shape[-1]), ptr(t[0]),
                                            ptr(t[1]), ptr(t[2]), ptr(t[3]), ptr(t[4]), ctypes.byref(sid)),
              "server_add_scene")
        self._scenes.append(t)
        return sid.value

    def add_resident(self, scene):
        """Registers a scene_io.ResidentScene (e.g. from load_checkpoint)."""
        return self.add_scene(*scene.params(), scene.C)

    def submit(self, scene_id, c2w, camera, out=None):
        """Queues one camera; returns (request id, output tensor [H,W,3]).
        The tensor is filled when the request completes (see poll)."""
        if out is None:
            out = torch.empty(camera.h, camera.w, 3, dtype=torch.float32, device=self.device)
        c = np.ascontiguousarray(
            (c2w.detach().cpu().numpy() if isinstance(c2w, torch.Tensor) else np.asarray(c2w)), np.float32
        ).reshape(12)
        rid = ctypes.c_int64()
        check(self._L.gsmi_server_submit(self._h, int(scene_id), c.ctypes.data_as(ctypes.c_void_p),
                                         float(camera.fx), float(camera.fy), float(camera.cx), float(camera.cy),
                                         int(camera.w), int(camera.h), float(camera.near_plane),
                                         float(camera.far_plane), ptr(out), ctypes.byref(rid)), "server_submit")
        self._outs[rid.value] = out
        return rid.value, out

    def poll(self, max_n=4096):
        """Completed requests since the last poll."""
        ids = np.zeros(max_n, np.int64)
        ts = np.zeros(max_n, np.float64)
        td = np.zeros(max_n, np.float64)
        stt = np.zeros(max_n, np.int32)
        bv = np.zeros(max_n, np.int32)
        n = ctypes.c_int()
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        check(self._L.gsmi_server_poll(self._h, vp(ids), vp(ts), vp(td), vp(stt), vp(bv), max_n, ctypes.byref(n)),
              "server_poll")
        res = [Completion(int(ids[i]), float(ts[i]), float(td[i]), int(stt[i]), int(bv[i])) for i in range(n.value)]
        for c in res:
            self._outs.pop(c.request_id, None)
        return res

    def drain(self, timeout_ms=600_000.0):
        check(self._L.gsmi_server_drain(self._h, float(timeout_ms)), "server_drain")

    def stats(self):
        s = np.zeros(7, np.int64)
        check(self._L.gsmi_server_stats_ex(self._h, s.ctypes.data_as(ctypes.c_void_p), 7), "server_stats")
        return dict(batches=int(s[0]), views=int(s[1]), max_batch=int(s[2]), splits=int(s[3]), failed=int(s[4]),
                    arena_bytes=int(s[5]), mixed_batches=int(s[6]))

    def close(self):
        if self._h:
            check(self._L.gsmi_server_destroy(self._h), "server_destroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
