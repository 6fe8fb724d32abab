"""GPU: gsmi_render_sh_backward_noalloc (tile-order scratch in a file-scope
device array) against gsmi_render_sh_backward with a caller workspace, both
through the C ABI, on two calls in a row, the second with a smaller view and
fewer Gaussians.  Gradients within 1e-3 (max-norm relative; both run the same
kernels, so in practice within float-atomic reordering).  Sizes beyond the
array and a missing saved state are refused before any launch."""
import numpy as np
import pytest
import torch

from gaussian_splatting_3d_amd import _lib, scenes
from gradmetrics import rel_err

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
P_ = _lib.ptr


def _scene(n, W, H, seed, C=4):
    g = torch.Generator().manual_seed(seed)
    mean = torch.randn(n, 3, generator=g) * 0.8
    ls = torch.randn(n, 3, generator=g) * 0.5 - 3.0
    q = torch.randn(n, 4, generator=g)
    q = q / q.norm(dim=-1, keepdim=True)
    logit = torch.randn(n, generator=g)
    sh = torch.randn(n, 3, C * C, generator=g) * 0.1
    sh[:, :, 0] = torch.randn(n, 3, generator=g) * 0.6
    c2w, cam = scenes.orbit_camera(4.0, 0.7, 90.0, 90.0, W, H, near=0.01, far=100.0, height=0.3)
    return scenes.Scene(mean, q, ls, logit, sh, C, [(c2w, cam)], f"noalloc{seed}")


def _setup(orc, sc):
    c2w, cam = sc.cameras[0]
    oc = orc.Camera(cam.fx, cam.fy, cam.cx, cam.cy, cam.w, cam.h, cam.near_plane, cam.far_plane)
    ref = orc.pipeline_forward(sc.mean, sc.qvec, sc.svec_before_activation, sc.alpha_before_activation,
                               sc.sh_coeffs, c2w, oc, sc.C)
    d = lambda x, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(x)).to(DEV, dt)  # noqa: E731
    t = dict(mean=d(ref["mean2d"]), cov=d(np.asarray(ref["cov2d"]).reshape(-1, 4)), sh=d(ref["sh_v"]),
             alpha=d(ref["a"][ref["idx"]]), start=d(ref["start"], torch.int32), end=d(ref["end"], torch.int32),
             ids=d(ref["ids"], torch.int32), topleft=d(ref["topleft"]), c2w=d(np.asarray(c2w, np.float32)))
    return ref, t


def _run(L, ref, t, C, gout, T, nd, out, noalloc, seed_grads=None):
    V, P = t["alpha"].numel(), t["ids"].numel()
    H, W, nth, ntw = ref["H"], ref["W"], ref["n_tiles_h"], ref["n_tiles_w"]
    g = [torch.zeros(V, 2, device=DEV), torch.zeros(V, 4, device=DEV), torch.zeros(V, 3, C * C, device=DEV),
         torch.zeros(V, device=DEV)]
    head = (V, P, P_(t["mean"]), P_(t["cov"]), P_(t["sh"]), 3 * C * C, C * C, None, P_(t["alpha"]), P_(t["start"]),
            P_(t["end"]), P_(t["ids"]), P_(out), P_(T), P_(nd), None, P_(gout), *[P_(x) for x in g],
            P_(t["topleft"]), P_(t["c2w"]), 16, nth, ntw, float(ref["psx"]), float(ref["psy"]), H, W, C, 1e-4,
            None, 0)
    st = _lib.stream_of()
    if noalloc:
        rc = L.gsmi_render_sh_backward_noalloc(*head, st)
    else:
        ws = torch.empty(L.gsmi_backward_order_workspace_size(nth * ntw), dtype=torch.uint8, device=DEV)
        rc = L.gsmi_render_sh_backward(*head, P_(ws), ws.numel(), st)
    assert rc == 0
    torch.cuda.synchronize()
    return g


def test_noalloc_backward_matches_workspace_entry(orc):
    L = _lib.load()
    st = _lib.stream_of()
    for sc in (_scene(6000, 160, 128, 21), _scene(2000, 96, 80, 22)):
        ref, t = _setup(orc, sc)
        V, P = t["alpha"].numel(), t["ids"].numel()
        H, W, nth, ntw = ref["H"], ref["W"], ref["n_tiles_h"], ref["n_tiles_w"]
        C = sc.C
        out = torch.zeros(H * W * 3, device=DEV)
        T = torch.empty(H * W, device=DEV)
        nd = torch.empty(H * W, dtype=torch.int32, device=DEV)
        assert L.gsmi_render_sh_forward(V, P, P_(t["mean"]), P_(t["cov"]), P_(t["sh"]), 3 * C * C, C * C, None,
                                        P_(t["alpha"]), P_(t["start"]), P_(t["end"]), P_(t["ids"]), P_(out),
                                        P_(t["topleft"]), P_(t["c2w"]), 16, nth, ntw, float(ref["psx"]),
                                        float(ref["psy"]), H, W, C, 1e-4, None, P_(T), P_(nd), None, 0, 0, st) == 0
        gout = torch.from_numpy(np.random.default_rng(3).random(H * W * 3).astype(np.float32) - 0.5).to(DEV)
        want = _run(L, ref, t, C, gout, T, nd, out, noalloc=False)
        got = _run(L, ref, t, C, gout, T, nd, out, noalloc=True)
        for a, b in zip(got, want):
            assert float(b.abs().max()) > 0
            assert rel_err(a.cpu(), b.cpu()) <= 1e-3


def test_noalloc_backward_refuses_before_launch():
    L = _lib.load()
    st = _lib.stream_of()
    dummy = torch.zeros(16, device=DEV)
    p = P_(dummy)
    args = [0, 0, p, p, p, 48, 16, None, p, p, p, p, p, p, p, None, p, p, p, p, p, p, p, 16]
    # 513 x 513 tiles exceeds the 262,144-tile array
    assert L.gsmi_render_sh_backward_noalloc(*args, 513, 513, 1.0, 1.0, 513 * 16, 513 * 16, 4, 1e-4, None, 0,
                                             st) == 4
    # no saved state
    args2 = list(args)
    args2[13] = None
    args2[14] = None
    assert L.gsmi_render_sh_backward_noalloc(*args2, 4, 4, 1.0, 1.0, 64, 64, 4, 1e-4, None, 0, st) == 1
