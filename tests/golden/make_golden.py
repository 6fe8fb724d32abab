"""Writes the committed golden fixtures (run from the repo root):

  kat_gaussian_2d.json  — test/gaussian_test.py:1-35 inputs with the 2D
                          Gaussian value and gradients from torch autograd;
  mock_scene_sh.npz     — the reference's debug scene (gs/debug.py:40-57
                          "default" layout, sh_render_sanity_check camera and
                          settings at gs/debug.py:1051-1110) rendered by the
                          CPU oracle for SH orders 1 and 2: tile lists, the
                          count and float64 sum of non-zero output values and
                          every 97th non-zero value.

The reference code is not imported or run; the scene is retyped from its
definition.
"""
import json
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from gaussian_splatting_3d_amd.camera import get_c2w_from_up_and_look_at  # noqa: E402


def kat():
    mean = torch.tensor([0.1, 0.2], dtype=torch.float64, requires_grad=True)
    q = torch.tensor([0.3, 0.4], dtype=torch.float64)
    cov = torch.tensor([[0.5, 0.2], [0.2, 0.8]], dtype=torch.float64, requires_grad=True)
    d = q - mean
    G = torch.exp(-0.5 * d @ torch.inverse(cov) @ d)
    G.backward()
    return dict(mean=[0.1, 0.2], query=[0.3, 0.4], cov=[0.5, 0.2, 0.2, 0.8], G=G.item(),
                grad_mean=mean.grad.tolist(), grad_cov=cov.grad.reshape(-1).tolist())


def mock_scene():
    mean = np.array([[0.0, 0.0, 0.0], [0.1, 0.07, 0.0]], np.float32)
    qvec = np.array([[1.0, 0, 0, 0], [1.0, 0, 0, 0]], np.float32)
    log_s = (np.array([[1.0, 1.0, 0.5], [1.1, 0.5, 1.1]], np.float32) * np.float32(np.log(20e-3))).astype(np.float32)
    color = np.array([[0.01, 0.01, 0.99], [0.01, 0.99, 0.01]], np.float32)
    logit = np.array([10000.0, 10000.0], np.float32)
    c2w = get_c2w_from_up_and_look_at(np.array([0, 0, 1.0]), np.array([0, 0, 0.0]), np.array([1.0, 0, 0]))
    cam = O.Camera(961.22, 963.09, 648.38, 420.12, 1297, 840, 0.0, 1000.0)
    return mean, qvec, log_s, color, logit, c2w, cam


def mock_render(C):
    mean, qvec, log_s, color, logit, c2w, cam = mock_scene()
    sh_base = 0.28209479177387814
    sh = (np.log(color / (1 - color)) / sh_base).astype(np.float32)[:, :, None].repeat(C * C, -1)
    r = O.pipeline_forward(mean, qvec, log_s, logit, sh, c2w, cam, C, D=10.0, frustum_radius=4.0, thresh=1e-4)
    nz = np.nonzero(r["out"])[0].astype(np.int32)
    pick = nz[::97]
    return dict(P=np.int64(r["P"]), ids=r["ids"], start=r["start"], end=r["end"], n_nonzero=np.int64(nz.size),
                checksum=np.float64(r["out"].astype(np.float64).sum()), pix=pick, vals=r["out"][pick],
                argmax=np.int64(np.argmax(r["out"])), vmax=r["out"].max(), c2w=c2w, sh=sh)


if __name__ == "__main__":
    with open(os.path.join(HERE, "kat_gaussian_2d.json"), "w") as f:
        json.dump(kat(), f, indent=1)
    out = {}
    for C in (1, 2):
        for k, v in mock_render(C).items():
            out[f"C{C}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, "mock_scene_sh.npz"), **out)
    print("wrote", sorted(os.listdir(HERE)))
