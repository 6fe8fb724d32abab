"""Scene checkpoints -> resident structure-of-arrays scenes.

The reference saves ``SHRenderer`` state with ``torch.save`` of a dict
(gs/sh_renderer.py:663-676): ``mean [N,3]``, ``qvec [N,4]``,
``svec_before_activation [N,3]``, ``sh_coeffs [N,3,C^2]``,
``alpha_before_activation [N]``, ``N`` and the Hydra ``cfg``; ``load``
(:678-708) restores the five tensors as parameters.

Here a checkpoint is read with torch's ``weights_only`` unpickler only, which
builds tensors, numbers, strings and plain containers and nothing else.  A
file that also names other classes or functions (the reference always stores
its Hydra/OmegaConf ``cfg``) is read the same way after every such name,
found by disassembling the pickle without running it, is bound to an inert
stand-in (``_Inert``): the stand-in takes any arguments and state and keeps
none, so the config comes back as an empty placeholder and nothing the file
names is imported or called.  Only the five tensors and ``N`` are used.
``save_checkpoint`` writes the reference's keys with the config as a plain
dict.
"""
from __future__ import annotations

import math
import os
import pickle
from dataclasses import dataclass

import torch

KEYS = ("mean", "qvec", "svec_before_activation", "sh_coeffs", "alpha_before_activation")


@dataclass
class ResidentScene:
    """Raw SHRenderer parameters as contiguous float32 device tensors (the
    layout every kernel and the render server read)."""

    mean: torch.Tensor
    qvec: torch.Tensor
    svec_before_activation: torch.Tensor
    alpha_before_activation: torch.Tensor
    sh_coeffs: torch.Tensor
    C: int

    @property
    def N(self):
        return int(self.mean.shape[0])

    def params(self):
        return [self.mean, self.qvec, self.svec_before_activation, self.alpha_before_activation, self.sh_coeffs]

    def nbytes(self):
        return sum(t.numel() * t.element_size() for t in self.params())


def _validate(sd):
    missing = [k for k in KEYS if k not in sd]
    if missing:
        raise ValueError(f"checkpoint lacks {missing}")
    N = int(sd["mean"].shape[0])
    shapes = {"mean": (N, 3), "qvec": (N, 4), "svec_before_activation": (N, 3)}
    for k, s in shapes.items():
        if tuple(sd[k].shape) != s:
            raise ValueError(f"{k} has shape {tuple(sd[k].shape)}, expected {s}")
    if tuple(sd["alpha_before_activation"].shape) not in ((N,), (N, 1)):
        raise ValueError(f"alpha_before_activation has shape {tuple(sd['alpha_before_activation'].shape)}")
    sh = sd["sh_coeffs"]
    if sh.dim() != 3 or sh.shape[0] != N or sh.shape[1] != 3:
        raise ValueError(f"sh_coeffs has shape {tuple(sh.shape)}, expected [N, 3, C^2]")
    C = int(round(math.sqrt(sh.shape[2])))
    if C * C != sh.shape[2] or not 1 <= C <= 4:
        raise ValueError(f"sh_coeffs has {sh.shape[2]} coefficients per channel; expected C^2 with C in 1..4")
    if "N" in sd and int(sd["N"]) != N:  # the reference asserts the same (sh_renderer.py:706)
        raise ValueError(f"N = {int(sd['N'])} but mean has {N} rows")
    return N, C


class _Inert:
    """Stand-in for a class or function a checkpoint names but this loader
    does not run (e.g. the reference's Hydra config classes): constructing,
    calling or restoring it accepts anything and keeps nothing."""

    def __new__(cls, *args, **kwargs):
        return object.__new__(cls)

    def __init__(self, *args, **kwargs):
        pass

    def __setstate__(self, state):
        pass


def read_checkpoint(path):
    """The checkpoint's top-level object, read with torch's weights_only
    unpickler; globals outside its allow-list are bound to _Inert (they are
    listed by disassembly, never imported or called).  A file naming a
    module torch blocks outright (os, subprocess, ...) raises RuntimeError."""
    path = os.fspath(path)
    try:
        return torch.load(path, map_location="cpu", weights_only=True)
    except pickle.UnpicklingError:
        names = torch.serialization.get_unsafe_globals_in_checkpoint(path)
        if not names:
            raise
    try:
        with torch.serialization.safe_globals([(_Inert, name) for name in names]):
            return torch.load(path, map_location="cpu", weights_only=True)
    except pickle.UnpicklingError:  # torch blocks some modules (os, subprocess, ...) outright
        raise RuntimeError(f"{path}: the checkpoint names {names}; the loader refuses such a file even with "
                           "inert placeholders") from None


def load_checkpoint(path, device="cuda:0", C=None):
    """Reads a checkpoint (the reference's SHRenderer.save format, config
    included) without executing anything from it and returns a
    ResidentScene on `device`.  C (SH order used for rendering) defaults to
    the order the stored coefficients hold."""
    sd = read_checkpoint(path)
    if not isinstance(sd, dict):
        raise ValueError(f"{path}: expected a dict, found {type(sd).__name__}")
    N, C_store = _validate(sd)
    C = C_store if C is None else int(C)
    if not 1 <= C <= C_store:
        raise ValueError(f"SH order {C} exceeds the stored order {C_store}")
    t = {k: sd[k].detach().to(device=device, dtype=torch.float32).contiguous() for k in KEYS}
    return ResidentScene(t["mean"], t["qvec"], t["svec_before_activation"],
                         t["alpha_before_activation"].reshape(N).contiguous(), t["sh_coeffs"], C)


def save_checkpoint(path, mean, qvec, svec_before_activation, sh_coeffs, alpha_before_activation, cfg=None):
    """Writes the reference's checkpoint keys (gs/sh_renderer.py:663-676);
    cfg must be a plain dict (or None) so that the safe loader reads it."""
    if cfg is not None and not isinstance(cfg, dict):
        raise TypeError("cfg must be a plain dict")
    path = os.fspath(path)
    parent = os.path.dirname(path)
    if parent:
        os.makedirs(parent, exist_ok=True)
    sd = {"mean": mean.detach().cpu(), "qvec": qvec.detach().cpu(),
          "svec_before_activation": svec_before_activation.detach().cpu(), "sh_coeffs": sh_coeffs.detach().cpu(),
          "alpha_before_activation": alpha_before_activation.detach().cpu(), "N": int(mean.shape[0]), "cfg": cfg}
    torch.save(sd, path)
