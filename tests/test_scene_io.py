"""Checkpoint -> resident scene (CPU): reference key layout round trip,
validation, and checkpoints whose config object names classes and
functions the loader must not import or call."""
import os
import pickle
import sys
import types

import pytest
import torch

from gaussian_splatting_3d_amd import scene_io


def _params(N=50, C=4, seed=0):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(N, 3, generator=g), torch.randn(N, 4, generator=g), torch.randn(N, 3, generator=g),
            torch.randn(N, 3, C * C, generator=g), torch.randn(N, generator=g))


def test_round_trip_reference_keys(tmp_path):
    mean, q, s, sh, a = _params()
    p = tmp_path / "ckpt" / "scene.pt"
    scene_io.save_checkpoint(p, mean, q, s, sh, a, cfg={"sh_order": 4, "lr": 1e-3})
    sc = scene_io.load_checkpoint(p, device="cpu")
    assert sc.N == 50 and sc.C == 4
    for got, want in zip(sc.params(), (mean, q, s, a, sh)):
        assert torch.equal(got, want) and got.is_contiguous() and got.dtype == torch.float32
    assert scene_io.load_checkpoint(p, device="cpu", C=2).C == 2
    with pytest.raises(ValueError):
        scene_io.load_checkpoint(p, device="cpu", C=5)


def test_reference_style_dict_without_cfg(tmp_path):
    mean, q, s, sh, a = _params(C=1)
    p = tmp_path / "ref.pt"
    torch.save({"mean": mean, "qvec": q, "svec_before_activation": s, "sh_coeffs": sh,
                "alpha_before_activation": a.unsqueeze(-1), "N": 50}, p)
    sc = scene_io.load_checkpoint(p, device="cpu")
    assert sc.C == 1 and tuple(sc.alpha_before_activation.shape) == (50,)


@pytest.mark.parametrize("bad", ["N", "shape", "sh"])
def test_validation(tmp_path, bad):
    mean, q, s, sh, a = _params()
    sd = {"mean": mean, "qvec": q, "svec_before_activation": s, "sh_coeffs": sh, "alpha_before_activation": a,
          "N": 50}
    if bad == "N":
        sd["N"] = 49
    elif bad == "shape":
        sd["qvec"] = q[:, :3]
    else:
        sd["sh_coeffs"] = torch.randn(50, 3, 5)
    p = tmp_path / "bad.pt"
    torch.save(sd, p)
    with pytest.raises(ValueError):
        scene_io.load_checkpoint(p, device="cpu")


_EXECUTED = []


def _mark_executed():
    _EXECUTED.append(1)
    return {}


class _Config:  # a config object whose restoration would call a function
    def __reduce__(self):
        return (_mark_executed, ())


def test_config_needing_code_is_not_executed(tmp_path):
    mean, q, s, sh, a = _params()
    p = tmp_path / "unsafe.pt"
    torch.save({"mean": mean, "qvec": q, "svec_before_activation": s, "sh_coeffs": sh,
                "alpha_before_activation": a, "N": 50, "cfg": _Config()}, p, pickle_module=pickle)
    sc = scene_io.load_checkpoint(p, device="cpu")
    assert not _EXECUTED
    assert torch.equal(sc.mean, mean) and sc.C == 4
    assert isinstance(scene_io.read_checkpoint(p)["cfg"], scene_io._Inert)


def _fake_config_library():
    """A config library registered under a module name only for pickling:
    DictConfig with nested metadata objects, and Hostile, whose restoration
    would run a shell command.  Returns (classes, unregister)."""
    pkg = types.ModuleType("fake_cfg_lib")
    mod = types.ModuleType("fake_cfg_lib.dictconfig")
    pkg.dictconfig = mod

    class Metadata:
        def __init__(self):
            self.ref_type = dict
            self.flags = {"readonly": False}

    class DictConfig:
        def __init__(self):
            self._metadata = Metadata()
            self._content = {"sh_order": 4, "lr": {"mean": 1e-3}, "tile_size": 16}

    class Hostile:
        flag = None

        def __reduce__(self):
            return (os.system, (f"touch {Hostile.flag}",))

    for c in (Metadata, DictConfig, Hostile):
        c.__module__ = mod.__name__
        c.__qualname__ = c.__name__
        setattr(mod, c.__name__, c)
    sys.modules[pkg.__name__] = pkg
    sys.modules[mod.__name__] = mod

    def unregister():
        del sys.modules[mod.__name__]
        del sys.modules[pkg.__name__]

    return (DictConfig, Hostile), unregister


def test_reference_checkpoint_with_hydra_like_config(tmp_path):
    """A checkpoint laid out as SHRenderer.save writes it: the five tensors,
    N and a config object of a class from a config library, with nested
    metadata objects.  The library is not importable when the file is read:
    the loader binds its names to inert stand-ins and never imports them."""
    (DictConfig, _), unregister = _fake_config_library()
    mean, q, s, sh, a = _params(N=64, C=3)
    p = tmp_path / "ckpt.pt"
    try:
        torch.save({"mean": mean, "qvec": q, "svec_before_activation": s, "sh_coeffs": sh,
                    "alpha_before_activation": a, "N": 64, "cfg": DictConfig()}, p)
    finally:
        unregister()
    names = torch.serialization.get_unsafe_globals_in_checkpoint(str(p))
    assert "fake_cfg_lib.dictconfig.DictConfig" in names and "fake_cfg_lib.dictconfig.Metadata" in names
    sc = scene_io.load_checkpoint(p, device="cpu")
    assert "fake_cfg_lib" not in sys.modules
    assert sc.N == 64 and sc.C == 3
    for got, want in zip(sc.params(), (mean, q, s, a, sh)):
        assert torch.equal(got, want)
    assert isinstance(scene_io.read_checkpoint(p)["cfg"], scene_io._Inert)


def test_checkpoint_naming_a_shell_call_is_refused_unexecuted(tmp_path):
    (DictConfig, Hostile), unregister = _fake_config_library()
    Hostile.flag = tmp_path / "ran"
    cfg = DictConfig()
    cfg._content["hook"] = Hostile()
    mean, q, s, sh, a = _params()
    p = tmp_path / "hostile.pt"
    try:
        torch.save({"mean": mean, "qvec": q, "svec_before_activation": s, "sh_coeffs": sh,
                    "alpha_before_activation": a, "N": 50, "cfg": cfg}, p)
    finally:
        unregister()
    with pytest.raises(RuntimeError, match="refuses"):
        scene_io.load_checkpoint(p, device="cpu")
    assert not (tmp_path / "ran").exists()
