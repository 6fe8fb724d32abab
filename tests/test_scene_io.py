"""Checkpoint -> resident scene (CPU): reference key layout round trip,
validation, and refusal of files that would need arbitrary unpickling."""
import pickle

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


class _Config:  # stands in for a config object that needs unpickling
    def __reduce__(self):
        return (_mark_executed, ())


_EXECUTED = []


def _mark_executed():
    _EXECUTED.append(1)
    return {}


def test_refuses_objects_needing_unpickling(tmp_path):
    mean, q, s, sh, a = _params()
    p = tmp_path / "unsafe.pt"
    torch.save({"mean": mean, "qvec": q, "svec_before_activation": s, "sh_coeffs": sh,
                "alpha_before_activation": a, "N": 50, "cfg": _Config()}, p, pickle_module=pickle)
    with pytest.raises(RuntimeError, match="cannot be loaded without executing"):
        scene_io.load_checkpoint(p, device="cpu")
    assert not _EXECUTED
