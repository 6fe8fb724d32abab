"""The photometric loss with the SSIM term (get_loss_fn, utils/loss.py:14-21)
on gfx950 against the float64 oracle: loss within 1e-5 relative, gradient
within 1e-4 (max-norm relative), on ragged sizes, several windows, both base
terms, one to four channels, and a full 1920x1080 image."""
import numpy as np
import pytest
import torch

from gaussian_splatting_3d_amd import training

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _imgs(H, W, C, seed):
    rng = np.random.default_rng(seed)
    a = rng.random((H, W, C)).astype(np.float32)
    b = np.clip(a + rng.normal(0.0, 0.2, (H, W, C)), 0.0, 1.0).astype(np.float32)
    return a, b


def _check(orc, H, W, C, win, mult, l2, seed=0, rtol_g=1e-4):
    a, b = _imgs(H, W, C, seed)
    loss, grad = training.photometric_loss(torch.from_numpy(a).to(DEV), torch.from_numpy(b).to(DEV),
                                           "l2" if l2 else "l1", ssim_mult=mult, ssim_win=win)
    want_l, want_g = orc.photometric_ssim_loss(a, b, win, mult, l2)
    assert abs(float(loss) - want_l) <= 1e-5 * abs(want_l)
    g = grad.cpu().numpy().astype(np.float64)
    assert np.abs(g - want_g).max() <= rtol_g * np.abs(want_g).max()


@pytest.mark.parametrize("H,W,C,win,mult,l2", [
    (37, 53, 3, 11, 0.2, False), (37, 53, 3, 11, 0.2, True), (64, 48, 3, 7, 0.5, False),
    (12, 15, 1, 11, 0.2, False), (16, 16, 4, 3, 1.0, False), (21, 90, 2, 21, 0.2, True)])
def test_ssim_loss_matches_oracle(orc, H, W, C, win, mult, l2):
    _check(orc, H, W, C, win, mult, l2, seed=H * W + win)


def test_ssim_loss_full_hd(orc):
    _check(orc, 1080, 1920, 3, 11, 0.2, False, seed=5)


def test_zero_ssim_weight_is_the_base_loss():
    a, b = _imgs(33, 47, 3, 9)
    ta, tb = torch.from_numpy(a).to(DEV), torch.from_numpy(b).to(DEV)
    l0, g0 = training.photometric_loss(ta, tb, "l1")
    l1, g1 = training.photometric_loss(ta, tb, "l1", ssim_mult=1e-30)
    assert abs(float(l0) - float(l1)) <= 1e-6 * float(l0)
    assert torch.allclose(g0, g1, rtol=1e-5, atol=1e-12)


def test_argument_errors():
    a = torch.rand(8, 8, 3, device=DEV)
    with pytest.raises(RuntimeError):
        training.photometric_loss(a, a, ssim_mult=0.2, ssim_win=10)  # even window
    with pytest.raises(RuntimeError):
        training.photometric_loss(a, a, ssim_mult=0.2, ssim_win=21)  # win // 2 >= 8
    with pytest.raises(RuntimeError):
        training.photometric_loss(a.reshape(-1), a.reshape(-1), ssim_mult=0.2)
