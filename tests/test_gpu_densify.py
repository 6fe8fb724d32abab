"""Adaptive density control on gfx950 against the oracle's restatement of
gs/sh_renderer.py:423-661 (same noise): copied rows bit-exact, sampled
means and shrunk scales to float rounding."""
import numpy as np
import pytest
import torch

from gaussian_splatting_3d_amd import densify

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _params(N, seed):
    g = torch.Generator().manual_seed(seed)
    mean = torch.randn(N, 3, generator=g)
    q = torch.randn(N, 4, generator=g)
    log_s = torch.randn(N, 3, generator=g) * 0.8 - 4.5
    logit = torch.randn(N, generator=g) * 2
    sh = torch.randn(N, 3, 16, generator=g)
    return [t.to(DEV).contiguous() for t in (mean, q, log_s, logit, sh)]


@pytest.mark.parametrize("reduction", ["mean", "max"])
def test_stats_split_clone_prune_match_oracle(orc, reduction):
    N = 5000
    p = _params(N, 3)
    cfg = densify.DensifyConfig(split_reduction=reduction, split_type="mean_grad")
    dc = densify.DensityControl(N, DEV, cfg)
    g = torch.Generator().manual_seed(4)
    acc = np.zeros(N, np.float32)
    cnt = np.zeros(N, np.int32)
    for step in range(3):
        grad = (torch.randn(N, 3, generator=g) * 2e-4).to(DEV)
        vis = (torch.rand(N, generator=g) > 0.3).to(DEV)
        dc.count_visible(vis)
        dc.update_grads(grad)
        acc = orc.grad_stats(acc, grad.cpu().numpy(), use_max=reduction == "max")
        cnt += vis.cpu().numpy().astype(np.int32)
    assert np.array_equal(dc.cnt.cpu().numpy(), cnt)
    assert np.allclose(dc.grad_mean.cpu().numpy(), acc, rtol=1e-6, atol=0)
    dc.grad_mean = torch.from_numpy(acc).to(DEV)  # identical inputs for the decision
    host = [t.cpu().numpy() for t in p]
    k_noise = torch.Generator(device=DEV).manual_seed(5)
    out, got_counts = dc.split_gaussians(*p, generator=k_noise)
    k = got_counts[2]
    noise = torch.randn(2 * k, 3, device=DEV, generator=torch.Generator(device=DEV).manual_seed(5)).cpu().numpy()
    want, counts = orc.split_clone(*host, acc, cnt, reduction == "mean", cfg.pos_grad_thresh, cfg.split_scale_thresh,
                                   cfg.scale_shrink_factor, noise)
    assert tuple(got_counts) == counts and counts[1] > 0 and counts[2] > 0
    got = [t.cpu().numpy() for t in out]
    n_copy = counts[0] + counts[1]
    for a, b in zip(got, want):
        assert a.shape == b.shape
        assert np.array_equal(a[:n_copy], b[:n_copy])
    for idx in (1, 3, 4):  # qvec, logit, sh of the split rows are copies
        assert np.array_equal(got[idx][n_copy:], want[idx][n_copy:])
    assert np.allclose(got[0][n_copy:], want[0][n_copy:], rtol=1e-6, atol=1e-7)
    assert np.allclose(got[2][n_copy:], want[2][n_copy:], rtol=1e-6, atol=1e-6)
    # pruning by opacity
    pruned = dc.remove_low_alpha(*out)
    w2, kept = orc.prune_alpha(*got, cfg.alpha_thresh)
    assert pruned[0].shape[0] == kept
    for a, b in zip(pruned, w2):
        assert np.array_equal(a.cpu().numpy(), b)
