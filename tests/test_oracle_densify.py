"""Pins the oracle's restatement of adaptive density control
(gs/sh_renderer.py:423-661) with hand-computed cases.  CPU only."""
import math

import numpy as np


def _scene():
    mean = np.array([[0, 0, 0], [1, 2, 3], [4, 5, 6], [7, 8, 9]], np.float32)
    q = np.array([[1, 0, 0, 0], [math.cos(math.pi / 4), 0, 0, math.sin(math.pi / 4)], [1, 0, 0, 0],
                  [1, 0, 0, 0]], np.float32)
    s = np.array([[0.005, 0.005, 0.005], [0.05, 0.01, 0.01], [0.01, 0.01, 0.01], [0.03, 0.001, 0.001]], np.float32)
    log_s = np.log(s).astype(np.float32)
    logit = np.array([0.5, -1.0, 2.0, -4.0], np.float32)
    sh = np.arange(4 * 3 * 4, dtype=np.float32).reshape(4, 3, 4)
    return mean, q, log_s, logit, sh


def test_split_clone_layout_and_transposed_rotation(orc):
    mean, q, log_s, logit, sh = _scene()
    # densify Gaussians 1 (big -> split), 2 (small -> clone), 3 (big -> split); 0 stays
    acc = np.array([1e-5, 1.0, 1.0, 1.0], np.float32)
    cnt = np.zeros(4, np.int32)
    noise = np.zeros((4, 3), np.float32)
    noise[0] = [1, 0, 0]  # first copy of Gaussian 1
    noise[3] = [0, 0, 2]  # second copy of Gaussian 3
    out, (n_unsplit, n_clone, k) = orc.split_clone(mean, q, log_s, logit, sh, acc, cnt, False, 2e-4, 0.02, 1.6, noise)
    assert (n_unsplit, n_clone, k) == (2, 1, 2)
    m, qq, ls, lg, shh = out
    assert m.shape == (2 + 1 + 4, 3)
    # unsplit block: Gaussians 0 and 2 in index order; then the clone of 2
    assert np.array_equal(m[:3], mean[[0, 2, 2]]) and np.array_equal(lg[:3], logit[[0, 2, 2]])
    # split rows: (1, 3, 1, 3) as .repeat(2, 1) lays them out
    assert np.array_equal(lg[3:], logit[[1, 3, 1, 3]]) and np.array_equal(shh[3:], sh[[1, 3, 1, 3]])
    # Gaussian 1 is rotated 90 degrees about z; the reference applies R^T:
    # offset R^T (0.05, 0, 0) = (0, -0.05, 0)   (R would give (0, +0.05, 0))
    assert np.allclose(m[3], mean[1] + [0, -0.05, 0], atol=1e-7)
    assert np.array_equal(m[4], mean[3]) and np.array_equal(m[5], mean[1])
    assert np.allclose(m[6], mean[3] + [0, 0, 2 * 0.001], atol=1e-7)
    # split scales shrink by 1.6, clones keep theirs
    assert np.allclose(np.exp(ls[3]), np.exp(log_s[1]) / 1.6, rtol=1e-6)
    assert np.array_equal(ls[2], log_s[2])


def test_mean_reduction_divides_by_count(orc):
    mean, q, log_s, logit, sh = _scene()
    acc = np.array([3e-4, 3e-4, 3e-4, 3e-4], np.float32)
    cnt = np.array([0, 1, 2, 3], np.int32)  # 3e-4 / (c + 1e-5): > 2e-4 only for c = 0, 1
    _, counts = orc.split_clone(mean, q, log_s, logit, sh, acc, cnt, True, 2e-4, 0.02, 1.6, np.zeros((2, 3)))
    assert counts == (3, 1, 1)  # Gaussian 0 (small) clones, 1 (big) splits


def test_grad_stats_and_prune(orc):
    acc = np.array([0.5, 0.0, 0.0], np.float32)
    g = np.array([[3, 4, 0], [0, 0, 1], [1, 1, 1]], np.float32)
    assert np.allclose(orc.grad_stats(acc, g, use_max=True), [5, 1, math.sqrt(3)])
    assert np.allclose(orc.grad_stats(acc, g, use_max=False), [5.5, 1, math.sqrt(3)])
    assert np.allclose(orc.grad_stats(acc, g[:, :2], index=[2, 0, 1]), [0.5, 1.4142135, 5])
    mean, q, log_s, logit, sh = _scene()
    out, kept = orc.prune_alpha(mean, q, log_s, logit, sh, 0.05)  # sigmoid(-4) = 0.018 removed
    assert kept == 3 and np.array_equal(out[3], logit[:3])
