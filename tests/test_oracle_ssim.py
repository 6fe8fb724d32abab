"""Pins of the SSIM-loss oracle (CPU).  kornia is absent here, so its
published algorithm is restated twice (torch float64 and scipy float64) and
both are checked against closed forms, each other, finite differences, and
the border adjoint that the HIP kernel uses."""
import numpy as np
import pytest
import torch

from oracle import oracle as O


def _imgs(H, W, C, seed):
    rng = np.random.default_rng(seed)
    a = rng.random((H, W, C))
    b = np.clip(a + rng.normal(0.0, 0.15, (H, W, C)), 0.0, 1.0)
    return a, b


def test_window_is_the_published_gaussian():
    w = O.ssim_window1d(11)
    assert w.dtype == np.float32 and abs(float(w.sum()) - 1.0) < 1e-6
    # sigma 1.5, 11 taps: the values every SSIM implementation quotes
    np.testing.assert_allclose(w[:6], [0.0010284, 0.0075988, 0.0360008, 0.1093607, 0.2130055, 0.2660117], atol=2e-7)
    assert np.array_equal(w, w[::-1])


@pytest.mark.parametrize("a,b", [(0.3, 0.3), (0.25, 0.75), (0.0, 1.0), (0.9, 0.1)])
def test_constant_images_closed_form(a, b):
    x = np.full((13, 17, 3), a)
    y = np.full((13, 17, 3), b)
    # the float32-normalised window sums to S = 1 +- 1e-7 (kornia's, too): mu = a S,
    # E[x^2] = a^2 S, so s11 = a^2 S (1 - S) is not exactly 0
    S = float(O.ssim_window1d(11).astype(np.float64).sum()) ** 2
    m1, m2 = a * S, b * S
    s11, s22, s12 = a * a * S - m1 * m1, b * b * S - m2 * m2, a * b * S - m1 * m2
    want = ((2 * m1 * m2 + O.SSIM_C1) * (2 * s12 + O.SSIM_C2)) / (
        (m1 * m1 + m2 * m2 + O.SSIM_C1) * (s11 + s22 + O.SSIM_C2) + 1e-12)
    np.testing.assert_allclose(O.ssim_map_scipy(x, y, 11), want, rtol=1e-9, atol=1e-12)
    t = O.ssim_map_t(torch.tensor(x), torch.tensor(y), 11).numpy()
    np.testing.assert_allclose(t, want, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("shape,win", [((23, 31, 3), 11), ((12, 15, 1), 11), ((40, 9, 2), 7), ((6, 6, 3), 11)])
def test_two_restatements_agree(shape, win):
    a, b = _imgs(*shape, seed=sum(shape) + win)
    s1 = O.ssim_map_scipy(a, b, win)
    s2 = O.ssim_map_t(torch.tensor(a), torch.tensor(b), win).numpy()
    assert np.abs(s1 - s2).max() < 1e-12
    # identity and symmetry
    assert np.abs(O.ssim_map_scipy(a, a, win) - 1.0).max() < 1e-8
    assert np.abs(O.ssim_map_scipy(b, a, win) - s1).max() < 1e-14


def test_gradient_matches_central_differences():
    a, b = _imgs(9, 10, 2, seed=3)
    for l2 in (False, True):
        _, g = O.photometric_ssim_loss(a, b, win=5, ssim_mult=0.3, l2=l2)
        rng = np.random.default_rng(8)
        for _ in range(12):
            idx = tuple(int(rng.integers(0, n)) for n in a.shape)
            h = 1e-6
            ap, am = a.copy(), a.copy()
            ap[idx] += h
            am[idx] -= h
            fd = (O.photometric_ssim_loss(ap, b, 5, 0.3, l2)[0] - O.photometric_ssim_loss(am, b, 5, 0.3, l2)[0]) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9


@pytest.mark.parametrize("n,win", [(16, 11), (6, 11), (7, 5), (40, 11), (11, 21)])
def test_border_adjoint_weights(n, win):
    """The adjoint weights the HIP kernel uses, W(q, p) = w(|q-p|) +
    [q >= 1, q+p <= R] w(q+p) + [q <= n-2, 2(n-1)-q-p <= R] w(2(n-1)-q-p),
    equal the transpose of the reflect-pad-and-filter operator."""
    R = win // 2
    if R >= n:
        pytest.skip("one reflection needs win//2 < n")
    w = O.ssim_window1d(win).astype(np.float64)[R:]  # w[|d|]
    M = O.reflect_filter_matrix(n, win)
    Wq = np.zeros((n, n))
    for q in range(n):
        for p in range(max(0, q - R), min(n, q + R + 1)):
            v = w[abs(q - p)]
            if q >= 1 and q + p <= R:
                v += w[q + p]
            r = 2 * (n - 1) - q - p
            if q <= n - 2 and r <= R:
                v += w[r]
            Wq[q, p] = v
    assert np.abs(Wq - M.T).max() < 1e-15
