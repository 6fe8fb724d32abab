"""Error metrics shared by the gradient tests.

rel_err: max-norm relative error of a whole tensor.
elementwise_excess: per element, |g - r| over rtol |r| + atol_abs * scale,
  where scale is the element's rounding scale from the oracle (the sum of its
  per-pixel terms' magnitudes, with the cancelling difference F - o of the
  reference's formula counted by its parts).  <= 1 means every element is in.
exact_excess: per element, |g - r| over rtol |r| + atol_max * max|r|, with r
  the float64 evaluation of the adjoint.  <= 1 means every element is in.
"""
import numpy as np


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _flat(x):
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.asarray(x, np.float64).reshape(-1)


def elementwise_excess(got, want, scale, rtol=1e-3, atol_abs=1e-5):
    g, r, a = _flat(got), _flat(want), _flat(scale)
    tiny = 1e-12 * max(np.abs(r).max(), 1e-30)
    return float((np.abs(g - r) / (rtol * np.abs(r) + atol_abs * a + tiny)).max())


def exact_excess(got, exact, rtol=1e-3, atol_max=1e-6):
    g, r = _flat(got), _flat(exact)
    floor = atol_max * max(np.abs(r).max(), 1e-30)
    return float((np.abs(g - r) / (rtol * np.abs(r) + floor)).max())
