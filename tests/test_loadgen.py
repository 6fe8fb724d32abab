"""Host logic of the serving measurement: the FIFO queue model of the plain
per-call renderer and the curve drawn from it."""
import numpy as np

from gaussian_splatting_3d_amd import loadgen


def test_fifo_latencies_closed_forms():
    # a burst queues behind itself; spaced arrivals wait for nothing
    assert np.allclose(loadgen.fifo_latencies(np.zeros(3), np.ones(3)), [1.0, 2.0, 3.0])
    assert np.allclose(loadgen.fifo_latencies(np.array([0.0, 5.0, 10.0]), np.array([1.0, 2.0, 3.0])), [1, 2, 3])
    # the second arrives while the first is in service
    assert np.allclose(loadgen.fifo_latencies(np.array([0.0, 0.5]), np.array([1.0, 1.0])), [1.0, 1.5])


def test_plain_curve_serves_at_most_its_capacity():
    """With a fixed 1 ms service time the plain renderer's capacity is 1000
    views/s: below it the served rate follows the offered one, above it the
    served rate is the capacity and the queue (the p99) grows."""
    pool = {(s, w): [1e-3] for s in range(2) for w, _ in loadgen.RESOLUTIONS}
    curve = loadgen.plain_curve(pool, 1000.0, (0.5, 2.0, 4.0), 2, n=2000)
    low, over, far = curve
    assert abs(low["views_s"] - 500.0) < 0.06 * 500.0
    assert low["p99_ms"] < 10.0
    for c in (over, far):
        assert abs(c["views_s"] - 1000.0) < 1.0, c  # the queue drains at exactly one request per ms
    assert far["p99_ms"] > over["p99_ms"] > 100.0
