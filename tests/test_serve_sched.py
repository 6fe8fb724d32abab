"""The render server's scheduling policy on the CPU, through host-only hooks
of the library (no device is touched): the arena allocator, the choice of a
batch from the queue and the admission of a counted batch."""
import ctypes

import numpy as np
import pytest

from gaussian_splatting_3d_amd import _lib

A = 256  # arena granularity


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def arena_run(size, ops):
    ops = np.ascontiguousarray(np.asarray(ops, np.int64).reshape(-1, 3))
    out = np.zeros(len(ops), np.int64)
    assert _lib.load().gsmi_sched_arena_run(size, len(ops), _ptr(ops), _ptr(out)) == 0
    return out


def up(n):
    return -(-max(n, 1) // A) * A


def test_arena_first_fit_and_coalescing():
    size = 16 * A
    # fill with three blocks, free the middle one, take two smaller pieces of it
    out = arena_run(size, [(0, 3 * A, 0), (0, 1, 0), (0, 5 * A, 0), (1, 3 * A, 1), (0, A, 0), (0, 2 * A, 0),
                           (2, 0, 0), (0, 10 * A, 0), (0, 7 * A, 0), (2, 0, 0)])
    # 0..3A, 3A..4A (1 byte rounds up), 4A..9A; free 3A..4A; A at 3A; 2A goes past 9A (first fit)
    assert list(out[:3]) == [0, 3 * A, 4 * A]
    assert out[4] == 3 * A and out[5] == 9 * A
    assert out[6] == 5 * A  # 11A..16A free
    assert out[7] == -1  # 10A does not fit anywhere
    assert out[8] == -1 and out[9] == 5 * A
    # freeing everything merges back to one range
    out = arena_run(size, [(0, 4 * A, 0), (0, 4 * A, 0), (0, 4 * A, 0), (1, 4 * A, 4 * A), (1, 0, 4 * A),
                           (1, 8 * A, 4 * A), (2, 0, 0), (0, 16 * A, 0)])
    assert out[6] == 16 * A and out[7] == 0


def test_arena_random_operations_keep_invariants():
    """Random allocate/free sequences: offsets are aligned, live ranges lie
    inside the arena and never overlap, a failed request is larger than the
    largest free range, and once all is freed one range holds the arena."""
    rng = np.random.default_rng(3)
    size = 1 << 22
    for trial in range(20):
        ops, live = [], []
        for _ in range(300):
            if live and rng.random() < 0.45:
                k = int(rng.integers(len(live)))
                ops.append((1, live[k][0], live[k][1]))
                live.pop(k)
                continue
            n = int(rng.integers(1, 1 << 18))
            ops.append((0, n, 0))
            out = arena_run(size, ops + [(2, 0, 0)])
            off, largest = int(out[-2]), int(out[-1])
            ops[-1] = (0, n, 0)
            if off < 0:
                ops.pop()
                before = arena_run(size, ops + [(2, 0, 0)])[-1]
                assert up(n) > before
                continue
            assert off % A == 0 and off + up(n) <= size
            for o, m in live:
                assert off + up(n) <= o or o + up(m) <= off
            live.append((off, n))
            assert largest <= size
        for o, m in live:
            ops.append((1, o, m))
        assert arena_run(size, ops + [(2, 0, 0)])[-1] == size


def select(scene, pixels, layout, max_views, max_pixels):
    scene = np.asarray(scene, np.int32)
    pixels = np.asarray(pixels, np.int64)
    layout = np.asarray(layout, np.int32)
    n = len(scene)
    batch, rest, nb = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(1, np.int32)
    assert _lib.load().gsmi_sched_select_batch(n, _ptr(scene), _ptr(pixels), len(layout), _ptr(layout), max_views,
                                               max_pixels, _ptr(batch), _ptr(nb), _ptr(rest)) == 0
    k = int(nb[0])
    return list(batch[:k]), list(rest[:n - k])


def test_select_batch_cases():
    # budgets: views, then pixels (an over-budget request is skipped, a later small one joins)
    b, r = select([0, 0, 0, 0], [10, 10, 10, 10], [0], 3, 100)
    assert (b, r) == ([0, 1, 2], [3])
    b, r = select([0, 0, 0, 0], [60, 50, 30, 10], [0], 8, 100)
    assert (b, r) == ([0, 2, 3], [1])
    # the oldest request always goes, even over the pixel budget
    b, r = select([1, 1], [500, 10], [0, 0], 8, 100)
    assert (b, r) == ([0], [1])
    # scenes with another SH layout wait; the batch is grouped by scene, stably:
    # lead scene 2 (layout 7) takes scene 0 (layout 7) along, not scene 1 (layout 9)
    b, r = select([2, 0, 1, 2, 0], [1] * 5, [7, 9, 7], 8, 100)
    assert r == [2]
    assert b == [1, 4, 0, 3]  # scene 0's views in arrival order, then scene 2's


def test_select_batch_random_properties():
    rng = np.random.default_rng(11)
    for _ in range(300):
        n = int(rng.integers(1, 40))
        n_sc = int(rng.integers(1, 5))
        layout = rng.integers(0, 2, n_sc)
        scene = rng.integers(0, n_sc, n)
        pixels = rng.integers(0, 3_000_000, n)
        max_views = int(rng.integers(1, 33))
        max_pixels = int(rng.integers(1, 8_000_000))
        b, r = select(scene, pixels, layout, max_views, max_pixels)
        assert 0 in b and 1 <= len(b) <= max_views
        assert sorted(b + r) == list(range(n)) and r == sorted(r)
        assert all(layout[scene[i]] == layout[scene[0]] for i in b)
        assert len(b) == 1 or pixels[b].sum() <= max_pixels
        # grouped by scene; arrival order within a scene
        assert [scene[i] for i in b] == sorted(scene[i] for i in b)
        for s in set(scene[i] for i in b):
            idx = [i for i in b if scene[i] == s]
            assert idx == sorted(idx)
        # a same-layout request left behind, examined before the view cap was
        # reached, did not fit the pixel budget next to those taken before it
        taken = sorted(b)
        for i in r:
            if layout[scene[i]] != layout[scene[0]]:
                continue
            before = [j for j in taken if j < i]
            if len(before) >= max_views:
                continue
            assert pixels[before].sum() + pixels[i] > max_pixels


def admit(vV, vP, nb, largest):
    vV, vP, nb = (np.ascontiguousarray(np.asarray(x, np.int64)) for x in (vV, vP, nb))
    B, fi, by = np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1, np.uint64)
    assert _lib.load().gsmi_sched_admit(len(vV), _ptr(vV), _ptr(vP), _ptr(nb), ctypes.c_uint64(largest), _ptr(B),
                                        _ptr(fi), _ptr(by)) == 0
    return int(B[0]), bool(fi[0]), int(by[0])


def test_admission_takes_the_longest_prefix_that_fits():
    vV = [300_000, 200_000, 900_000, 50_000]
    vP = [2_000_000, 1_000_000, 9_000_000, 100_000]
    nb = [600, 600, 2000, 300]
    B_all, fi, need_all = admit(vV, vP, nb, 1 << 50)
    assert (B_all, fi) == (4, True)
    sizes = [admit(vV[:k], vP[:k], nb[:k], 1 << 50)[2] for k in range(1, 5)]
    assert sizes == sorted(sizes) and sizes[-1] == need_all
    for k in range(1, 5):  # exactly the k-view prefix fits
        assert admit(vV, vP, nb, sizes[k - 1])[:2] == (k, True)
        if k > 1:
            assert admit(vV, vP, nb, sizes[k - 1] - 1)[0] == k - 1
    # one view is returned as it is even when it does not fit
    B, fi, need = admit(vV, vP, nb, 1)
    assert (B, fi, need) == (1, True, sizes[0])


def test_admission_keeps_pair_totals_in_int32():
    big = (1 << 31) - 10
    B, fi, _ = admit([10, 10, 10], [big - 100, 50, 100], [1, 1, 1], 1 << 60)
    assert (B, fi) == (2, True)
    B, fi, _ = admit([10, 10], [1 << 31, 1], [1, 1], 1 << 60)
    assert (B, fi) == (1, False)  # the caller fails such a request


def test_hooks_refuse_bad_arguments():
    L = _lib.load()
    ops = np.array([[5, 0, 0]], np.int64)
    out = np.zeros(1, np.int64)
    assert L.gsmi_sched_arena_run(1024, 1, _ptr(ops), _ptr(out)) == 1
    one = np.zeros(1, np.int32)
    assert L.gsmi_sched_select_batch(1, _ptr(np.array([3], np.int32)), _ptr(np.array([1], np.int64)), 2,
                                     _ptr(np.array([0, 0], np.int32)), 4, 10, _ptr(one), _ptr(one), _ptr(one)) == 1
    with pytest.raises(AssertionError):
        admit([], [], [], 1)
