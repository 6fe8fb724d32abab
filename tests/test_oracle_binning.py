"""Pins the oracle's tile bounds and binning (gs/culling.py:8-37,
aabb_culling.h:15-103) with hand-computed bounds and invariants checked by
brute force over tiles.  CPU only."""
import numpy as np
import pytest


def cam300(orc, W=256, H=256):
    return orc.Camera(300.0, 300.0, W / 2, H / 2, W, H, 0.01, 5.0)


def test_aabb_hand_computed(orc):
    """sigma 0.02 at the image centre: e = sqrt(6 * 4e-4) = 0.04899; x*300+128:
    113.30 -> 113 -> tile 7; 142.70 -> 142 -> tile 8."""
    cam = cam300(orc)
    P, tl, br = orc.aabb(np.zeros((1, 2), np.float32), np.array([[4e-4, 0, 0, 4e-4]], np.float32), cam, 6.0)
    assert tl.tolist() == [[7, 7]] and br.tolist() == [[8, 8]] and P == 4


def test_aabb_clamp_saturation_and_nan(orc):
    cam = cam300(orc)
    m = np.array([[1e9, -1e9], [0.0, 0.0], [-0.3, 0.2]], np.float32)
    c = np.array([[1.0, 0, 0, 1.0], [np.nan, 0, 0, 4e-4], [1e-6, 0, 0, 1e-6]], np.float32)
    P, tl, br = orc.aabb(m, c, cam, 6.0)
    # huge x saturates to INT_MAX -> clamp 255 -> tile 15; huge negative y -> 0
    assert tl[0].tolist() == [15, 0] and br[0].tolist() == [15, 0]
    # NaN extent: NaN -> 0 -> tile 0 on x; y normal
    assert tl[1, 0] == 0 and br[1, 0] == 0
    # small Gaussian: x = -0.3*300+128 = 38 -> tile 2, y = 0.2*300+128 = 188 -> tile 11
    assert tl[2].tolist() == [2, 11] and br[2].tolist() == [2, 11]
    assert P == sum((b - a + 1).prod() for a, b in zip(tl, br))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_binning_invariants_and_brute_force(orc, seed):
    rng = np.random.default_rng(seed)
    nth, ntw = 9, 13
    V = 400
    lo = np.stack([rng.integers(0, ntw, V), rng.integers(0, nth, V)], 1)
    hi = np.minimum(lo + rng.integers(0, 5, (V, 2)), [ntw - 1, nth - 1])
    depth = rng.choice(np.array([0.1, 0.5, 0.5, 2.0, 7.0, -1.0], np.float32), V) * \
        (1 + (rng.random(V) < 0.5)).astype(np.float32)
    b = orc.bin_gaussians(lo.astype(np.int32), hi.astype(np.int32), depth, nth, ntw)
    P = int(((hi - lo + 1).prod(1)).sum())
    sk, ids, st, en = b["sorted_keys"], b["ids"], b["start"], b["end"]
    assert sk.size == P and ids.size == P
    assert np.all(sk[1:] >= sk[:-1])  # sorted
    tiles = (sk >> np.uint64(32)).astype(np.int64)
    covered = 0
    for t in range(nth * ntw):
        ty, tx = divmod(t, ntw)
        members = [v for v in range(V) if lo[v, 0] <= tx <= hi[v, 0] and lo[v, 1] <= ty <= hi[v, 1]]
        if not members:
            assert st[t] == -1 and en[t] == -1
            continue
        seg = ids[st[t]:en[t]]
        covered += en[t] - st[t]
        assert np.all(tiles[st[t]:en[t]] == t)
        # order: (depth bits as unsigned, index)
        want = sorted(members, key=lambda v: (int(depth[v:v + 1].view(np.uint32)[0]), v))
        assert seg.tolist() == want
    assert covered == P  # the ranges partition [0, P)


def test_negative_depth_sorts_after_positive(orc):
    """Unsigned comparison of the low word: negative depths (sign bit set)
    come after every positive depth of the same tile (A.2)."""
    lo = np.zeros((3, 2), np.int32)
    b = orc.bin_gaussians(lo, lo, np.array([-0.5, 3.0, 0.25], np.float32), 1, 1)
    assert b["ids"].tolist() == [2, 1, 0]


@pytest.mark.parametrize("W,H", [(256, 256), (600, 344)])
def test_two_level_block_lists_filter_to_tile_lists(orc, W, H):
    """Two-level binning: bounds counted with 64-px tiles are the 16-px bounds
    shifted right by 2 (so the block pair count is the shifted bounds'
    area), and filtering each 4 x 4-tile block's depth-ordered list by the
    records' 16-px bounds gives exactly each tile's own list, ties in depth
    and ragged block edges (600 x 344: 38 x 22 tiles, 10 x 6 blocks)
    included."""
    rng = np.random.default_rng(W)
    V = 3000
    cam = orc.Camera(300.0, 300.0, W / 2, H / 2, W, H, 0.01, 5.0)
    m = np.stack([rng.uniform(-W / 560, W / 560, V), rng.uniform(-H / 560, H / 560, V)], 1).astype(np.float32)
    sx, sy = np.exp(rng.uniform(-6.5, -3.0, (2, V)))
    r = rng.uniform(-0.9, 0.9, V)
    cov = np.stack([sx * sx, r * sx * sy, r * sx * sy, sy * sy], 1).astype(np.float32)
    depth = rng.uniform(0.1, 5.0, V).astype(np.float32)
    depth[::7] = 1.0  # ties: order by index
    P16, tl, br = orc.aabb(m, cov, cam, 6.0, 16)
    P64, tl64, br64 = orc.aabb(m, cov, cam, 6.0, 64)
    assert np.array_equal(tl64, tl >> 2) and np.array_equal(br64, br >> 2)
    assert P64 == int(((br64 - tl64 + 1).prod(1)).sum()) and P64 < P16
    nth, ntw = -(-H // 16), -(-W // 16)
    nbh, nbw = -(-nth // 4), -(-ntw // 4)
    tiles = orc.bin_gaussians(tl, br, depth, nth, ntw)
    blocks = orc.bin_gaussians(tl >> 2, br >> 2, depth, nbh, nbw)
    for ty in range(nth):
        for tx in range(ntw):
            t, b = ty * ntw + tx, (ty >> 2) * nbw + (tx >> 2)
            s0, e0 = blocks["start"][b], blocks["end"][b]
            cand = blocks["ids"][s0:e0] if s0 >= 0 else np.zeros(0, np.int64)
            hit = (tl[cand, 0] <= tx) & (tx <= br[cand, 0]) & (tl[cand, 1] <= ty) & (ty <= br[cand, 1])
            s1, e1 = tiles["start"][t], tiles["end"][t]
            want = tiles["ids"][s1:e1] if s1 >= 0 else np.zeros(0, np.int64)
            assert np.array_equal(cand[hit], want), (tx, ty)
