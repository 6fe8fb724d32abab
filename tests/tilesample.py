"""Helpers for checking full-size views on a sample of tiles.

sample_tiles: a stratified sample of non-empty tiles plus the heaviest ones.
oracle_records: the oracle's preprocessing of one view (cull, projection,
  tile bounds) without binning.
tile_counts: pairs per tile from the tile rectangles.
tile_lists: the sorted list of each sampled tile by brute force: the
  Gaussians whose tile rectangle covers the tile, in (depth bits, index)
  order, i.e. the order of a stable sort of the reference's keys
  (tile << 32 | depth bits) emitted in index order (aabb_culling.h:15-41,
  235-241).
tile_pixels: flat pixel indices of tiles.
"""
import numpy as np


def sample_tiles(counts, n_done_tile, every=10, top=16):
    """Every `every`-th non-empty tile in order of list length, plus the
    `top` longest lists and the `top` largest per-tile n_done."""
    nonempty = np.nonzero(counts)[0]
    by_len = nonempty[np.argsort(counts[nonempty], kind="stable")]
    by_nd = nonempty[np.argsort(n_done_tile[nonempty], kind="stable")]
    return np.unique(np.concatenate([by_len[every // 2::every], by_len[-top:], by_nd[-top:]]))


def tile_max(per_pixel, H, W, tile=16):
    """Per-tile maximum of a per-pixel array [H*W] (tiles in row-major order)."""
    nth, ntw = -(-H // tile), -(-W // tile)
    a = np.zeros((nth * tile, ntw * tile), np.int64)
    a[:H, :W] = np.asarray(per_pixel).reshape(H, W)
    return a.reshape(nth, tile, ntw, tile).max(axis=(1, 3)).reshape(-1)


def oracle_records(O, mean, qvec, log_s, logit, c2w, cam, D=6.0, frustum_radius=1.0):
    oc = O.Camera(cam.fx, cam.fy, cam.cx, cam.cy, cam.w, cam.h, cam.near_plane, cam.far_plane)
    s, a = O.activate(log_s, logit)
    normals, pts = O.frustum(c2w, oc)
    mask = O.cull(mean, s, normals, pts, frustum_radius)
    idx = np.nonzero(mask)[0]
    mean_np = mean.numpy() if hasattr(mean, "numpy") else np.asarray(mean)
    q_np = qvec.numpy() if hasattr(qvec, "numpy") else np.asarray(qvec)
    m2d, cov, depth = O.project(mean_np[idx], q_np[idx], s[idx], c2w)
    P, tl, br = O.aabb(m2d, cov, oc, D)
    return dict(mask=mask, idx=idx, m2d=m2d, cov=cov, depth=depth, a=a, P=int(P), tl=tl, br=br,
                ic=O.image_constants(oc))


def tile_counts(tl, br, n_tiles_h, n_tiles_w):
    """Pairs per tile (row-major) from the tile rectangles [tl, br] (a 2-D
    difference array)."""
    cnt = np.zeros((n_tiles_h + 1, n_tiles_w + 1), np.int64)
    np.add.at(cnt, (tl[:, 1], tl[:, 0]), 1)
    np.add.at(cnt, (tl[:, 1], br[:, 0] + 1), -1)
    np.add.at(cnt, (br[:, 1] + 1, tl[:, 0]), -1)
    np.add.at(cnt, (br[:, 1] + 1, br[:, 0] + 1), 1)
    return cnt.cumsum(0).cumsum(1)[:n_tiles_h, :n_tiles_w].reshape(-1)


def tile_lists(tl, br, depth, tiles, n_tiles_w, n_tiles_total):
    """(start [n_tiles_total], end, ids) holding the sampled tiles' lists
    (other tiles -1), built by brute-force membership."""
    dbits = np.ascontiguousarray(np.asarray(depth, np.float32).reshape(-1)).view(np.uint32)
    start = -np.ones(n_tiles_total, np.int32)
    end = -np.ones(n_tiles_total, np.int32)
    parts = []
    pos = 0
    for t in np.sort(tiles):
        ty, tx = divmod(int(t), n_tiles_w)
        sel = np.nonzero((tl[:, 0] <= tx) & (br[:, 0] >= tx) & (tl[:, 1] <= ty) & (br[:, 1] >= ty))[0]
        sel = sel[np.argsort(dbits[sel], kind="stable")]
        start[t], end[t] = pos, pos + sel.size
        parts.append(sel.astype(np.int32))
        pos += sel.size
    ids = np.concatenate(parts) if parts else np.zeros(0, np.int32)
    return start, end, ids


def tile_pixels(tiles, ntw, H, W, tile=16):
    pix = []
    for t in tiles:
        ty, tx = divmod(int(t), ntw)
        ys = np.arange(ty * tile, min(ty * tile + tile, H))
        xs = np.arange(tx * tile, min(tx * tile + tile, W))
        pix.append((ys[:, None] * W + xs[None, :]).reshape(-1))
    return np.concatenate(pix) if pix else np.zeros(0, np.int64)
