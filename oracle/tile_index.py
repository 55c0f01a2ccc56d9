"""NumPy integer restatement of the tile index.  Test infrastructure only.

Tile ``t`` (row-major over a ``ceil(W/ts) x ceil(H/ts)`` grid) owns the Gaussians whose clipped
support rectangle meets it, in ascending Gaussian id: the contract of the reference's torch
builder (cuda_render.py:177-226) with the within-tile order of its stable GPU radix sort
(render_ext.cu:135-196, :927-944).  With ``ellipse_cull`` a pair is dropped when the exact
minimum of the conic quadratic over the tile-clipped rectangle exceeds
``cutoff^2 + 1e-5 max(cutoff^2, 1)`` (render_ext.cu:97-133, :178-191), evaluated here in float32
with the same operation order so the decision is bit-identical to the GPU's.
"""
from __future__ import annotations

import numpy as np

CENTER_CLAMP = np.float32(1.0e12)
RADIUS_CLAMP = np.int64(2**36)
FLT_MAX = np.float32(np.finfo(np.float32).max)


def support_rects(means, radii, H: int, W: int):
    """Half-open clipped support rectangles (x0, x1, y0, y1) int64 and a non-empty mask.

    render.py:62-78 (round half to even, +-radius, clip to the image) plus the non-finite drop
    and overflow clamps of render_ext.cu:25-41.
    """
    means = np.asarray(means, dtype=np.float32).reshape(-1, 2)
    radii = np.asarray(radii, dtype=np.int64).reshape(-1, 2)
    finite = np.isfinite(means).all(axis=1)
    safe = np.where(finite[:, None], means, np.float32(0))
    c = np.clip(np.rint(safe), -CENTER_CLAMP, CENTER_CLAMP).astype(np.int64)
    r = np.clip(radii, -RADIUS_CLAMP, RADIUS_CLAMP)
    x0 = np.maximum(c[:, 0] - r[:, 0], 0)
    x1 = np.minimum(c[:, 0] + r[:, 0], W - 1)
    y0 = np.maximum(c[:, 1] - r[:, 1], 0)
    y1 = np.minimum(c[:, 1] + r[:, 1], H - 1)
    valid = finite & (x1 >= x0) & (y1 >= y0)
    x0 = np.where(valid, x0, 0)
    y0 = np.where(valid, y0, 0)
    x1 = np.where(valid, x1 + 1, 0)
    y1 = np.where(valid, y1 + 1, 0)
    return x0, x1, y0, y1, valid


def cull_threshold(sigma_cutoff: float) -> np.float32:
    """float(cutoff^2) + 1e-5 * max(float(cutoff^2), 1) in float32 (render_ext.cu:187-188, :920)."""
    cutoff2 = np.float32(float(sigma_cutoff) * float(sigma_cutoff))
    guard = np.float32(np.float32(1e-5) * max(cutoff2, np.float32(1.0)))
    return np.float32(cutoff2 + guard)


def _conic_q(dx, dy, a, b, c):
    """a*dx*dx + 2*b*dx*dy + c*dy*dy, float32, left-to-right, no fused multiply-add."""
    t1 = (a * dx) * dx
    t2 = ((np.float32(2.0) * b) * dx) * dy
    t3 = (c * dy) * dy
    return (t1 + t2) + t3


def tile_min_q(mx, my, a, b, c, x_lo, x_hi, y_lo, y_hi):
    """Vectorized float32 restatement of render_ext.cu:101-133 (exact minimum of the conic
    quadratic over [x_lo, x_hi] x [y_lo, y_hi])."""
    f = np.float32
    mx, my, a, b, c = (np.asarray(v, dtype=f) for v in (mx, my, a, b, c))
    x_lo, x_hi, y_lo, y_hi = (np.asarray(v, dtype=f) for v in (x_lo, x_hi, y_lo, y_hi))
    with np.errstate(all="ignore"):
        det = a * c - b * b
        malformed = ~((a > 0) & (c > 0) & (det > 0))
        inside = (mx >= x_lo) & (mx <= x_hi) & (my >= y_lo) & (my <= y_hi)
        slope_y = -(b / c)
        slope_x = -(b / a)
        best = np.full(np.broadcast(mx, x_lo).shape, FLT_MAX, dtype=f)
        dx = x_lo - mx
        dy = np.fmin(np.fmax(slope_y * dx, y_lo - my), y_hi - my)
        best = np.fmin(best, _conic_q(dx, dy, a, b, c))
        dx = x_hi - mx
        dy = np.fmin(np.fmax(slope_y * dx, y_lo - my), y_hi - my)
        best = np.fmin(best, _conic_q(dx, dy, a, b, c))
        dy = y_lo - my
        dx = np.fmin(np.fmax(slope_x * dy, x_lo - mx), x_hi - mx)
        best = np.fmin(best, _conic_q(dx, dy, a, b, c))
        dy = y_hi - my
        dx = np.fmin(np.fmax(slope_x * dy, x_lo - mx), x_hi - mx)
        best = np.fmin(best, _conic_q(dx, dy, a, b, c))
    best = np.where(inside, f(0.0), best)
    return np.where(malformed, -FLT_MAX, best).astype(f)


def tile_grid(H: int, W: int, tile_size: int):
    tiles_x = (W + tile_size - 1) // tile_size
    tiles_y = (H + tile_size - 1) // tile_size
    return tiles_x, tiles_y, tiles_x * tiles_y


def build_tile_index(means, conics, radii, H: int, W: int, tile_size: int,
                     ellipse_cull: bool = False, sigma_cutoff: float = 3.0):
    """(gids int32 (P,), offsets int64 (T+1,)); ascending Gaussian id within each tile."""
    if tile_size <= 0 or tile_size * tile_size > 1024:
        raise ValueError("tile_size must be positive with tile_size^2 <= 1024")
    means = np.asarray(means, dtype=np.float32).reshape(-1, 2)
    conics = np.asarray(conics, dtype=np.float32).reshape(-1, 3)
    tiles_x, _, T = tile_grid(H, W, tile_size)
    x0, x1, y0, y1, valid = support_rects(means, radii, H, W)
    ids = np.nonzero(valid)[0]
    if ids.size == 0:
        return np.zeros(0, np.int32), np.zeros(T + 1, np.int64)
    tx0 = x0[ids] // tile_size
    ty0 = y0[ids] // tile_size
    tw = (x1[ids] - 1) // tile_size - tx0 + 1
    th = (y1[ids] - 1) // tile_size - ty0 + 1
    counts = tw * th
    gid = np.repeat(ids, counts)
    start = np.cumsum(counts) - counts
    local = np.arange(int(counts.sum()), dtype=np.int64) - np.repeat(start, counts)
    twg = np.repeat(tw, counts)
    tile_x = np.repeat(tx0, counts) + local % twg
    tile_y = np.repeat(ty0, counts) + local // twg
    if ellipse_cull:
        ts = tile_size
        rx0 = np.maximum(np.repeat(x0[ids], counts), tile_x * ts)
        rx1 = np.minimum(np.repeat(x1[ids], counts) - 1, (tile_x + 1) * ts - 1)
        ry0 = np.maximum(np.repeat(y0[ids], counts), tile_y * ts)
        ry1 = np.minimum(np.repeat(y1[ids], counts) - 1, (tile_y + 1) * ts - 1)
        mq = tile_min_q(means[gid, 0], means[gid, 1], conics[gid, 0], conics[gid, 1],
                        conics[gid, 2], rx0.astype(np.float32), rx1.astype(np.float32),
                        ry0.astype(np.float32), ry1.astype(np.float32))
        keep = ~(mq > cull_threshold(sigma_cutoff))
        gid, tile_x, tile_y = gid[keep], tile_x[keep], tile_y[keep]
    key = tile_y * tiles_x + tile_x
    order = np.argsort(key, kind="stable")
    offsets = np.zeros(T + 1, np.int64)
    offsets[1:] = np.cumsum(np.bincount(key, minlength=T))
    return gid[order].astype(np.int32), offsets


def brute_force_tile_index(means, conics, radii, H: int, W: int, tile_size: int,
                           ellipse_cull: bool = False, sigma_cutoff: float = 3.0):
    """Loop over every (tile, Gaussian) pair; for tiny inputs only."""
    means = np.asarray(means, dtype=np.float32).reshape(-1, 2)
    conics = np.asarray(conics, dtype=np.float32).reshape(-1, 3)
    tiles_x, tiles_y, T = tile_grid(H, W, tile_size)
    x0, x1, y0, y1, valid = support_rects(means, radii, H, W)
    thr = cull_threshold(sigma_cutoff)
    lists = []
    for ty in range(tiles_y):
        for tx in range(tiles_x):
            lo_x, hi_x = tx * tile_size, (tx + 1) * tile_size - 1
            lo_y, hi_y = ty * tile_size, (ty + 1) * tile_size - 1
            members = []
            for i in range(means.shape[0]):
                if not valid[i]:
                    continue
                if x0[i] > hi_x or x1[i] - 1 < lo_x or y0[i] > hi_y or y1[i] - 1 < lo_y:
                    continue
                if ellipse_cull:
                    mq = tile_min_q(means[i, 0], means[i, 1], conics[i, 0], conics[i, 1],
                                    conics[i, 2], max(x0[i], lo_x), min(x1[i] - 1, hi_x),
                                    max(y0[i], lo_y), min(y1[i] - 1, hi_y))
                    if mq > thr:
                        continue
                members.append(i)
            lists.append(members)
    offsets = np.zeros(T + 1, np.int64)
    offsets[1:] = np.cumsum([len(m) for m in lists])
    gids = np.array([g for m in lists for g in m], dtype=np.int32)
    return gids, offsets
