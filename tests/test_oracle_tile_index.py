"""Pins oracle.tile_index: the reference golden, a brute-force loop over every (tile, Gaussian)
pair, and the exact-zero property the ellipse cull relies on."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import tile_index as TI
from tests.helpers import field_tensors, random_field, to_np


def test_reference_golden_tile_index():
    # test_render.py:167-172
    gids, offsets = TI.build_tile_index(np.array([[2.0, 2.0], [20.0, 20.0]]),
                                        np.ones((2, 3)), np.array([[1, 1], [1, 1]]), 8, 8, 4)
    assert offsets.tolist() == [0, 1, 1, 1, 1]
    assert gids.tolist() == [0]


def _case(seed, N, H, W, ts):
    f = random_field(seed, N, H, W, scale_lo=0.5, scale_hi=7.0, margin=8.0,
                     aniso=(1.0, 9.0, 0.3, 2.0))
    if seed % 2:
        f["means"][:3] = np.array([[np.nan, 1.0], [1e12, 5.0], [-np.inf, 3.0]], np.float32)
    means, conics, colors, radii, _ = field_tensors(f, "cpu")
    return to_np(means), to_np(conics), to_np(radii)


@pytest.mark.parametrize("seed,N,H,W,ts", [
    (1, 60, 21, 23, 4), (2, 80, 33, 29, 8), (3, 70, 40, 44, 16), (4, 50, 17, 70, 32),
    (5, 40, 9, 9, 3), (6, 30, 13, 5, 1),
])
@pytest.mark.parametrize("cull", [False, True])
def test_vectorized_matches_brute_force(seed, N, H, W, ts, cull):
    means, conics, radii = _case(seed, N, H, W, ts)
    g, o = TI.build_tile_index(means, conics, radii, H, W, ts, cull, 3.0)
    bg, bo = TI.brute_force_tile_index(means, conics, radii, H, W, ts, cull, 3.0)
    assert o.tolist() == bo.tolist()
    assert g.tolist() == bg.tolist()
    assert o[-1] == len(g)
    for t in range(len(o) - 1):
        seg = g[o[t]:o[t + 1]]
        assert (np.diff(seg) > 0).all()  # ascending Gaussian id inside every tile


def test_cull_drops_only_pairs_with_zero_faded_weight():
    """Every culled pair has q > cutoff^2 (weight exactly 0 under fade) on every pixel of the
    tile-clipped rectangle, evaluated in float64; and the cull really fires."""
    H, W, ts, cutoff = 40, 44, 8, 3.0
    f = random_field(22, 48, H, W, margin=3.0, aniso=(2.0, 8.0, 0.5, 1.2))
    means, conics, colors, radii, _ = field_tensors(f, "cpu")
    means, conics, radii = to_np(means), to_np(conics), to_np(radii)
    full_g, full_o = TI.build_tile_index(means, conics, radii, H, W, ts, False)
    cull_g, cull_o = TI.build_tile_index(means, conics, radii, H, W, ts, True, cutoff)
    x0, x1, y0, y1, _ = TI.support_rects(means, radii, H, W)
    tiles_x = (W + ts - 1) // ts
    dropped = 0
    for t in range(len(full_o) - 1):
        kept = set(cull_g[cull_o[t]:cull_o[t + 1]].tolist())
        tx, ty = t % tiles_x, t // tiles_x
        for i in full_g[full_o[t]:full_o[t + 1]].tolist():
            if i in kept:
                continue
            dropped += 1
            xs = np.arange(max(x0[i], tx * ts), min(x1[i], (tx + 1) * ts), dtype=np.float64)
            ys = np.arange(max(y0[i], ty * ts), min(y1[i], (ty + 1) * ts), dtype=np.float64)
            dx = xs[None, :] - np.float64(means[i, 0])
            dy = ys[:, None] - np.float64(means[i, 1])
            a, b, c = (np.float64(v) for v in conics[i])
            q = a * dx * dx + 2 * b * dx * dy + c * dy * dy
            assert q.min() > cutoff * cutoff, (t, i)
    assert dropped > 0


def test_tile_min_q_is_the_rectangle_minimum():
    """The float32 edge-minimiser formula equals a dense float64 minimum over the rectangle."""
    rng = np.random.default_rng(5)
    for _ in range(300):
        sx, sy = np.exp(rng.uniform(np.log(0.5), np.log(6.0), 2))
        th = rng.uniform(0, math.pi)
        cth, sth = math.cos(th), math.sin(th)
        a = cth * cth / sx ** 2 + sth * sth / sy ** 2
        b = cth * sth * (1 / sx ** 2 - 1 / sy ** 2)
        c = sth * sth / sx ** 2 + cth * cth / sy ** 2
        mx, my = rng.uniform(-10, 30, 2)
        xl, yl = rng.integers(0, 16, 2)
        xh, yh = xl + rng.integers(0, 12), yl + rng.integers(0, 12)
        got = float(TI.tile_min_q(mx, my, a, b, c, xl, xh, yl, yh))
        xs = np.linspace(xl, xh, 401)
        ys = np.linspace(yl, yh, 401)
        dx = xs[None, :] - np.float64(np.float32(mx))
        dy = ys[:, None] - np.float64(np.float32(my))
        a64, b64, c64 = (np.float64(np.float32(v)) for v in (a, b, c))
        dense = (a64 * dx * dx + 2 * b64 * dx * dy + c64 * dy * dy).min()
        assert got <= dense * (1 + 1e-5) + 1e-5
        assert got >= dense * (1 - 1e-3) - 1e-3  # the dense grid only approaches the minimum


def test_malformed_conic_is_never_culled():
    assert TI.tile_min_q(0.0, 0.0, -1.0, 0.0, 1.0, 5, 6, 5, 6) == -TI.FLT_MAX
    assert TI.tile_min_q(0.0, 0.0, 1.0, 2.0, 1.0, 5, 6, 5, 6) == -TI.FLT_MAX


def test_support_rects_follow_round_half_even_and_clip():
    means = np.array([[2.5, 3.5], [-0.5, 0.5], [9.6, 9.4]], np.float32)
    radii = np.array([[1, 1], [2, 0], [3, 3]])
    x0, x1, y0, y1, valid = TI.support_rects(means, radii, 10, 10)
    # round(2.5) = 2, round(3.5) = 4, round(-0.5) = -0 = 0, round(0.5) = 0
    assert (x0.tolist(), x1.tolist()) == ([1, 0, 7], [4, 3, 10])
    assert (y0.tolist(), y1.tolist()) == ([3, 0, 6], [6, 1, 10])
    assert valid.all()
