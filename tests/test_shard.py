"""Row-band shards on the host: band bounds, and band membership against the oracle's support
rectangles (render.py:62-78): every Gaussian whose rectangle meets a band is a member of it."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle.tile_index import support_rects
from structsplat_amd import shard
from tests.helpers import field_tensors, random_field


def test_band_bounds_cover_the_rows_in_tile_rows():
    for H, bands, ts in ((384, 6, 16), (2048, 8, 16), (100, 3, 32), (17, 2, 16)):
        b = shard.band_bounds(H, bands, ts)
        assert len(b) == bands and b[0][0] == 0 and b[-1][1] == H
        assert all(y1 > y0 and y0 % ts == 0 for y0, y1 in b)
        assert all(b[k][1] == b[k + 1][0] for k in range(bands - 1))
    with pytest.raises(ValueError, match="tile rows"):
        shard.band_bounds(40, 4, 16)


@pytest.mark.parametrize("seed", [0, 1])
def test_band_members_hold_every_gaussian_reaching_the_band(seed):
    H, W = 200, 150
    f = random_field(seed, 700, H, W, scale_hi=9.0, margin=30.0)
    means, conics, colors, radii, op = field_tensors(f, "cpu")
    means[3] = float("nan")
    x0, x1, y0, y1, ok = support_rects(means.numpy(), radii.numpy(), H, W)
    owners = np.zeros(len(means), dtype=int)
    for a, b in shard.band_bounds(H, 5, 16):
        idx = shard.band_members(means, radii, a, b).numpy()
        assert (np.diff(idx) > 0).all()
        need = np.nonzero(ok & (y0 < b) & (y1 > a))[0]
        assert set(need) <= set(idx)
        assert 3 not in set(idx)
        owners[need] += 1
    assert (owners[ok] >= 1).all()  # every visible Gaussian reaches some band


@pytest.mark.parametrize("y0", [64, 128, 1024])
def test_shift_rows_keeps_support_rectangles_exact(y0):
    """Halo means far above the band (below y0/2, or above the image) lose bits in the float32
    shift; the rectangle centre round(my) - y0 must still be exact, so the band's rectangles
    are the full image's rows minus y0 (oracle support_rects)."""
    rng = np.random.default_rng(y0)
    my = np.concatenate([
        rng.uniform(-20.0, y0 / 2, 4000),                     # inexact shifts
        np.arange(-20.0, y0 / 2, 0.5) + 0.5,                  # exact half-integers (ties)
        np.float32(5.4999995) + np.zeros(1),                  # rounds across a half by one ulp
        np.nextafter(np.arange(-20.0, 32.0) + 0.5, -np.inf),  # just below ties
        np.nextafter(np.arange(-20.0, 32.0) + 0.5, np.inf),   # just above ties
        rng.uniform(y0 / 2, 3 * y0, 2000),                     # exact shifts
    ]).astype(np.float32)
    n = my.size
    means = torch.from_numpy(np.stack([rng.uniform(0, 64, n).astype(np.float32), my], 1))
    shifted = shard.shift_rows(means, y0)
    sy = shifted[:, 1].numpy()
    assert np.array_equal(np.round(sy.astype(np.float64)), np.round(my.astype(np.float64)) - y0)
    exact = my.astype(np.float64) - y0
    assert np.all(np.abs(sy - exact) <= 2 * np.spacing(np.abs(exact).astype(np.float32)))
    ok = my >= y0 / 2
    assert np.array_equal(sy[ok], (my[ok] - np.float32(y0)).astype(np.float32))
    assert torch.equal(shifted[:, 0], means[:, 0])
    # Band rectangles equal the full image's, clipped to the band, shifted by y0.
    H, W, h = y0 + 256, 64, 256
    radii = torch.from_numpy(rng.integers(1, y0 + 40, (n, 2)).astype(np.int64))
    fx0, fx1, fy0, fy1, fok = support_rects(means.numpy(), radii.numpy(), H, W)
    bx0, bx1, by0, by1, bok = support_rects(shifted.numpy(), radii.numpy(), h, W)
    band_y0 = np.maximum(fy0, y0) - y0
    band_y1 = np.minimum(fy1, y0 + h) - y0
    meets = fok & (band_y1 > band_y0)
    assert np.array_equal(bok, meets)
    assert np.array_equal(by0[meets], band_y0[meets])
    assert np.array_equal(by1[meets], band_y1[meets])
    assert np.array_equal(bx0[meets], fx0[meets]) and np.array_equal(bx1[meets], fx1[meets])


def test_shift_rows_is_differentiable():
    means = torch.tensor([[3.0, 5.4999995], [7.0, 200.25]], requires_grad=True)
    out = shard.shift_rows(means, 64)
    (out * torch.tensor([[1.0, 2.0], [3.0, 4.0]])).sum().backward()
    assert torch.equal(means.grad, torch.tensor([[1.0, 2.0], [3.0, 4.0]]))
