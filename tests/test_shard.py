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
