"""bench.py's counting (CPU): the visit count E of the VALU roof."""
from __future__ import annotations

import numpy as np
import torch

import bench
from oracle.tile_index import support_rects


def test_rect_visits_equal_the_oracle_rectangles():
    """E sums the clipped support rectangles of render.py:62-78 (half-to-even centres, NaN means
    dropped, rectangles outside the image empty), as the oracle's support_rects builds them."""
    rng = np.random.default_rng(0)
    m = np.stack([rng.uniform(-40, 140, 800), rng.uniform(-40, 110, 800)], 1).astype(np.float32)
    m[3] = [np.nan, 1.0]
    m[4] = [5.5, 2.5]      # ties round to even
    m[5] = [6.5, -0.5]
    m[6] = [-60.0, 10.0]   # rectangle left of the image
    r = rng.integers(0, 30, (800, 2)).astype(np.int64)
    x0, x1, y0, y1, valid = support_rects(m, r, 70, 100)
    expect = int(((x1 - x0) * (y1 - y0))[valid].sum())
    assert bench.rect_visits(torch.from_numpy(m), torch.from_numpy(r), 70, 100) == expect

