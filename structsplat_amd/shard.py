"""One GPU's share of a row-band split of a large image.

A large image shards into horizontal bands of whole tile rows, one band per GPU.  Each GPU
renders its band from the Gaussians whose support rectangle meets the band's rows (its own plus
a halo from the bands above and below), and its backward yields partial gradients for those
Gaussians; summing the partials over all bands (the all-reduce of a multi-GPU run) gives the
full image's gradients.

The band is rendered as an image of its own with the Gaussians' means shifted up by the band's
first row y0.  For a mean at or below row y0/2 (every mean of the band's own rows and most of
the halo) the float32 shift is exact (Sterbenz's lemma, or an integer subtracted from a larger
mean), every pixel offset ``p - mean`` is the same rounded number as in the full image, and the
band's pixels are bit-identical to the full render's rows.  A halo mean further above (or above
the image) lands in a larger binade and is rounded: its pixel offsets then differ from the full
image's by at most one float32 ulp of ~y0.  Its support rectangle is kept exact: where the
rounding would move the half-to-even centre ``round(my)`` by a row, the shifted mean is moved by
the fewest ulps that restore ``round(my - y0) == round(my) - y0``, so the tile lists and every
pixel's set of Gaussians are the full image's.

The support rectangle is the reference's (render.py:62-78, render_ext.cu:13-50): rows
``round(my) - ry .. round(my) + ry`` with round-half-to-even.  Membership only needs a superset
of the Gaussians that reach the band: a member whose rectangle misses the band renders nothing
there.
"""
from __future__ import annotations

import math

import torch

from . import render as HR


def band_bounds(height: int, bands: int, tile_size: int = HR.DEFAULT_TILE_SIZE):
    """Row ranges ``[(y0, y1), ...]`` of `bands` bands of whole tile rows covering `height`
    rows; the first bands take one tile row more when the rows do not divide evenly."""
    height, bands, ts = int(height), int(bands), int(tile_size)
    if height <= 0 or bands <= 0 or ts <= 0:
        raise ValueError("height, bands and tile_size must be positive")
    tile_rows = (height + ts - 1) // ts
    if bands > tile_rows:
        raise ValueError(f"{bands} bands need at least {bands} tile rows, have {tile_rows}")
    out, row = [], 0
    for b in range(bands):
        take = tile_rows // bands + (1 if b < tile_rows % bands else 0)
        y0, y1 = row * ts, min((row + take) * ts, height)
        out.append((y0, y1))
        row += take
    return out


def band_members(means: torch.Tensor, radii: torch.Tensor, y0: int, y1: int) -> torch.Tensor:
    """Indices (ascending, int64, on the means' device) of the Gaussians whose support rows
    ``round(my) +- ry`` meet rows ``[y0, y1)``."""
    iy = torch.round(means.detach()[:, 1])
    ry = radii.detach()[:, 1].to(iy.dtype)
    hit = (iy - ry <= float(y1 - 1)) & (iy + ry >= float(y0))
    return torch.nonzero(hit, as_tuple=False).reshape(-1)


def shift_rows(means: torch.Tensor, y0: int) -> torch.Tensor:
    """``means - (0, y0)`` in float32 with every support-rectangle centre kept exact:
    ``round(result_y) == round(my) - y0`` (round half to even, the reference's rounding,
    render.py:62-78) for every finite ``|my| < 2^22``.  A rounded shift that breaks it is moved
    by single ulps toward the exact difference until it holds (at most a few steps: the rounded
    difference is within half an ulp of the exact one).  Differentiable: the moves are
    constants."""
    shifted = means - means.new_tensor([0.0, float(y0)])
    with torch.no_grad():
        my = means.detach()[:, 1]
        target = torch.round(my) - float(y0)
        sy = shifted.detach()[:, 1].clone()
        live = torch.isfinite(my) & (my.abs() < 2.0 ** 22)
        for _ in range(8):
            bad = live & (torch.round(sy) != target)
            if not bool(bad.any()):
                break
            up = torch.round(sy) < target
            step = torch.nextafter(sy, torch.where(up, torch.full_like(sy, math.inf),
                                                   torch.full_like(sy, -math.inf)))
            sy = torch.where(bad, step, sy)
        moved = sy != shifted.detach()[:, 1]
    if not bool(moved.any()):
        return shifted
    y = shifted[:, 1]
    y = torch.where(moved, sy + (y - y.detach()), y)  # value sy, gradient of the shift
    return torch.stack([shifted[:, 0], y], dim=1)


def band_inputs(means, conics, colors, radii, opacities, y0: int, y1: int):
    """The band's field: ``(idx, shift_rows(means, y0), conics, colors, radii, opacities)``
    restricted to the band's members (differentiable through the gather and the shift)."""
    idx = band_members(means, radii, y0, y1)
    has_op = opacities is not None and opacities.numel() > 0
    return (idx, shift_rows(means.index_select(0, idx), y0), conics.index_select(0, idx),
            colors.index_select(0, idx), radii.index_select(0, idx),
            opacities.index_select(0, idx) if has_op else None)


def render_band(means, conics, colors, radii, H: int, W: int, y0: int, y1: int,
                opacities=None, **kwargs):
    """Rows ``[y0, y1)`` of ``render_hip_exact(means, ..., H, W, **kwargs)``, rendered from the
    band's members only; returns ``(image (y1 - y0, W, 3), member indices)``.  `y0` must be a
    multiple of the tile size and ``y1 <= H``."""
    ts = int(kwargs.get("tile_size", HR.DEFAULT_TILE_SIZE))
    if not (0 <= y0 < y1 <= H) or y0 % ts != 0:
        raise ValueError(f"band rows [{y0}, {y1}) must lie in [0, {H}) and start on a "
                         f"{ts}-pixel tile row")
    idx, m, q, c, r, o = band_inputs(means, conics, colors, radii, opacities, y0, y1)
    return HR.render_hip_exact(m, q, c, r, int(y1 - y0), int(W), opacities=o, **kwargs), idx
