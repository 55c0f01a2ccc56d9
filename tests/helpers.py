"""Test helpers: seeded small fields in the reference tests' style and error metrics."""
from __future__ import annotations

import math

import numpy as np
import torch

from oracle import reference_render as R
from oracle.tile_index import support_rects


def random_field(seed: int, N: int, H: int, W: int, scale_lo=0.7, scale_hi=4.0, margin=2.0,
                 opacity=False, aniso=None):
    """Field arrays in the reference tests' recipe (test_render.py:178-184, :236-243)."""
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(-margin, W + margin, N), rng.uniform(-margin, H + margin, N)], 1)
    if aniso is None:
        scales = np.exp(rng.uniform(math.log(scale_lo), math.log(scale_hi), (N, 2)))
    else:
        sx = np.exp(rng.uniform(math.log(aniso[0]), math.log(aniso[1]), N))
        sy = np.exp(rng.uniform(math.log(aniso[2]), math.log(aniso[3]), N))
        scales = np.stack([sx, sy], 1)
    angles = rng.uniform(0, math.pi, N)
    colors = rng.random((N, 3))
    opac = rng.uniform(0.1, 1.0, N) if opacity else None
    return {
        "means": means.astype(np.float32),
        "scales": np.clip(scales, 1e-3, None).astype(np.float32),
        "angles": angles.astype(np.float32),
        "colors": colors.astype(np.float32),
        "opacity": None if opac is None else opac.astype(np.float32),
    }


def field_tensors(f, device, sigma_cutoff=3.0, dilation=0.0):
    """(means, conics, colors, radii, opacities) as float32/int64 tensors on `device`,
    produced by the reference formulas (gaussians.py:275-321) on the CPU in float32."""
    means = torch.tensor(f["means"], dtype=torch.float32)
    log_scales = torch.log(torch.tensor(f["scales"], dtype=torch.float32))
    rot = torch.tensor(f["angles"], dtype=torch.float32)
    conics = R.conics_from_rs(log_scales, rot, dilation)
    radii = R.radii_from_rs(log_scales, rot, sigma_cutoff, dilation)
    colors = torch.tensor(f["colors"], dtype=torch.float32)
    op = None if f.get("opacity") is None else torch.tensor(f["opacity"], dtype=torch.float32)
    out = [means, conics, colors, radii, op]
    return [None if t is None else t.to(device) for t in out]


def to_np(t):
    return None if t is None else t.detach().cpu().numpy()


def rel_err(got, ref, floor: float):
    """max |got - ref| / max(|ref|, floor), elementwise."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor))) if ref.size else 0.0


def scaled_err(got, ref):
    """max |got - ref| / max |ref| (error relative to the array's scale)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = float(np.max(np.abs(ref))) if ref.size else 0.0
    return float(np.max(np.abs(got - ref))) / max(scale, 1e-30) if ref.size else 0.0


def window_field(means, conics, colors, radii, opacities, H: int, W: int, x0: int, x1: int,
                 y0: int, y1: int):
    """The Gaussians that render into the pixel window [x0, x1) x [y0, y1) of an H x W image,
    moved into window coordinates (for the float64 oracle with ``float64_params=True``).

    A pixel sums exactly the Gaussians whose clipped support rectangle holds it, so rendering
    these Gaussians over the window, with their image rectangles clipped to the window and
    shifted (``rects``), gives the image's pixels there.  Means are shifted in float64, which
    holds every float32 mean minus an integer exactly.  Returns ``(ids, means, conics, colors,
    radii, opacities, rects)``; ``ids`` are the Gaussians' indices in the field, ascending.
    """
    m32 = np.asarray(means, dtype=np.float32).reshape(-1, 2)
    rx0, rx1, ry0, ry1, valid = support_rects(m32, radii, H, W)
    hit = valid & (rx0 < x1) & (rx1 > x0) & (ry0 < y1) & (ry1 > y0)
    ids = np.nonzero(hit)[0]
    rects = (np.clip(rx0[ids], x0, x1) - x0, np.clip(rx1[ids], x0, x1) - x0,
             np.clip(ry0[ids], y0, y1) - y0, np.clip(ry1[ids], y0, y1) - y0,
             np.ones(ids.size, dtype=bool))
    shift = np.array([x0, y0], dtype=np.float64)
    f64 = lambda a, k: np.asarray(a, dtype=np.float32).reshape(-1, k)[ids].astype(np.float64)
    op = None if opacities is None or np.size(opacities) == 0 else f64(opacities, 1)[:, 0]
    return (ids, m32[ids].astype(np.float64) - shift, f64(conics, 3), f64(colors, 3),
            np.asarray(radii).reshape(-1, 2)[ids], op, rects)
