"""Test helpers: seeded small fields in the reference tests' style and error metrics."""
from __future__ import annotations

import math

import numpy as np
import torch

from oracle import exact64
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


def _host_np(x):
    return x.detach().cpu().numpy() if torch.is_tensor(x) else np.asarray(x)


def pixel_rel_err(got, ref, floor: float = 1e-6):
    """Per-pixel (elementwise) relative error max |got - ref| / max(|ref|, floor * max |ref|):
    the forward's 1e-5 relative bar taken pixel by pixel, with values below `floor` of the
    image's scale measured against that floor (exact zeros of both sides count as 0)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if ref.size == 0:
        return 0.0
    scale = float(np.max(np.abs(ref)))
    denom = np.maximum(np.abs(ref), max(floor * scale, 1e-300))
    return float(np.max(np.abs(got - ref) / denom))


def local_gradient(host, H, W, G, i, normalize):
    """Gaussian i's float64 gradient from the oracle on its own support rectangle (a pixel sums
    only the Gaussians whose rectangle holds it, so these pixels give its whole gradient), with
    the window problem (for finite differences).  `host` = (means, conics, colors, radii,
    opacities) as arrays or tensors."""
    m, q, c, r, o = (_host_np(t) for t in host)
    x0, x1, y0, y1, valid = support_rects(m, r, H, W)
    assert valid[i]
    win = window_field(m, q, c, r, o, H, W, x0[i], x1[i], y0[i], y1[i])
    ids, mw, qw, cw, rw, ow, rects = win
    k = int(np.searchsorted(ids, i))
    assert ids[k] == i
    h, w = int(y1[i] - y0[i]), int(x1[i] - x0[i])
    Gw = G[y0[i]:y1[i], x0[i]:x1[i]]
    g = exact64.backward(Gw, mw, qw, cw, rw, ow, h, w, normalize, 1e-8, False, 3.0, rects=rects,
                         float64_params=True)
    return [x[k] for x in g], (k, Gw, h, w, mw, qw, cw, rw, ow, rects)


def local_gradient_conditioning(host, H, W, G, i, normalize):
    """Gaussian i's float32 sensitivity (``exact64.backward_conditioning``) on its own support
    rectangle, shaped like ``local_gradient``'s gradients."""
    _, (k, Gw, h, w, mw, qw, cw, rw, ow, rects) = local_gradient(host, H, W, G, i, normalize)
    s = exact64.backward_conditioning(Gw, mw, qw, cw, rw, ow, h, w, normalize, 1e-8, False, 3.0,
                                      rects=rects, float64_params=True)
    return [x[k] for x in s]


def grad_bar_err(got, ref, sens, rel: float = 1e-4, units: float = 64.0):
    """The gradients' bar as a ratio (pass <= 1): every value must satisfy
    ``|got - ref| <= max(rel * max |ref|, units * u * sens)``, with ``sens`` from
    ``exact64.backward_conditioning``.  Where the float32 chain is well conditioned this is the
    ``rel`` bar relative to the array's scale; where its terms cancel (needle-like Gaussians, a
    Gaussian's color near the pixels' output) it is what any float32 evaluation can meet."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = float(np.max(np.abs(ref))) if ref.size else 0.0
    allow = np.maximum(rel * scale, units * F32_UNIT * np.asarray(sens, dtype=np.float64))
    diff = np.abs(got - ref)
    ratio = np.where(allow > 0, diff / np.maximum(allow, 1e-300), np.where(diff > 0, np.inf, 0))
    return float(np.max(ratio)) if ratio.size else 0.0


def sample_gaussians(host, H, W, count, seed):
    """`count` random visible Gaussians (ascending ids)."""
    m, r = _host_np(host[0]), _host_np(host[3])
    valid = support_rects(m, r, H, W)[4]
    ids = np.nonzero(valid)[0]
    return np.sort(np.random.default_rng(seed).choice(ids, size=count, replace=False))


F32_UNIT = 2.0 ** -24  # float32 unit roundoff


def pixel_bar_err(got, ref, sens, rel: float = 1e-5, units: float = 64.0):
    """The forward's per-pixel bar as a ratio (pass <= 1): each value must satisfy
    ``|got - ref| <= max(rel * |ref|, units * u * sens)``, with ``sens`` from
    ``exact64.forward_conditioning`` (the pixel's float32 sensitivity, u = 2^-24).  Where the
    reference's own float32 formula is well conditioned (sens ~ |ref|) this is the 1e-5 relative
    bar itself; where it is not (faded weights near the cutoff, needle-like Gaussians) the bar is
    what any float32 evaluation of that formula can meet."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    allow = np.maximum(rel * np.abs(ref), units * F32_UNIT * np.asarray(sens, dtype=np.float64))
    diff = np.abs(got - ref)
    ratio = np.where(allow > 0, diff / np.maximum(allow, 1e-300), np.where(diff > 0, np.inf, 0))
    return float(np.max(ratio)) if ratio.size else 0.0
