"""Torch restatement of the reference renderer (render.py).  Test infrastructure only.

Same tensor operations as the reference, so it runs in float32 (the reference's own numerics
and autograd gradient) or float64 (an accuracy oracle):
clipped support rectangles, element-budget slices, ragged per-pixel expansion, ``index_add``
accumulation and the normalized / additive epilogue.

One deliberate addition: non-finite means are dropped and centres / radii are clamped
(+-1e12 / +-2^36) before the integer arithmetic, following render_ext.cu:25-41.  The reference's
``torch.round(nan).long()`` is undefined behaviour; for every finite mean and radius inside the
clamps the rectangles are exactly those of render.py:62-78.
"""
from __future__ import annotations

import math

import torch

EPS = 1e-8
ELEMENTS_PER_RENDER_CHUNK = 4096
CENTER_CLAMP = 1.0e12
RADIUS_CLAMP = 2**36


def validate_normalization_eps(value: float) -> float:
    """render.py:33-37."""
    eps = float(value)
    if not math.isfinite(eps) or eps <= 0.0:
        raise ValueError(f"normalization_eps must be finite and > 0, got {value}")
    return eps


def element_budget(chunk: int) -> int:
    """render.py:40-45: `chunk` in 4096-element units with a 64-unit floor."""
    return max(int(chunk), 64) * ELEMENTS_PER_RENDER_CHUNK


def flat_tile_slices(n: torch.Tensor, budget: int):
    """render.py:48-59: Gaussian index ranges whose summed rectangle area fits the budget."""
    csum = torch.cumsum(n, dim=0)
    N = n.shape[0]
    start = 0
    base = 0
    while start < N:
        end = int(torch.searchsorted(csum, base + budget, right=True))
        end = max(end, start + 1)
        yield start, min(end, N)
        start = min(end, N)
        base = int(csum[start - 1])


def tile_bounds(means: torch.Tensor, radii: torch.Tensor, H: int, W: int):
    """render.py:62-78 (+ the non-finite drop of render_ext.cu:28-34): (x0, y0, Tx, n)."""
    m = means.detach()
    finite = torch.isfinite(m[:, 0]) & torch.isfinite(m[:, 1])
    mx = torch.where(finite, m[:, 0], torch.zeros_like(m[:, 0]))
    my = torch.where(finite, m[:, 1], torch.zeros_like(m[:, 1]))
    ix = torch.round(mx).clamp(-CENTER_CLAMP, CENTER_CLAMP).long()
    iy = torch.round(my).clamp(-CENTER_CLAMP, CENTER_CLAMP).long()
    r = radii.long().clamp(-RADIUS_CLAMP, RADIUS_CLAMP)
    x0 = (ix - r[:, 0]).clamp(min=0)
    x1 = (ix + r[:, 0]).clamp(max=W - 1)
    y0 = (iy - r[:, 1]).clamp(min=0)
    y1 = (iy + r[:, 1]).clamp(max=H - 1)
    Tx = (x1 - x0 + 1).clamp(min=0)
    n = Tx * (y1 - y0 + 1).clamp(min=0)
    n = torch.where(finite, n, torch.zeros_like(n))
    return x0, y0, Tx, n


def tile_coords(x0, y0, Tx, n, s: int, e: int, dev):
    """render.py:81-91: flat per-pixel Gaussian ids and integer pixel coordinates for [s, e)."""
    ns = n[s:e]
    ends = torch.cumsum(ns, dim=0)
    total = int(ends[-1]) if ns.numel() else 0
    gid = torch.repeat_interleave(torch.arange(s, e, device=dev), ns)
    t = torch.arange(total, device=dev) - (ends - ns)[gid - s]
    Txg = Tx[gid]
    px = x0[gid] + t % Txg
    py = y0[gid] + t // Txg
    return gid, px, py


def resolve_support_fade_alpha(support_fade: bool, support_fade_alpha: float | None = None
                               ) -> float:
    """render.py:94-98."""
    if support_fade_alpha is None:
        return 1.0 if support_fade else 0.0
    return max(0.0, min(1.0, float(support_fade_alpha)))


def support_weight(q, sigma_cutoff: float, support_fade: bool,
                   support_fade_alpha: float | None = None):
    """render.py:101-107: exp(-q/2), minus alpha * exp(-cutoff^2/2) clamped at 0 under fade."""
    w = torch.exp(-0.5 * q)
    alpha = resolve_support_fade_alpha(support_fade, support_fade_alpha)
    if alpha > 0.0:
        w = torch.clamp(w - alpha * math.exp(-0.5 * float(sigma_cutoff) ** 2), min=0.0)
    return w


def pixel_colors(colors, color_grads, gid, dx, dy, scales, rotations):
    """render.py:110-123: the Gaussian's color, plus with affine colors
    G[0] * local_x + G[1] * local_y in its own frame (local = R(-theta) d / scale)."""
    base = colors[gid]
    if color_grads is None:
        return base
    if scales is None or rotations is None:
        raise ValueError("affine color rendering requires scales and rotations")
    theta = rotations[gid]
    c, s = torch.cos(theta), torch.sin(theta)
    sx = scales[gid, 0].clamp_min(1e-6)
    sy = scales[gid, 1].clamp_min(1e-6)
    local_x = (c * dx + s * dy) / sx
    local_y = (-s * dx + c * dy) / sy
    grads = color_grads[gid]
    return base + grads[:, 0, :] * local_x[:, None] + grads[:, 1, :] * local_y[:, None]


def accumulate(means, conics, colors, radii, H: int, W: int, chunk: int, opacities,
               normalize: bool, support_fade: bool = False, sigma_cutoff: float = 3.0,
               support_fade_alpha: float | None = None, normalization_eps: float = EPS,
               color_grads=None, scales=None, rotations=None):
    """render.py:180-235 without the optional per-slice checkpointing (a memory knob that
    leaves values and gradients unchanged, test_render.py:53-81).  With ``color_grads``
    (N, 2, 3), ``scales`` (N, 2) and ``rotations`` (N,) the colors are affine in each
    Gaussian's local frame (render.py:110-123)."""
    dev, dt = means.device, means.dtype
    num = torch.zeros(H * W, 3, device=dev, dtype=dt)
    den = torch.zeros(H * W, 1, device=dev, dtype=dt) if normalize else None
    x0, y0, Tx, n = tile_bounds(means, radii, H, W)
    for s, e in flat_tile_slices(n, element_budget(chunk)):
        gid, px, py = tile_coords(x0, y0, Tx, n, s, e, dev)
        dx = px.to(dt) - means[gid, 0]
        dy = py.to(dt) - means[gid, 1]
        a, b, c = conics[gid, 0], conics[gid, 1], conics[gid, 2]
        q = a * dx * dx + 2.0 * b * dx * dy + c * dy * dy
        w = support_weight(q, sigma_cutoff, support_fade, support_fade_alpha)
        if opacities is not None:
            w = w * opacities[gid]
        flat = py * W + px
        col = pixel_colors(colors, color_grads, gid, dx, dy, scales, rotations)
        num = num.index_add(0, flat, w[:, None] * col)
        if normalize:
            den = den.index_add(0, flat, w[:, None])
    if normalize:
        return (num / (den + normalization_eps)).view(H, W, 3)
    return num.view(H, W, 3)


def render(means, conics, colors, radii, H: int, W: int, chunk: int = 4096, opacities=None,
           support_fade: bool = False, sigma_cutoff: float = 3.0,
           support_fade_alpha: float | None = None, normalization_eps: float = EPS,
           color_grads=None, scales=None, rotations=None):
    """Normalized weighted sum (render.py:238-249)."""
    normalization_eps = validate_normalization_eps(normalization_eps)
    return accumulate(means, conics, colors, radii, H, W, chunk, opacities, True, support_fade,
                      sigma_cutoff, support_fade_alpha, normalization_eps, color_grads, scales,
                      rotations)


def render_additive(means, conics, colors, radii, H: int, W: int, chunk: int = 4096,
                    opacities=None, support_fade: bool = False, sigma_cutoff: float = 3.0,
                    support_fade_alpha: float | None = None, color_grads=None, scales=None,
                    rotations=None):
    """Additive accumulation (render.py:252-263)."""
    return accumulate(means, conics, colors, radii, H, W, chunk, opacities, False, support_fade,
                      sigma_cutoff, support_fade_alpha, EPS, color_grads, scales, rotations)


@torch.no_grad()
def gaussian_activity(means, conics, radii, H: int, W: int, chunk: int = 4096,
                      support_fade: bool = False, sigma_cutoff: float = 3.0,
                      support_fade_alpha: float | None = None):
    """Summed unnormalized, opacity-free weight per Gaussian (render.py:417-442)."""
    dev, dt = means.device, means.dtype
    activity = torch.zeros(means.shape[0], device=dev, dtype=dt)
    x0, y0, Tx, n = tile_bounds(means, radii, H, W)
    for s, e in flat_tile_slices(n, element_budget(chunk)):
        gid, px, py = tile_coords(x0, y0, Tx, n, s, e, dev)
        dx = px.to(dt) - means[gid, 0]
        dy = py.to(dt) - means[gid, 1]
        a, b, c = conics[gid, 0], conics[gid, 1], conics[gid, 2]
        q = a * dx * dx + 2.0 * b * dx * dy + c * dy * dy
        activity.index_add_(0, gid, support_weight(q, sigma_cutoff, support_fade,
                                                   support_fade_alpha))
    return activity


def conics_from_rs(log_scales, rotations, dilation: float = 0.0):
    """Inverse covariance entries (a, b, c) of an RS field (gaussians.py:275-300)."""
    s = torch.exp(log_scales)
    inv_sx2 = 1.0 / (s[:, 0] ** 2 + dilation)
    inv_sy2 = 1.0 / (s[:, 1] ** 2 + dilation)
    c = torch.cos(rotations)
    sn = torch.sin(rotations)
    return torch.stack([c * c * inv_sx2 + sn * sn * inv_sy2,
                        c * sn * (inv_sx2 - inv_sy2),
                        sn * sn * inv_sx2 + c * c * inv_sy2], dim=1)


@torch.no_grad()
def radii_from_rs(log_scales, rotations, sigma_cutoff: float, dilation: float = 0.0):
    """Integer support half-widths (gaussians.py:302-321)."""
    s2 = torch.exp(log_scales) ** 2 + dilation
    c = torch.cos(rotations)
    sn = torch.sin(rotations)
    var_x = c * c * s2[:, 0] + sn * sn * s2[:, 1]
    var_y = sn * sn * s2[:, 0] + c * c * s2[:, 1]
    r = sigma_cutoff * torch.sqrt(torch.stack([var_x, var_y], dim=1))
    return torch.clamp(r.ceil().long(), min=1)
