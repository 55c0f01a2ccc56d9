"""Float64 forward and analytic backward of the render path.  Test infrastructure only.

Every Gaussian is evaluated on every pixel of its clipped support rectangle
(``tile_index.support_rects``), in float64 from the float32 inputs:

    q = a dx^2 + 2 b dx dy + c dy^2,  raw = exp(-q/2),  tail = alpha exp(-cutoff^2/2)
    base = max(raw - tail, 0) under support fade, else raw;   w = base * o
(alpha = 1 for support_fade=True, a number in [0, 1] for a fractional fade, render.py:94-107)
    N_k = sum w c_k,  D = sum w;   out = N/(D+eps) (normalized) or N (additive)

Backward (the per-pixel chain of render_ext.cu:485-530 and :784-818), with G = dL/dout:
    normalized: dN_k = G_k/(D+eps), dD = -sum_k G_k out_k/(D+eps);  additive: dN_k = G_k, dD = 0
    dc_ik += dN_k w_i;  dw_i = sum_k dN_k c_ik + dD;  do_i += dw_i base_i
    dq = dw_i o_i * (0 if fade and raw <= tail else -raw/2)
    dmx += dq (-2a dx - 2b dy);  dmy += dq (-2b dx - 2c dy)
    da += dq dx^2;  db += dq 2 dx dy;  dc += dq dy^2
The support rectangles (rounded means, integer radii) are held fixed, as in the reference
(render.py:70-71 detaches the means; gaussians.py:306 detaches the radii).
"""
from __future__ import annotations

import math

import numpy as np

from .tile_index import support_rects

_CHUNK_VISITS = 1 << 22


def _visit_chunks(x0, x1, y0, y1, valid):
    """Yield (gid, px, py) int64 arrays covering every (Gaussian, pixel) visit, in chunks."""
    ids = np.nonzero(valid)[0]
    if ids.size == 0:
        return
    n = (x1[ids] - x0[ids]) * (y1[ids] - y0[ids])
    csum = np.cumsum(n)
    start = 0
    while start < ids.size:
        base = csum[start - 1] if start > 0 else 0
        end = int(np.searchsorted(csum, base + _CHUNK_VISITS, side="right"))
        end = max(end, start + 1)
        sel = ids[start:end]
        ns = n[start:end]
        gid = np.repeat(sel, ns)
        first = np.cumsum(ns) - ns
        t = np.arange(int(ns.sum()), dtype=np.int64) - np.repeat(first, ns)
        tx = np.repeat(x1[sel] - x0[sel], ns)
        yield gid, np.repeat(x0[sel], ns) + t % tx, np.repeat(y0[sel], ns) + t // tx
        start = end


def _prepare(means, conics, colors, radii, opacities, H, W, rects, float64_params=False):
    """Parameters as float64 arrays: rounded to float32 first (the renderer's inputs), or taken
    as given with ``float64_params`` (finite differences with steps below float32's spacing;
    the support rectangles must then be passed in)."""
    if float64_params and rects is None:
        raise ValueError("float64_params needs explicit support rectangles")
    inp = np.float64 if float64_params else np.float32
    means = np.asarray(means, dtype=inp).reshape(-1, 2)
    rect = support_rects(means, radii, H, W) if rects is None else rects
    m = means.astype(np.float64)
    q = np.asarray(conics, dtype=inp).reshape(-1, 3).astype(np.float64)
    col = np.asarray(colors, dtype=inp).reshape(-1, 3).astype(np.float64)
    if opacities is None or np.size(opacities) == 0:
        op = np.ones(m.shape[0])
    else:
        op = np.asarray(opacities, dtype=inp).reshape(-1).astype(np.float64)
    return m, q, col, op, rect


def _tail(support_fade, sigma_cutoff) -> tuple[bool, float]:
    if isinstance(support_fade, bool):
        alpha = 1.0 if support_fade else 0.0
    else:
        alpha = max(0.0, min(1.0, float(support_fade)))
    return alpha > 0.0, alpha * math.exp(-0.5 * float(sigma_cutoff) ** 2)


def _weights(m, q, op, gid, px, py, fade, tail):
    dx = px.astype(np.float64) - m[gid, 0]
    dy = py.astype(np.float64) - m[gid, 1]
    a, b, c = q[gid, 0], q[gid, 1], q[gid, 2]
    quad = a * dx * dx + 2.0 * b * dx * dy + c * dy * dy
    raw = np.exp(-0.5 * quad)
    base = np.maximum(raw - tail, 0.0) if fade else raw
    return dx, dy, a, b, c, raw, base, base * op[gid]


def accumulate(means, conics, colors, radii, opacities, H: int, W: int, support_fade: bool,
               sigma_cutoff: float, rects=None, float64_params: bool = False):
    """Float64 numerator (H*W, 3) and denominator (H*W,) sums."""
    m, q, col, op, rect = _prepare(means, conics, colors, radii, opacities, H, W, rects,
                                   float64_params)
    fade, tail = _tail(support_fade, sigma_cutoff)
    num = np.zeros((H * W, 3))
    den = np.zeros(H * W)
    for gid, px, py in _visit_chunks(*rect):
        *_, w = _weights(m, q, op, gid, px, py, fade, tail)
        flat = py * W + px
        for k in range(3):
            num[:, k] += np.bincount(flat, weights=w * col[gid, k], minlength=H * W)
        den += np.bincount(flat, weights=w, minlength=H * W)
    return num, den


def forward(means, conics, colors, radii, opacities, H: int, W: int, normalize: bool = True,
            eps: float = 1e-8, support_fade: bool = False, sigma_cutoff: float = 3.0,
            rects=None, float64_params: bool = False):
    """Float64 image (H, W, 3) and raw coverage (H, W)."""
    num, den = accumulate(means, conics, colors, radii, opacities, H, W, support_fade,
                          sigma_cutoff, rects, float64_params)
    out = num / (den[:, None] + eps) if normalize else num
    return out.reshape(H, W, 3), den.reshape(H, W)


def forward_conditioning(means, conics, colors, radii, opacities, H: int, W: int,
                         normalize: bool = True, eps: float = 1e-8, support_fade: bool = False,
                         sigma_cutoff: float = 3.0, rects=None, float64_params: bool = False):
    """Per-pixel float32 sensitivity (H, W, 3) of the forward: a float32 evaluation of the
    reference's formula (render.py:101-107, :180-235) errs at a pixel by a few units of float32
    rounding times this value.  Per visit, the rounding of q = a dx^2 + 2b dx dy + c dy^2 is
    relative to qabs = |a| dx^2 + 2|b dx dy| + |c| dy^2 (large for needle-like Gaussians, whose
    terms cancel), so raw = exp(-q/2) carries raw (1 + qabs/2) of it, and a faded weight
    raw - tail carries that plus the tail's own rounding (large relative to the weight near the
    cutoff, where the difference cancels).  With s = o (raw (1 + qabs/2) + tail):
        additive:   S_k = sum s c_k
        normalized: (S_k + |out_k| sum s) / (D + eps)
    """
    m, q, col, op, rect = _prepare(means, conics, colors, radii, opacities, H, W, rects,
                                   float64_params)
    fade, tail = _tail(support_fade, sigma_cutoff)
    num = np.zeros((H * W, 3))
    den = np.zeros(H * W)
    sens = np.zeros((H * W, 3))
    sens_d = np.zeros(H * W)
    for gid, px, py in _visit_chunks(*rect):
        dx, dy, a, b, c, raw, base, w = _weights(m, q, op, gid, px, py, fade, tail)
        qabs = np.abs(a) * dx * dx + 2.0 * np.abs(b * dx * dy) + np.abs(c) * dy * dy
        sv = op[gid] * (raw * (1.0 + 0.5 * qabs) + (tail if fade else 0.0))
        flat = py * W + px
        for k in range(3):
            num[:, k] += np.bincount(flat, weights=w * col[gid, k], minlength=H * W)
            sens[:, k] += np.bincount(flat, weights=sv * np.abs(col[gid, k]), minlength=H * W)
        den += np.bincount(flat, weights=w, minlength=H * W)
        sens_d += np.bincount(flat, weights=sv, minlength=H * W)
    if normalize:
        out = num / (den[:, None] + eps)
        sens = (sens + np.abs(out) * sens_d[:, None]) / (den[:, None] + eps)
    return sens.reshape(H, W, 3)


def backward(grad_out, means, conics, colors, radii, opacities, H: int, W: int,
             normalize: bool = True, eps: float = 1e-8, support_fade: bool = False,
             sigma_cutoff: float = 3.0, rects=None, float64_params: bool = False):
    """Analytic float64 gradients (g_means (N,2), g_conics (N,3), g_colors (N,3), g_op (N,))."""
    m, q, col, op, rect = _prepare(means, conics, colors, radii, opacities, H, W, rects,
                                   float64_params)
    fade, tail = _tail(support_fade, sigma_cutoff)
    G = np.asarray(grad_out, dtype=np.float64).reshape(H * W, 3)
    if normalize:
        num, den = accumulate(means, conics, colors, radii, opacities, H, W, support_fade,
                              sigma_cutoff, rect, float64_params)
        denom = den + eps
        out = num / denom[:, None]
        dN = G / denom[:, None]
        dD = -(G * out).sum(axis=1) / denom
    else:
        dN = G
        dD = np.zeros(H * W)
    n = m.shape[0]
    g_m = np.zeros((n, 2))
    g_q = np.zeros((n, 3))
    g_c = np.zeros((n, 3))
    g_o = np.zeros(n)
    for gid, px, py in _visit_chunks(*rect):
        dx, dy, a, b, c, raw, base, w = _weights(m, q, op, gid, px, py, fade, tail)
        flat = py * W + px
        dw = (dN[flat] * col[gid]).sum(axis=1) + dD[flat]
        for k in range(3):
            g_c[:, k] += np.bincount(gid, weights=dN[flat, k] * w, minlength=n)
        g_o += np.bincount(gid, weights=dw * base, minlength=n)
        dvis = -0.5 * raw
        if fade:
            dvis = np.where(raw <= tail, 0.0, dvis)
        dq = dw * op[gid] * dvis
        g_m[:, 0] += np.bincount(gid, weights=dq * (-2.0 * a * dx - 2.0 * b * dy), minlength=n)
        g_m[:, 1] += np.bincount(gid, weights=dq * (-2.0 * b * dx - 2.0 * c * dy), minlength=n)
        g_q[:, 0] += np.bincount(gid, weights=dq * dx * dx, minlength=n)
        g_q[:, 1] += np.bincount(gid, weights=dq * 2.0 * dx * dy, minlength=n)
        g_q[:, 2] += np.bincount(gid, weights=dq * dy * dy, minlength=n)
    return g_m, g_q, g_c, g_o

def backward_conditioning(grad_out, means, conics, colors, radii, opacities, H: int, W: int,
                          normalize: bool = True, eps: float = 1e-8, support_fade: bool = False,
                          sigma_cutoff: float = 3.0, rects=None, float64_params: bool = False):
    """Per-Gaussian float32 sensitivity of the backward, shaped like ``backward``'s gradients: a
    float32 evaluation of the per-pixel chain above errs in a gradient by a few units of float32
    rounding times this value.  Per visit, with dw = sum_k dN_k c_k + dD and its rounding
    relative to dw_abs = sum_k |dN_k c_k| + |dD| (large where the Gaussian's color is near the
    pixel's output, whose terms cancel), and raw's rounding relative to rs = 1 + qabs/2 (as in
    ``forward_conditioning``; plus tail/raw under support fade):
        dq carries sq = o raw/2 (dw_abs + |dw| rs)
        means:  sq |2a dx + 2b dy| + |dq| (|2a dx| + |2b dy|)   (and the y analogue)
        conics: sq dx^2, sq 2|dx dy|, sq dy^2
        colors: |dN_k| w rs;   opacity: base (dw_abs + |dw| rs)
    summed in absolute value over the Gaussian's pixels.  The pixel offsets dx = px - mx are
    exact in float32 for the integer pixels and float32 means used here.
    """
    m, q, col, op, rect = _prepare(means, conics, colors, radii, opacities, H, W, rects,
                                   float64_params)
    fade, tail = _tail(support_fade, sigma_cutoff)
    G = np.asarray(grad_out, dtype=np.float64).reshape(H * W, 3)
    if normalize:
        num, den = accumulate(means, conics, colors, radii, opacities, H, W, support_fade,
                              sigma_cutoff, rect, float64_params)
        denom = den + eps
        out = num / denom[:, None]
        dN = G / denom[:, None]
        dD = -(G * out).sum(axis=1) / denom
        dD_abs = np.abs(G * out).sum(axis=1) / denom
    else:
        dN = G
        dD = np.zeros(H * W)
        dD_abs = np.zeros(H * W)
    n = m.shape[0]
    s_m = np.zeros((n, 2))
    s_q = np.zeros((n, 3))
    s_c = np.zeros((n, 3))
    s_o = np.zeros(n)
    for gid, px, py in _visit_chunks(*rect):
        dx, dy, a, b, c, raw, base, w = _weights(m, q, op, gid, px, py, fade, tail)
        flat = py * W + px
        qabs = np.abs(a) * dx * dx + 2.0 * np.abs(b * dx * dy) + np.abs(c) * dy * dy
        rs = 1.0 + 0.5 * qabs + (tail / np.maximum(raw, 1e-300) if fade else 0.0)
        dw = (dN[flat] * col[gid]).sum(axis=1) + dD[flat]
        dw_abs = np.abs(dN[flat] * col[gid]).sum(axis=1) + dD_abs[flat]
        live = ~(fade & (raw <= tail)) if fade else np.ones_like(raw, dtype=bool)
        o = np.abs(op[gid])
        sq = np.where(live, o * 0.5 * raw * (dw_abs + np.abs(dw) * rs), 0.0)
        dq = np.where(live, np.abs(dw) * o * 0.5 * raw, 0.0)
        for k in range(3):
            s_c[:, k] += np.bincount(gid, weights=np.abs(dN[flat, k]) * np.abs(w) * rs,
                                     minlength=n)
        s_o += np.bincount(gid, weights=np.abs(base) * (dw_abs + np.abs(dw) * rs), minlength=n)
        s_m[:, 0] += np.bincount(gid, weights=sq * np.abs(2.0 * a * dx + 2.0 * b * dy)
                                 + dq * (np.abs(2.0 * a * dx) + np.abs(2.0 * b * dy)), minlength=n)
        s_m[:, 1] += np.bincount(gid, weights=sq * np.abs(2.0 * b * dx + 2.0 * c * dy)
                                 + dq * (np.abs(2.0 * b * dx) + np.abs(2.0 * c * dy)), minlength=n)
        s_q[:, 0] += np.bincount(gid, weights=sq * dx * dx, minlength=n)
        s_q[:, 1] += np.bincount(gid, weights=sq * 2.0 * np.abs(dx * dy), minlength=n)
        s_q[:, 2] += np.bincount(gid, weights=sq * dy * dy, minlength=n)
    return s_m, s_q, s_c, s_o
