"""Seeded synthetic fields and targets of the benchmark workloads.

All generators use ``numpy.random.default_rng`` with an explicit seed and return float32 NumPy
arrays in the reference's field layout (means (N,2) pixel coordinates, scales (N,2) standard
deviations in pixels, angles (N,), colors (N,3), opacity values (N,)).

Workloads (image side / Gaussian count):

* ``cpu_reference``: 256^2 / 2,000 - log-scales U[ln 0.5, ln 8] per axis, angle U[0, pi),
  colors U[0,1], opacity values U[0.1, 1] (the reference's own CPU-runnable case).
* ``single``: 1024^2 / 50,000 and ``metric``: 2048^2 / 200,000 - scale s log-uniform in
  [0.5, 12] px, aspect ratio a log-uniform in [1, 10], major axis s and minor axis s/a,
  orientation from a smoothed random angle field (doubled-angle vector field blurred with a
  64 px Gaussian, theta = atan2/2 at each mean), colors U[0,1], opacity values U[0.1, 1].
* ``stress``: 4096^2 / 1,000,000 - s log-normal around 2 px (sigma 1.0) clipped to 64 px,
  aspect ratio log-uniform in [1, 20], same angle field.
* ``fit_target``: sum of 32 random sinusoids plus 200 random straight step edges, mapped
  affinely to [0, 1]; ``fit_init_field`` the fit's initial field.
* ``batched_fit_workload``: 64 fields, sides uniform over {256, 512, 768, 1024}, Gaussian
  counts log-uniform in [1k, 60k], field k's target and initial field seeded with 1000 + k.
"""
from __future__ import annotations

import math

import numpy as np


def _angle_field(rng, H: int, W: int, means: np.ndarray, corr_px: float = 64.0,
                 cell: int = 16) -> np.ndarray:
    """Orientation per mean from a smooth random doubled-angle field (blurred on a coarse grid
    of `cell`-pixel cells, sampled bilinearly)."""
    from scipy.ndimage import gaussian_filter, map_coordinates

    gh = max(2, (H + cell - 1) // cell + 1)
    gw = max(2, (W + cell - 1) // cell + 1)
    phi = rng.uniform(0.0, math.pi, (gh, gw))
    vx = gaussian_filter(np.cos(2 * phi), corr_px / cell, mode="wrap")
    vy = gaussian_filter(np.sin(2 * phi), corr_px / cell, mode="wrap")
    coords = np.stack([means[:, 1] / cell, means[:, 0] / cell])
    sx = map_coordinates(vx, coords, order=1, mode="nearest")
    sy = map_coordinates(vy, coords, order=1, mode="nearest")
    return (0.5 * np.arctan2(sy, sx)).astype(np.float32)


def _finish(means, scales, angles, colors, opacity) -> dict[str, np.ndarray]:
    return {
        "means": means.astype(np.float32),
        "scales": scales.astype(np.float32),
        "angles": angles.astype(np.float32),
        "colors": colors.astype(np.float32),
        "opacity": opacity.astype(np.float32),
    }


def cpu_reference_field(seed: int = 0, H: int = 256, W: int = 256, n: int = 2000):
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(0, W, n), rng.uniform(0, H, n)], 1)
    scales = np.exp(rng.uniform(math.log(0.5), math.log(8.0), (n, 2)))
    angles = rng.uniform(0.0, math.pi, n)
    colors = rng.uniform(0.0, 1.0, (n, 3))
    opacity = rng.uniform(0.1, 1.0, n)
    return _finish(means, scales, angles, colors, opacity)


def anisotropic_field(seed: int, H: int, W: int, n: int, scale_lo: float = 0.5,
                      scale_hi: float = 12.0, aspect_hi: float = 10.0):
    """The single-image / metric recipe (major = s, minor = s / a)."""
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(0, W, n), rng.uniform(0, H, n)], 1)
    s = np.exp(rng.uniform(math.log(scale_lo), math.log(scale_hi), n))
    a = np.exp(rng.uniform(0.0, math.log(aspect_hi), n))
    scales = np.stack([s, s / a], 1)
    angles = _angle_field(rng, H, W, means)
    colors = rng.uniform(0.0, 1.0, (n, 3))
    opacity = rng.uniform(0.1, 1.0, n)
    return _finish(means, scales, angles, colors, opacity)


def single_field(seed: int = 0):
    return anisotropic_field(seed, 1024, 1024, 50_000)


def metric_field(seed: int = 0):
    return anisotropic_field(seed, 2048, 2048, 200_000)


def stress_field(seed: int = 0, H: int = 4096, W: int = 4096, n: int = 1_000_000):
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(0, W, n), rng.uniform(0, H, n)], 1)
    s = np.minimum(np.exp(rng.normal(math.log(2.0), 1.0, n)), 64.0)
    a = np.exp(rng.uniform(0.0, math.log(20.0), n))
    scales = np.stack([s, s / a], 1)
    angles = _angle_field(rng, H, W, means)
    colors = rng.uniform(0.0, 1.0, (n, 3))
    opacity = rng.uniform(0.1, 1.0, n)
    return _finish(means, scales, angles, colors, opacity)


def _target_terms(seed: int, n_waves: int, n_edges: int):
    """The random draws of a fit target, in draw order: per wave (theta, freq, phase, amp[3]),
    per edge (px, py, theta, step[3]) with px, py in units of W and H."""
    rng = np.random.default_rng(seed)
    waves = []
    for _ in range(n_waves):
        theta = rng.uniform(0, 2 * math.pi)
        freq = rng.uniform(1.0, 64.0)
        phase = rng.uniform(0, 2 * math.pi)
        amp = rng.uniform(0.0, 1.0, 3) / n_waves
        waves.append((theta, freq, phase, amp))
    edges = []
    for _ in range(n_edges):
        px, py = rng.uniform(0, 1), rng.uniform(0, 1)
        theta = rng.uniform(0, 2 * math.pi)
        step = rng.uniform(-0.3, 0.3, 3)
        edges.append((px, py, theta, step))
    return waves, edges


def fit_target(seed: int, H: int, W: int, n_waves: int = 32, n_edges: int = 200,
               device=None):
    """(H, W, 3) float32 target in [0, 1]: random sinusoids plus random straight step edges,
    evaluated in float64.  With ``device`` (a torch device) the target is evaluated there with
    torch and returned as a tensor on it; otherwise a NumPy array."""
    waves, edges = _target_terms(seed, n_waves, n_edges)
    if device is not None:
        import torch

        y, x = torch.meshgrid(torch.arange(H, dtype=torch.float64, device=device),
                              torch.arange(W, dtype=torch.float64, device=device),
                              indexing="ij")
        sin, zeros = torch.sin, torch.zeros((H, W, 3), dtype=torch.float64, device=device)
        as_vec = lambda a: torch.tensor(a, dtype=torch.float64, device=device)  # noqa: E731
    else:
        y, x = np.mgrid[0:H, 0:W].astype(np.float64)
        sin, zeros = np.sin, np.zeros((H, W, 3))
        as_vec = np.asarray
    u = x / W
    v = y / H
    img = zeros
    for theta, freq, phase, amp in waves:
        arg = 2 * math.pi * freq * (math.cos(theta) * u + math.sin(theta) * v) + phase
        img = img + sin(arg)[..., None] * as_vec(amp)
    for px, py, theta, step in edges:
        side = (x - px * W) * math.cos(theta) + (y - py * H) * math.sin(theta) > 0
        img = img + side[..., None] * as_vec(step)
    lo, hi = float(img.min()), float(img.max())
    out = (img - lo) / max(hi - lo, 1e-12)
    return out.float() if device is not None else out.astype(np.float32)


def fit_init_field(target, n: int, seed: int):
    """Fit initialisation: uniform means, isotropic s = 0.5 sqrt(HW/N), angle 0, colors sampled
    from the target (array or tensor) at the nearest pixel, opacity logits 0 (opacity 0.5)."""
    rng = np.random.default_rng(seed)
    H, W = target.shape[:2]
    means = np.stack([rng.uniform(0, W, n), rng.uniform(0, H, n)], 1)
    s = 0.5 * math.sqrt(H * W / n)
    scales = np.full((n, 2), s)
    angles = np.zeros(n)
    ix = np.clip(np.rint(means[:, 0]).astype(np.int64), 0, W - 1)
    iy = np.clip(np.rint(means[:, 1]).astype(np.int64), 0, H - 1)
    if isinstance(target, np.ndarray):
        colors = target[iy, ix]
    else:
        import torch

        flat = torch.from_numpy(iy * W + ix).to(target.device)
        colors = target.reshape(-1, 3)[flat].cpu().numpy()
    return _finish(means, scales, angles, colors, np.zeros(n))


def batched_fit_workload(num_fields: int = 64, seed: int = 0, device=None):
    """The ragged batched-fit workload: ``(sides, counts, targets, inits)`` with fields sorted by
    side (stable), ``targets[k]`` an (s, s, 3) fit target (a tensor on ``device`` when given)
    and ``inits[k]`` its initial field (``fit_init_field``), both seeded with 1000 + k."""
    rng = np.random.default_rng(seed)
    sides = rng.choice([256, 512, 768, 1024], num_fields)
    counts = np.exp(rng.uniform(np.log(1000), np.log(60000), num_fields)).astype(int)
    order = np.argsort(sides, kind="stable")
    sides, counts = sides[order], counts[order]
    targets, inits = [], []
    for k, (side, n) in enumerate(zip(sides, counts)):
        t = fit_target(1000 + k, int(side), int(side), device=device)
        targets.append(t)
        inits.append(fit_init_field(t, int(n), 1000 + k))
    return sides, counts, targets, inits
