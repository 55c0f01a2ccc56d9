"""Autotuner of the render path's tile shape and schedule.

The render kernels run one workgroup per square tile, one thread per pixel and one wave per 8x8
quadrant, for 16-pixel tiles (4 waves) or 32-pixel tiles (16 waves).  Larger tiles give fewer
(tile, Gaussian) pairs: less binning work and fewer per-pair gradient reductions in the backward.
They also give longer lists that every wave of a tile filters.  The schedule is a grid launch
(one workgroup per tile, dispatched by the hardware in XCD-aware order) or a persistent launch
(as many workgroups as the device holds, taking tiles from per-XCD queues, render.TILE_PERSISTENT).
Which configuration wins depends on the Gaussian sizes and, for batches, on the spread of field
sizes, so the tuner times the fit step's render work on the actual field: binning, forward,
backward in the fit's normalizer mode, for each candidate, and keeps the fastest.

Results are cached for the process by (H, W, N, overlap bucket) and the render mode: the overlap
is the mean number of Gaussians per 16-pixel tile, bucketed in half octaves, so fields of the
same size and density reuse one measurement.  A batch's key is its field count, total pixels,
total Gaussians and overlap.
"""
from __future__ import annotations

import math
import statistics

import torch

from . import render as HR

CANDIDATE_TILE_SIZES = (16, 32, 16 | HR.TILE_PERSISTENT, 32 | HR.TILE_PERSISTENT)
_CACHE: dict = {}


def describe(config: int) -> str:
    """'16', '32', '16/persistent', ... for a tile configuration."""
    edge = HR.tile_edge(config)
    return f"{edge}/persistent" if config & HR.TILE_PERSISTENT else str(edge)


def overlap_bucket(pairs16: int, tiles16: int) -> int:
    """Half-octave bucket of the mean Gaussians per 16-pixel tile."""
    per_tile = max(pairs16, 1) / max(tiles16, 1)
    return int(round(2.0 * math.log2(per_tile)))


def _time_ms(fn, reps: int) -> float:
    stream = torch.cuda.current_stream()
    fn()  # warm-up
    times = []
    for _ in range(reps):
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        fn()
        stop.record(stream)
        stop.synchronize()
        times.append(start.elapsed_time(stop))
    return statistics.median(times)


@torch.no_grad()
def tune_tile_size(sizes, counts, means, conics, colors, radii, opacities, *, normalize=True,
                   eps=1e-8, support_fade=False, sigma_cutoff=3.0, recompute_normalizer=True,
                   candidates=CANDIDATE_TILE_SIZES, reps=5):
    """Fastest tile configuration (edge, with render.TILE_PERSISTENT for the persistent
    schedule) for rendering (fwd+bwd) the fields `sizes` [(H, W)] with `counts` Gaussians each
    (rows of `means`, ... concatenated in field order).

    Returns ``(config, {config: ms})``; cached by (H, W, N, overlap bucket) and mode.
    """
    from .batch import Batch

    dev = means.device
    alpha = HR._fade_alpha(support_fade)
    cull = alpha == 1.0
    op = opacities if opacities is not None else means.new_empty(0)
    means, conics, colors = means.contiguous(), conics.contiguous(), colors.contiguous()
    radii = HR._radii64(radii)
    b16 = Batch(sizes, counts, dev, tile_size=16)
    pairs16 = b16.pair_count(means, conics, radii, cull, sigma_cutoff)
    bucket = overlap_bucket(pairs16, b16.total_tiles)
    if len(sizes) == 1:
        shape = (int(sizes[0][0]), int(sizes[0][1]), int(counts[0]))
    else:
        shape = (len(sizes), b16.total_pixels, int(sum(int(c) for c in counts)))
    key = (shape, bucket, bool(normalize), alpha, float(sigma_cutoff),
           bool(recompute_normalizer), tuple(candidates))
    hit = _CACHE.get(key)
    if hit is not None:
        return hit
    steps = []
    for ts in candidates:
        b = Batch(sizes, counts, dev, tile_size=ts)
        pairs = b.pair_count(means, conics, radii, cull, sigma_cutoff)
        gids, offsets, status = b.build_tile_index(means, conics, radii, pairs + 1, cull,
                                                   sigma_cutoff)
        den = (torch.empty(b.total_pixels, device=dev)
               if normalize and not recompute_normalizer else None)
        grad = torch.rand(b.total_pixels, 3, device=dev,
                          generator=torch.Generator(device=dev).manual_seed(0))
        out = torch.empty(b.total_pixels, 3, device=dev)
        bufs = (gids, offsets, status)
        slots = HR.grad_slots(pairs + 1, dev)

        def step(b=b, gids=gids, offsets=offsets, bufs=bufs, den=den, grad=grad, out=out,
                 slots=slots, status=status, cap=pairs + 1):
            b.build_tile_index(means, conics, radii, cap, cull, sigma_cutoff, out=bufs)
            b.forward(means, conics, colors, radii, op, gids, offsets, normalize, eps, alpha,
                      sigma_cutoff, out=out, den=den)
            b.backward(grad, means, conics, colors, radii, op, den, out, gids, offsets,
                       normalize, eps, alpha, sigma_cutoff, slots=slots, slot_capacity=cap,
                       status=status, after_forward=True)

        steps.append((ts, step))
    # Every candidate runs twice before any is timed: the first launches of a process meet a GPU
    # that is still raising its clocks, which would penalize whichever candidate came first.
    for _, step in steps:
        step()
        step()
    torch.cuda.synchronize()
    timings = {ts: _time_ms(step, reps) for ts, step in steps}
    best = min(timings, key=timings.get)
    _CACHE[key] = (best, timings)
    return best, timings


def tune_field(field, H: int, W: int, cfg):
    """Tile configuration for fitting one field (the fit step's render configuration)."""
    conics, radii, op = field.produce(cfg.sigma_cutoff)
    return tune_tile_size([(H, W)], [field.n], field.means.detach(), conics.detach(),
                          field.colors.detach(), radii,
                          None if op is None else op.detach(), normalize=cfg.normalize,
                          eps=cfg.normalization_eps, support_fade=cfg.support_fade,
                          sigma_cutoff=cfg.sigma_cutoff,
                          recompute_normalizer=cfg.recompute_normalizer)
