"""The persistent render schedule (SSPLAT_TILE_PERSISTENT: one launch of resident workgroups
taking tiles from per-XCD queues) gives the grid launch's results bit for bit: forwards,
coverage and every gradient, for single fields and ragged batches, tile edges 16 and 32."""
from __future__ import annotations

import pytest
import torch

from structsplat_amd import batch as B
from structsplat_amd import render as HR
from tests.helpers import field_tensors, random_field

pytestmark = pytest.mark.gpu

P = HR.TILE_PERSISTENT


@pytest.mark.parametrize("ts", [16, 32])
@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True, 0.5])
def test_persistent_single_field_equals_grid(cuda, ts, normalize, fade):
    print(f"case: persistent single field ts={ts} normalize={normalize} fade={fade}", flush=True)
    N, H, W = 3000, 211, 187
    f = random_field(31, N, H, W, scale_lo=0.6, scale_hi=10.0, opacity=True)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    alpha = HR._fade_alpha(fade)
    gids, offsets = HR.build_tile_index(means, conics, radii, H, W, ts, alpha == 1.0, 3.0)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(2))
    results = []
    for cfg in (ts, ts | P):
        den = torch.empty(H, W, device=cuda)
        out, _ = HR.forward_tiled(means, conics, colors, radii, op, gids, offsets, H, W,
                                  normalize, 1e-8, cfg, alpha, 3.0, den=den)
        grads = []
        for d, o in ((den, out), (None, out), (None, None)) if normalize else ((None, None),):
            grads.append(HR.backward_tiled(g, means, conics, colors, radii, op, d, o, gids,
                                           offsets, H, W, normalize, 1e-8, cfg, alpha, 3.0))
        torch.cuda.synchronize()
        results.append((out, den, grads))
    (o0, d0, g0), (o1, d1, g1) = results
    assert torch.equal(o0, o1) and torch.equal(d0, d1)
    for a, b in zip(g0, g1):
        for x, y in zip(a, b):
            assert torch.equal(x, y)


@pytest.mark.parametrize("ts", [16, 32])
def test_persistent_batch_equals_grid(cuda, ts):
    print(f"case: persistent batch ts={ts}", flush=True)
    sizes = [(37, 53), (64, 64), (16, 16), (9, 130), (200, 170)]
    counts = [150, 300, 0, 40, 2200]
    parts = []
    for k, ((h, w), n) in enumerate(zip(sizes, counts)):
        f = random_field(300 + k, max(n, 1), h, w, scale_lo=0.6, scale_hi=8.0, opacity=True)
        parts.append([x[:n] for x in field_tensors(f, cuda)])
    means, conics, colors, radii, op = (torch.cat([p[i] for p in parts], 0).contiguous()
                                        for i in range(5))
    outs = []
    for cfg in (ts, ts | P):
        b = B.Batch(sizes, counts, cuda, tile_size=cfg)
        gids, offsets, status = b.build_tile_index(means, conics, radii, 1 << 17)
        grad = torch.rand(b.total_pixels, 3, device=cuda,
                          generator=torch.Generator(device=cuda).manual_seed(6))
        out = b.forward(means, conics, colors, radii, op, gids, offsets)
        grads = b.backward(grad, means, conics, colors, radii, op, None, out, gids, offsets,
                           after_forward=True)
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        outs.append((out, grads))
    assert torch.equal(outs[0][0], outs[1][0])
    for x, y in zip(outs[0][1], outs[1][1]):
        assert torch.equal(x, y)


def test_persistent_empty_field_renders_zero(cuda):
    """No Gaussians: the schedule falls back to the grid launch, which writes zeros."""
    H, W = 40, 24
    empty = torch.empty(0, 2, device=cuda)
    gids, offsets = HR.build_tile_index(empty, torch.empty(0, 3, device=cuda),
                                        torch.empty(0, 2, dtype=torch.int64, device=cuda), H, W,
                                        16)
    out, _ = HR.forward_tiled(empty, torch.empty(0, 3, device=cuda),
                              torch.empty(0, 3, device=cuda),
                              torch.empty(0, 2, dtype=torch.int64, device=cuda), None, gids,
                              offsets, H, W, True, 1e-8, 16 | P, False, 3.0,
                              out=torch.full((H, W, 3), 7.0, device=cuda))
    torch.cuda.synchronize()
    assert torch.count_nonzero(out) == 0
