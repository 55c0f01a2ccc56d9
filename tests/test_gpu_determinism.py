"""GPU tests of the deterministic backward: the gradients do not depend on scheduling.

Every (tile, Gaussian) pair writes its partial gradients to its own slot and each Gaussian's
slots are summed in an order fixed by the Gaussian alone, so repeated backwards, a batched
backward and its per-field backwards, and graph replays and eager steps agree bit for bit.
"""
from __future__ import annotations

import pytest
import torch

from structsplat_amd import batch as B
from structsplat_amd import fit as F
from structsplat_amd import render as HR
from structsplat_amd import synth
from structsplat_amd.gaussians import GaussianField
from tests.helpers import field_tensors, random_field

pytestmark = pytest.mark.gpu


def _case(**kw):
    print("case:", ", ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _field(device, seed, N, H, W, scale_hi=9.0):
    f = random_field(seed, N, H, W, scale_lo=0.6, scale_hi=scale_hi, opacity=True)
    return field_tensors(f, device)


@pytest.mark.parametrize("ts", [16, 32])
@pytest.mark.parametrize("fade", [False, True])
def test_repeated_backwards_are_bitwise_equal(cuda, ts, fade):
    """Dense overlap (many Gaussians per pixel, many tiles per Gaussian, some with more than 64
    slots): two backwards with fresh slot buffers and a third reusing one are identical."""
    _case(test="repeated backward", tile=ts, fade=fade)
    H, W, N = 150, 190, 1500
    means, conics, colors, radii, op = _field(cuda, 3, N, H, W, scale_hi=40.0)
    alpha = HR._fade_alpha(fade)
    cull = alpha == 1.0
    gids, offsets = HR.build_tile_index(means, conics, radii, H, W, ts, cull, 3.0)
    out, _ = HR.forward_tiled(means, conics, colors, radii, op, gids, offsets, H, W, True, 1e-8,
                              ts, alpha, 3.0)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(1))
    slots_needed = HR.rect_pair_count(means, radii, H, W, ts)
    assert slots_needed >= gids.numel()
    if not cull:
        assert slots_needed == gids.numel()
    runs = []
    shared = HR.grad_slots(slots_needed, cuda)
    for k in range(3):
        slots = shared if k == 2 else None
        grads = HR.backward_tiled(g, means, conics, colors, radii, op, None, out, gids, offsets,
                                  H, W, True, 1e-8, ts, alpha, 3.0, slots=slots,
                                  slot_capacity=slots_needed if slots is not None else None)
        torch.cuda.synchronize()
        runs.append([t.clone() for t in grads])
        if k == 2:  # the shared buffer's second call must ignore the first call's slots
            again = HR.backward_tiled(g, means, conics, colors, radii, op, None, out, gids,
                                      offsets, H, W, True, 1e-8, ts, alpha, 3.0, slots=shared,
                                      slot_capacity=slots_needed)
            torch.cuda.synchronize()
            runs.append([t.clone() for t in again])
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert torch.equal(a, b)


@pytest.mark.parametrize("ts", [16, 32])
def test_batched_backward_is_bitwise_equal_to_per_field(cuda, ts):
    _case(test="batched vs per-field backward bits", tile=ts)
    sizes = [(37, 53), (64, 64), (9, 130), (70, 41)]
    counts = [150, 300, 40, 220]
    parts = [_field(cuda, 40 + k, n, h, w) for k, ((h, w), n) in enumerate(zip(sizes, counts))]
    b = B.Batch(sizes, counts, cuda, tile_size=ts)
    cat = [torch.cat([p[i] for p in parts], 0).contiguous() for i in range(5)]
    means, conics, colors, radii, op = cat
    gen = torch.Generator(device=cuda).manual_seed(5)
    grad = torch.rand(b.total_pixels, 3, device=cuda, generator=gen)
    gids, offsets, _ = b.build_tile_index(means, conics, radii, 1 << 16)
    out = b.forward(means, conics, colors, radii, op, gids, offsets)
    gm, gq, gc, go = b.backward(grad, means, conics, colors, radii, op, None, out, gids, offsets)
    torch.cuda.synchronize()
    for k, (p, (h, w)) in enumerate(zip(parts, sizes)):
        fg, fo = HR.build_tile_index(p[0], p[1], p[3], h, w, ts)
        fout, _ = HR.forward_tiled(*p[:5], fg, fo, h, w, True, 1e-8, ts, False, 3.0)
        ref = HR.backward_tiled(grad[b.pixel_slices[k]].view(h, w, 3).contiguous(), *p[:5],
                                None, fout, fg, fo, h, w, True, 1e-8, ts, False, 3.0)
        sl = b.gauss_slices[k]
        for got, exp in zip((gm[sl], gq[sl], gc[sl], go[sl]), ref):
            assert torch.equal(got, exp), k


@pytest.mark.parametrize("ts", [16, 32])
@pytest.mark.parametrize("fade", [False, True, 0.5])
def test_after_forward_backward_is_bitwise_equal(cuda, ts, fade):
    """The backward that reuses the forward's packed records (single field and batch) gives the
    plain backward's gradients bit for bit, in every normalizer mode, and twice in a row."""
    _case(test="after-forward backward", tile=ts, fade=fade)
    H, W, N = 96, 120, 700
    means, conics, colors, radii, op = _field(cuda, 23, N, H, W, scale_hi=20.0)
    alpha = HR._fade_alpha(fade)
    gids, offsets = HR.build_tile_index(means, conics, radii, H, W, ts, alpha == 1.0, 3.0)
    cap = HR.rect_pair_count(means, radii, H, W, ts)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(8))
    for keep_den in (False, True):
        ws = HR.workspace(N, H, W, ts, cuda)
        den = torch.empty(H, W, device=cuda) if keep_den else None
        out, _ = HR.forward_tiled(means, conics, colors, radii, op, gids, offsets, H, W, True,
                                  1e-8, ts, alpha, 3.0, den=den, ws=ws)
        ref = HR.backward_tiled(g, means, conics, colors, radii, op, den, out, gids, offsets, H,
                                W, True, 1e-8, ts, alpha, 3.0, slot_capacity=cap)
        slots = HR.grad_slots(cap, cuda)
        for _ in range(2):
            got = HR.backward_tiled(g, means, conics, colors, radii, op, den, out, gids, offsets,
                                    H, W, True, 1e-8, ts, alpha, 3.0, ws=ws, slots=slots,
                                    slot_capacity=cap, after_forward=True)
            torch.cuda.synchronize()
            for a, e in zip(got, ref):
                assert torch.equal(a, e), keep_den
    # Batch: the batch's own workspace after its forward.
    sizes, counts = [(40, 70), (96, 120)], [150, N]
    b = B.Batch(sizes, counts, cuda, tile_size=ts)
    small = _field(cuda, 24, 150, 40, 70, scale_hi=20.0)
    cat = [torch.cat([x, y], 0).contiguous()
           for x, y in zip(small, (means, conics, colors, radii, op))]
    bg, bo, _ = b.build_tile_index(cat[0], cat[1], cat[3], 1 << 16, alpha == 1.0, 3.0)
    grad = torch.rand(b.total_pixels, 3, device=cuda,
                      generator=torch.Generator(device=cuda).manual_seed(9))
    bout = b.forward(*cat, bg, bo, True, 1e-8, alpha, 3.0)
    bref = b.backward(grad, *cat, None, bout, bg, bo, True, 1e-8, alpha, 3.0)
    bout = b.forward(*cat, bg, bo, True, 1e-8, alpha, 3.0)
    bgot = b.backward(grad, *cat, None, bout, bg, bo, True, 1e-8, alpha, 3.0, after_forward=True)
    torch.cuda.synchronize()
    for a, e in zip(bgot, bref):
        assert torch.equal(a, e)


def test_slot_capacity_overflow_sets_status(cuda):
    _case(test="slot capacity overflow")
    H, W, N = 64, 80, 200
    means, conics, colors, radii, op = _field(cuda, 9, N, H, W)
    gids, offsets = HR.build_tile_index(means, conics, radii, H, W, 16)
    out, _ = HR.forward_tiled(means, conics, colors, radii, op, gids, offsets, H, W, True, 1e-8,
                              16, False, 3.0)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(2))
    P = gids.numel()
    for cap, expect in ((P, 0), (P - 1, 2)):
        status = torch.zeros(1, dtype=torch.int32, device=cuda)
        HR.backward_tiled(g, means, conics, colors, radii, op, None, out, gids, offsets, H, W,
                          True, 1e-8, 16, False, 3.0, slots=HR.grad_slots(cap, cuda),
                          slot_capacity=cap, status=status)
        torch.cuda.synchronize()
        assert int(status.item()) == expect, cap


def test_graph_replays_and_eager_steps_are_bitwise_equal(cuda):
    """Forty graph-captured fit steps and forty eager steps from the same start end in the same
    parameters, bit for bit."""
    _case(test="graph vs eager fit bits")
    H, W, n = 96, 128, 700
    t = synth.fit_target(21, H, W)
    target = torch.tensor(t, device=cuda)
    init = synth.fit_init_field(t, n, 21)

    def field():
        return GaussianField.from_numpy(init["means"], init["scales"], init["angles"],
                                        init["colors"], opacities=init["opacity"], device=cuda)

    a, b = field(), field()
    cfg = F.FitConfig(iters=40, window=10, tile_size=16)
    ra = F.fit(a, target, cfg, graph=True)
    rb = F.fit(b, target, cfg, graph=False)
    assert ra["psnr"] == rb["psnr"]
    for pa, pb in zip(a.parameters(), b.parameters()):
        assert torch.equal(pa.detach(), pb.detach())
