"""GPU tests of the ragged batch path against per-field renders and the plain per-image loop."""
from __future__ import annotations

import dataclasses

import pytest
import torch

from structsplat_amd import batch as B
from structsplat_amd import fit as F
from structsplat_amd import render as HR
from structsplat_amd import synth
from structsplat_amd.gaussians import GaussianField
from tests.helpers import field_tensors, random_field

pytestmark = pytest.mark.gpu

SIZES = [(37, 53), (64, 64), (16, 16), (9, 130), (70, 41)]
COUNTS = [150, 300, 0, 40, 220]


def _fields(device):
    out = []
    for k, ((h, w), n) in enumerate(zip(SIZES, COUNTS)):
        f = random_field(100 + k, max(n, 1), h, w, scale_lo=0.6, scale_hi=6.0, opacity=True)
        t = field_tensors(f, device)
        out.append([x[:n] for x in t])
    return out


def _cat(parts, i):
    return torch.cat([p[i] for p in parts], 0).contiguous()


@pytest.mark.parametrize("ts", [16, 32])
@pytest.mark.parametrize("fade", [False, True])
def test_batched_index_and_forward_match_per_field(cuda, fade, ts):
    print(f"case: batch forward, fade={fade}, tile {ts}", flush=True)
    parts = _fields(cuda)
    b = B.Batch(SIZES, COUNTS, cuda, tile_size=ts)
    means, conics, colors, radii, op = (_cat(parts, i) for i in range(5))
    gids, offsets, status = b.build_tile_index(means, conics, radii, 1 << 16, fade, 3.0)
    out = b.forward(means, conics, colors, radii, op, gids, offsets, True, 1e-8, fade, 3.0)
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    offs = offsets.cpu().numpy()
    g_all = gids.cpu().numpy()
    for k, (p, (h, w)) in enumerate(zip(parts, SIZES)):
        fg, fo = HR.build_tile_index(p[0], p[1], p[3], h, w, ts, fade, 3.0)
        t0 = b.table[k].first_tile
        g0 = b.table[k].first_gaussian
        seg = offs[t0:t0 + b.table[k].num_tiles + 1]
        assert (seg - seg[0]).tolist() == fo.cpu().tolist()
        assert (g_all[seg[0]:seg[-1]] - g0).tolist() == fg.cpu().tolist()
        single = HR.render_hip_exact(*p[:4], h, w, opacities=p[4], support_fade=fade)
        assert torch.equal(b.images(out)[k], single), k


@pytest.mark.parametrize("ts", [16, 32])
def test_batched_backward_matches_per_field(cuda, ts):
    print(f"case: batch backward, tile {ts}", flush=True)
    parts = _fields(cuda)
    b = B.Batch(SIZES, COUNTS, cuda, tile_size=ts)
    means, conics, colors, radii, op = (_cat(parts, i) for i in range(5))
    gen = torch.Generator(device=cuda).manual_seed(5)
    grad = torch.rand(b.total_pixels, 3, device=cuda, generator=gen)
    gids, offsets, _ = b.build_tile_index(means, conics, radii, 1 << 16)
    den = torch.empty(b.total_pixels, device=cuda)
    out = b.forward(means, conics, colors, radii, op, gids, offsets, den=den)
    gm, gq, gc, go = b.backward(grad, means, conics, colors, radii, op, den, out, gids, offsets)
    torch.cuda.synchronize()
    for k, (p, (h, w)) in enumerate(zip(parts, SIZES)):
        if COUNTS[k] == 0:
            continue
        fg, fo = HR.build_tile_index(p[0], p[1], p[3], h, w, ts)
        fden = torch.empty(h, w, device=cuda)
        fout, _ = HR.forward_tiled(*p[:5], fg, fo, h, w, True, 1e-8, ts, False, 3.0, den=fden)
        ref = HR.backward_tiled(grad[b.pixel_slices[k]].view(h, w, 3).contiguous(), *p[:5], fden,
                                fout, fg, fo, h, w, True, 1e-8, ts, False, 3.0)
        sl = b.gauss_slices[k]
        for got, exp in zip((gm[sl], gq[sl], gc[sl], go[sl]), ref):
            scale = exp.abs().max().clamp_min(1e-30)
            assert float((got - exp).abs().max() / scale) < 1e-5, k


def test_batched_fit_matches_plain_loop(cuda):
    """Same fields, same steps: the batched graph-captured fit reaches the per-image PSNRs of the
    plain loop within 0.01 dB."""
    print("case: batched fit vs plain loop", flush=True)
    sizes = [(64, 64), (64, 64), (96, 80), (96, 80), (96, 80)]
    counts = [300, 500, 400, 800, 250]
    targets, init = [], []
    for k, ((h, w), n) in enumerate(zip(sizes, counts)):
        t = synth.fit_target(50 + k, h, w)
        targets.append(torch.tensor(t, device=cuda))
        init.append(synth.fit_init_field(t, n, 50 + k))

    def fields():
        return [GaussianField.from_numpy(f["means"], f["scales"], f["angles"], f["colors"],
                                         opacities=f["opacity"], device=cuda) for f in init]

    cfg = F.FitConfig(iters=30, window=10)
    batched = B.fit_batch(fields(), targets, cfg, graph=True)
    plain = B.fit_each(fields(), targets, dataclasses.replace(cfg, tile_size=batched.tile_size))
    print("plain", plain.psnr, plain.seconds, "batched", batched.psnr, batched.seconds)
    # Deterministic gradients and the same tile size: the same fits, bit for bit.
    assert plain.psnr == batched.psnr


def test_batched_fit_with_filter_variance_and_caps_matches_plain_loop(cuda):
    """Fields with a covariance filter variance or scale caps (and fields without either) fit in
    one batch exactly as they fit one by one: the batch carries both with the rows."""
    print("case: batched fit with filter variance and scale caps", flush=True)
    sizes = [(64, 64), (48, 80), (64, 64)]
    counts = [300, 250, 200]
    targets, init = [], []
    for k, ((h, w), n) in enumerate(zip(sizes, counts)):
        t = synth.fit_target(70 + k, h, w)
        targets.append(torch.tensor(t, device=cuda))
        init.append(synth.fit_init_field(t, n, 70 + k))

    def fields():
        out = []
        for k, f in enumerate(init):
            g = GaussianField.from_numpy(f["means"], f["scales"], f["angles"], f["colors"],
                                         opacities=f["opacity"], device=cuda)
            if k == 0:
                g.filter_variance = torch.full((g.n,), 0.35, device=cuda)
            if k == 1:
                g.scale_max = g.scales().detach() * 1.02
            out.append(g)
        return out

    cfg = F.FitConfig(iters=20, window=10, lr_scales=0.03)
    batched = B.fit_batch(fields(), targets, cfg, graph=True)
    plain = B.fit_each(fields(), targets, dataclasses.replace(cfg, tile_size=batched.tile_size))
    print("plain", plain.psnr, "batched", batched.psnr)
    assert plain.psnr == batched.psnr
