"""One GPU's share of a row-band split: each band rendered from its members alone equals the full
render's rows (bit for bit when every member's row shift is exact, within float32 rounding of
the shift otherwise), and the bands' partial gradients sum to the full gradients."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from structsplat_amd import render as HR
from structsplat_amd import shard
from tests.helpers import field_tensors, random_field, scaled_err, to_np

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True, 0.5])
@pytest.mark.parametrize("tile_size", [16, 32])
def test_bands_equal_full_render_and_gradients_sum(cuda, normalize, fade, tile_size):
    print(f"case: bands normalize={normalize} fade={fade} tile={tile_size}", flush=True)
    H, W = 384, 320
    f = random_field(21, 3000, H, W, scale_hi=12.0, opacity=True)
    leaves = field_tensors(f, cuda)
    means, conics, colors, radii, op = leaves
    kw = dict(normalize=normalize, support_fade=fade, tiled=True, tile_size=tile_size)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(4))

    full_leaves = [t.detach().clone().requires_grad_(True) for t in (means, conics, colors, op)]
    full = HR.render_hip_exact(full_leaves[0], full_leaves[1], full_leaves[2], radii, H, W,
                               opacities=full_leaves[3], **kw)
    (full * g).sum().backward()

    band_leaves = [t.detach().clone().requires_grad_(True) for t in (means, conics, colors, op)]
    total, members = 0.0, 0
    for y0, y1 in shard.band_bounds(H, 6, 32):
        img, idx = shard.render_band(band_leaves[0], band_leaves[1], band_leaves[2], radii, H, W,
                                     y0, y1, opacities=band_leaves[3], **kw)
        assert torch.equal(img, full.detach()[y0:y1]), (y0, y1)
        total = total + (img * g[y0:y1]).sum()
        members += idx.numel()
    total.backward()
    torch.cuda.synchronize()
    assert members > means.shape[0]  # Gaussians crossing band edges belong to several bands
    for name, a, b in zip(("means", "conics", "colors", "opacities"), full_leaves, band_leaves):
        err = scaled_err(to_np(b.grad), to_np(a.grad))
        print(name, "scaled err", err)
        assert err <= 1e-5, name


def test_band_without_members_renders_zero(cuda):
    """A band no Gaussian reaches has no members: its image is zero and the gradients of the
    other bands are untouched by it."""
    print("case: empty band", flush=True)
    H, W = 128, 96
    f = random_field(5, 200, 40, W, scale_hi=2.0, margin=0.0)  # means in rows [0, 40)
    means, conics, colors, radii, _ = field_tensors(f, cuda)
    means = means.detach().clone().requires_grad_(True)
    top, idx_top = shard.render_band(means, conics, colors, radii, H, W, 0, 64)
    low, idx_low = shard.render_band(means, conics, colors, radii, H, W, 96, 128)
    assert idx_low.numel() == 0 and low.shape == (32, W, 3) and torch.count_nonzero(low) == 0
    full = HR.render_hip_exact(means, conics, colors, radii, H, W)
    assert torch.equal(top, full.detach()[0:64])
    (top.sum() + low.sum()).backward()
    torch.cuda.synchronize()
    assert torch.isfinite(means.grad).all()


@pytest.mark.parametrize("normalize", [True, False])
def test_bands_with_far_halo_members(cuda, normalize):
    """Large Gaussians (sigma up to 30 px) with means above the image and far above each band
    (below y0/2, where the float32 shift rounds): the band rows match the full render's within
    float32 rounding of the shifted means, no support row is gained or lost (a moved rectangle
    centre would change edge pixels by ~1e-2 of the image scale), and the gradients still sum."""
    print(f"case: far halo bands normalize={normalize}", flush=True)
    H, W = 512, 96
    rng = np.random.default_rng(77)
    N = 400
    f = random_field(77, N, H, W, scale_lo=8.0, scale_hi=30.0, opacity=True)
    f["means"][:, 1] = rng.uniform(-60.0, H + 20.0, N).astype(np.float32)
    f["means"][:8, 1] = np.float32(5.4999995) + np.arange(8, dtype=np.float32) * 16
    leaves = field_tensors(f, cuda)
    means, conics, colors, radii, op = leaves
    kw = dict(normalize=normalize, tiled=True, tile_size=16)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(8))
    full_leaves = [t.detach().clone().requires_grad_(True) for t in (means, conics, colors, op)]
    full = HR.render_hip_exact(full_leaves[0], full_leaves[1], full_leaves[2], radii, H, W,
                               opacities=full_leaves[3], **kw)
    (full * g).sum().backward()
    band_leaves = [t.detach().clone().requires_grad_(True) for t in (means, conics, colors, op)]
    total = 0.0
    for y0, y1 in shard.band_bounds(H, 4, 16):
        img, idx = shard.render_band(band_leaves[0], band_leaves[1], band_leaves[2], radii, H, W,
                                     y0, y1, opacities=band_leaves[3], **kw)
        err = scaled_err(to_np(img), to_np(full.detach()[y0:y1]))
        print("band", y0, y1, "members", idx.numel(), "scaled err", err)
        assert err <= 1e-6, (y0, y1)
        total = total + (img * g[y0:y1]).sum()
    total.backward()
    torch.cuda.synchronize()
    for name, a, b in zip(("means", "conics", "colors", "opacities"), full_leaves, band_leaves):
        err = scaled_err(to_np(b.grad), to_np(a.grad))
        print(name, "scaled err", err)
        assert err <= 1e-5, name
