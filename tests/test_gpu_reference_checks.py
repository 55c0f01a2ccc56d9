"""The reference's own GPU checks of this path, restated for the HIP renderer.

* Ellipse cull (tests/test_render.py:417-454): with support fade the culled pairs carry weights
  that are exactly zero, so images and gradients with the cull equal those without it (the
  reference's tolerance is 1e-6; the cull must also actually drop pairs on this field).
* Parity matrix (tests/test_render.py:457-497): normalized / additive x opacities x dilation
  {0, 0.3} x support fade, image and the gradients of means, log-scales, colors and opacity
  logits through render_field, against the reference's float32 CPU renderer (the oracle's
  restatement of render.py with torch autograd) at the reference's tolerances (2e-4 image,
  3e-4 gradients), and the image against the float64 oracle at 1e-5.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import exact64
from oracle import reference_render as R
from structsplat_amd import render as HR
from structsplat_amd.gaussians import GaussianField
from tests.helpers import scaled_err, to_np

pytestmark = pytest.mark.gpu


def test_ellipse_cull_matches_unculled(cuda):
    print("case: ellipse cull on vs off (seed 22, 40x44, N=48)", flush=True)
    rng = np.random.default_rng(22)
    N, H, W = 48, 40, 44
    means_np = np.stack([rng.uniform(-3, W + 3, N), rng.uniform(-3, H + 3, N)], 1)
    sx = np.exp(rng.uniform(np.log(2.0), np.log(8.0), N))
    sy = np.exp(rng.uniform(np.log(0.5), np.log(1.2), N))
    scales = np.stack([sx, sy], 1)
    angles = rng.uniform(0, np.pi, N)
    colors = rng.random((N, 3))
    target = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(2))

    def run(cull: bool):
        f = GaussianField.from_numpy(means_np, scales, angles, colors, device=cuda).trainable()
        img = HR.render_cuda_exact(f.means, f.conics(), f.colors, f.radii(3.0), H, W,
                                   tiled=True, support_fade=True, sigma_cutoff=3.0,
                                   tile_ellipse_cull=cull)
        (img - target).square().mean().backward()
        torch.cuda.synchronize()
        return (img.detach(), f.means.grad.detach(), f.log_scales.grad.detach(),
                f.rotations.grad.detach(), f.colors.grad.detach())

    unculled = run(False)
    culled = run(True)
    for got, exp in zip(culled, unculled):
        assert torch.allclose(got, exp, atol=1e-6, rtol=1e-6)
    # The cull fires on this field: fewer (tile, Gaussian) pairs with it than without.
    f = GaussianField.from_numpy(means_np, scales, angles, colors, device=cuda)
    with torch.no_grad():
        q, r = f.conics(), f.radii(3.0)
    g_off, _ = HR.build_tile_index(f.means, q, r, H, W, 16, False, 3.0)
    g_on, _ = HR.build_tile_index(f.means, q, r, H, W, 16, True, 3.0)
    print("pairs without / with cull", g_off.numel(), g_on.numel())
    assert g_on.numel() < g_off.numel()


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("use_opacities", [False, True])
@pytest.mark.parametrize("dilation", [0.0, 0.3])
@pytest.mark.parametrize("support_fade", [False, True])
def test_parity_matrix(cuda, normalize, use_opacities, dilation, support_fade):
    print(f"case: parity matrix normalize={normalize} opacities={use_opacities} "
          f"dilation={dilation} fade={support_fade}", flush=True)
    rng = np.random.default_rng(7)
    N, H, W = 20, 16, 18
    means = np.stack([rng.uniform(-2, W + 2, N), rng.uniform(-2, H + 2, N)], 1)
    scales = np.exp(rng.uniform(np.log(0.7), np.log(3.5), (N, 2)))
    angles = rng.uniform(0, np.pi, N)
    colors = rng.random((N, 3))
    opac = rng.uniform(0.1, 1.0, N) if use_opacities else None
    target = torch.rand(H, W, 3, generator=torch.Generator().manual_seed(7))

    # The integer radii (constants of the render) from the GPU's float32 formula for both runs,
    # so a ceil that CPU and GPU libm round differently cannot move a rectangle edge.
    radii_all = GaussianField.from_numpy(means, scales, angles, colors,
                                         device=cuda).radii(3.0, dilation)

    def run(device, mode):
        f = GaussianField.from_numpy(means, scales, angles, colors, opacities=opac,
                                     device=device).trainable()
        o = None if opac is None else f.opacity_values()
        conics, radii = f.conics(dilation), radii_all.to(device)
        if device == "cpu":
            fn = R.render if normalize else R.render_additive
            img = fn(f.means, conics, f.colors, radii, H, W, opacities=o,
                     support_fade=support_fade, sigma_cutoff=3.0)
        else:
            img = HR.render_field(f.means, conics, f.colors, radii, H, W, mode=mode, opacities=o,
                                  support_fade=support_fade, sigma_cutoff=3.0)
        loss = (img - target.to(device)).square().mean()
        loss.backward()
        torch.cuda.synchronize()
        grads = [f.means.grad, f.log_scales.grad, f.colors.grad]
        if opac is not None:
            grads.append(f.opacities.grad)
        inputs = [to_np(t) for t in (f.means, conics, f.colors, radii)]
        inputs.append(None if o is None else to_np(o))
        return img.detach().cpu(), [g.detach().cpu() for g in grads], inputs

    ref_img, ref_g, _ = run("cpu", None)
    hip_img, hip_g, inputs = run(cuda, "hip" if normalize else "hip_additive")
    assert torch.allclose(hip_img, ref_img, atol=2e-4, rtol=2e-4)
    for got, exp in zip(hip_g, ref_g):
        assert torch.allclose(got, exp, atol=3e-4, rtol=3e-4)
    exact, _ = exact64.forward(*inputs, H, W, normalize, 1e-8, support_fade, 3.0)
    err = scaled_err(hip_img.numpy(), exact)
    print("image vs float64 scaled err", err)
    assert err <= 1e-5
