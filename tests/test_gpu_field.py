"""GPU parity of the fused field producer (conics, radii, opacities and their backward) against
the torch expressions of the reference's GaussianField (gaussians.py:270-321)."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from structsplat_amd.gaussians import GaussianField

pytestmark = pytest.mark.gpu


def _field(device, n, seed, opacity=True, fv=False):
    rng = np.random.default_rng(seed)
    means = rng.uniform(0, 100, (n, 2))
    scales = np.exp(rng.uniform(np.log(0.2), np.log(40.0), (n, 2)))
    angles = rng.uniform(-2 * np.pi, 2 * np.pi, n)
    colors = rng.random((n, 3))
    logits = rng.normal(0, 2, n) if opacity else None
    f = GaussianField.from_numpy(means, scales, angles, colors, opacities=logits, device=device,
                                 filter_variance=rng.uniform(0, 0.5, n) if fv else None)
    return f


@pytest.mark.parametrize("fv", [False, True])
@pytest.mark.parametrize("dilation", [0.0, 0.3])
def test_radii_bit_identical_and_conics_close(cuda, fv, dilation):
    print(f"case: produce fv={fv} dilation={dilation}", flush=True)
    f = _field(cuda, 200_000, 7, fv=fv)
    # adversarial rotations: exact multiples of pi/4 and values near them
    f.rotations[:1000] = torch.arange(1000, device=cuda, dtype=torch.float32) * (math.pi / 4)
    conics, radii, op = f.produce(3.0, dilation)
    ref_r = f.radii(3.0, dilation)
    ref_q = f.conics(dilation).detach()
    ref_o = f.opacity_values().detach()
    torch.cuda.synchronize()
    assert radii.dtype == torch.int64
    mismatch = int((radii != ref_r).any(dim=1).sum())
    assert mismatch == 0, f"{mismatch} radii differ"
    assert torch.allclose(conics, ref_q, rtol=2e-6, atol=1e-12)
    assert torch.allclose(op, ref_o, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("opacity", [True, False])
def test_backward_matches_torch_autograd(cuda, opacity):
    f = _field(cuda, 5000, 3, opacity=opacity)
    f.trainable()
    g = torch.Generator(device="cpu").manual_seed(0)
    gq = torch.randn(5000, 3, generator=g).to(cuda)
    go = torch.randn(5000, generator=g).to(cuda)
    conics, _, op = f.produce(3.0)
    loss = (conics * gq).sum() + ((op * go).sum() if opacity else 0.0)
    loss.backward()
    mine = [p.grad.clone() for p in f.parameters()[1:3]] + ([f.opacities.grad.clone()]
                                                             if opacity else [])
    for p in f.parameters():
        p.grad = None
    ref_loss = (f.conics() * gq).sum() + ((f.opacity_values() * go).sum() if opacity else 0.0)
    ref_loss.backward()
    ref = [p.grad for p in f.parameters()[1:3]] + ([f.opacities.grad] if opacity else [])
    torch.cuda.synchronize()
    for a, b in zip(mine, ref):
        scale = b.abs().max().clamp_min(1e-30)
        assert float((a - b).abs().max() / scale) < 1e-5
