"""GPU parity of the fused L1 + SSIM loss against the torch statement of the reference's loss
(fit.py:800-823, metrics.py:88-151) evaluated in float64 on the CPU, value and gradient."""
from __future__ import annotations

import pytest
import torch

from structsplat_amd import batch as B
from structsplat_amd.loss import batch_l1_ssim_loss, l1_ssim_loss
from structsplat_amd.metrics import ssim

pytestmark = pytest.mark.gpu


def _ref_loss(img64, tgt64, w):
    l1 = (img64 - tgt64).abs().mean()
    if w == 0.0:
        return l1
    return (1.0 - w) * l1 + w * (1.0 - ssim(img64, tgt64))


def _ref(img, tgt, w):
    """Float64 loss and gradient on the CPU."""
    x = img.detach().cpu().double().requires_grad_(True)
    loss = _ref_loss(x, tgt.detach().cpu().double(), w)
    loss.backward()
    return float(loss.detach()), x.grad


def _pair(seed, H, W, device):
    g = torch.Generator(device="cpu").manual_seed(seed)
    t = torch.rand(H, W, 3, generator=g)
    # a rendered-looking image: the target plus smooth and pixel noise, some values equal
    p = (t + 0.1 * torch.randn(H, W, 3, generator=g)).clamp(0, 1)
    p[::7, ::5] = t[::7, ::5]
    return p.to(device), t.to(device)


@pytest.mark.parametrize("H,W", [(37, 53), (16, 16), (1, 9), (7, 1), (130, 70)])
@pytest.mark.parametrize("w", [0.0, 0.2, 1.0])
def test_loss_and_gradient_match_float64(cuda, H, W, w):
    print(f"case: loss H={H} W={W} w={w}", flush=True)
    p, t = _pair(H * 1000 + W, H, W, cuda)
    x = p.clone().requires_grad_(True)
    loss = l1_ssim_loss(x, t, w)
    loss.backward()
    torch.cuda.synchronize()
    ref_loss, ref_grad = _ref(p, t, w)
    assert abs(float(loss) - ref_loss) <= 2e-6 * max(1.0, abs(ref_loss))
    g = x.grad.detach().cpu().double()
    scale = ref_grad.abs().max().clamp_min(1e-30)
    err = float((g - ref_grad).abs().max() / scale)
    print("loss", float(loss), ref_loss, "grad scaled err", err)
    assert err <= 2e-5


def test_backward_scales_by_the_incoming_gradient(cuda):
    p, t = _pair(5, 33, 40, cuda)
    x = p.clone().requires_grad_(True)
    l1_ssim_loss(x, t, 0.2).backward()
    y = p.clone().requires_grad_(True)
    (2.5 * l1_ssim_loss(y, t, 0.2)).backward()
    torch.cuda.synchronize()
    assert torch.allclose(y.grad, 2.5 * x.grad, rtol=1e-6, atol=0.0)


def test_loss_only_matches_loss_with_gradient(cuda):
    p, t = _pair(3, 45, 61, cuda)
    with torch.no_grad():
        a = l1_ssim_loss(p, t, 0.3)
    b = l1_ssim_loss(p.clone().requires_grad_(True), t, 0.3)
    torch.cuda.synchronize()
    assert float(a) == float(b)


def test_batched_loss_is_the_sum_of_field_losses(cuda):
    sizes = [(20, 33), (64, 64), (5, 70), (48, 17)]
    pairs = [_pair(10 + k, h, w, cuda) for k, (h, w) in enumerate(sizes)]
    batch = B.Batch(sizes, [0] * len(sizes), cuda)
    img = torch.cat([p.reshape(-1, 3) for p, _ in pairs]).contiguous()
    tgt = torch.cat([t.reshape(-1, 3) for _, t in pairs]).contiguous()
    x = img.clone().requires_grad_(True)
    loss = batch_l1_ssim_loss(batch, x, tgt, 0.2)
    loss.backward()
    torch.cuda.synchronize()
    total = 0.0
    for k, (p, t) in enumerate(pairs):
        y = p.clone().requires_grad_(True)
        lk = l1_ssim_loss(y, t, 0.2)
        lk.backward()
        torch.cuda.synchronize()
        total += float(lk)
        # same tiles, same arithmetic: the batched gradient equals the single-image one
        assert torch.equal(x.grad[batch.pixel_slices[k]].view_as(y), y.grad), k
    assert abs(float(loss) - total) <= 1e-6 * max(1.0, abs(total))


@pytest.mark.parametrize("H,W", [(37, 53), (1, 9), (130, 70)])
def test_target_moments_match_float64_blur(cuda, H, W):
    """G*t and G*t^2 against the float64 separable blur (metrics.py's zero-padded window)."""
    from structsplat_amd.loss import target_moments
    from structsplat_amd.metrics import _blur, _to_bchw, gaussian_window

    print(f"case: target moments H={H} W={W}", flush=True)
    _, t = _pair(H * 7 + W, H, W, cuda)
    mom = target_moments(t)
    torch.cuda.synchronize()
    t64 = _to_bchw(t.detach().cpu().double())
    wh, wv = gaussian_window(11, 1.5, "cpu", torch.float64)
    for k, x in enumerate((t64, t64 * t64)):
        ref = _blur(x, wh, wv, 5)[0].permute(1, 2, 0)
        err = float((mom[k].cpu().double() - ref).abs().max())
        print("moment", k, "max abs err", err)
        assert err <= 2e-7


@pytest.mark.parametrize("w", [0.2, 1.0])
def test_loss_with_target_moments_is_bitwise_the_inline_loss(cuda, w):
    """The forward that reads precomputed target moments gives the same loss and gradient, bit
    for bit, as the one that blurs the target itself (single image and batch)."""
    from structsplat_amd.loss import batch_target_moments, target_moments

    p, t = _pair(11, 70, 45, cuda)
    x = p.clone().requires_grad_(True)
    a = l1_ssim_loss(x, t, w)
    a.backward()
    y = p.clone().requires_grad_(True)
    b = l1_ssim_loss(y, t, w, target_moments(t))
    b.backward()
    torch.cuda.synchronize()
    assert float(a) == float(b)
    assert torch.equal(x.grad, y.grad)

    sizes = [(20, 33), (64, 64), (5, 70)]
    pairs = [_pair(20 + k, h, ww, cuda) for k, (h, ww) in enumerate(sizes)]
    batch = B.Batch(sizes, [0] * len(sizes), cuda)
    img = torch.cat([q.reshape(-1, 3) for q, _ in pairs]).contiguous()
    tgt = torch.cat([r.reshape(-1, 3) for _, r in pairs]).contiguous()
    x = img.clone().requires_grad_(True)
    a = batch_l1_ssim_loss(batch, x, tgt, w)
    a.backward()
    y = img.clone().requires_grad_(True)
    b = batch_l1_ssim_loss(batch, y, tgt, w, batch_target_moments(batch, tgt))
    b.backward()
    torch.cuda.synchronize()
    assert float(a) == float(b)
    assert torch.equal(x.grad, y.grad)
