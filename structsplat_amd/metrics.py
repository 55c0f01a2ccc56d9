"""Loss and quality metrics of the fit step: PSNR and the separable-window SSIM.

Same definitions as the reference's metrics.py: PSNR from the MSE with a 1e-12 floor
(metrics.py:40-50); SSIM with an 11-tap Gaussian window (sigma 1.5) applied as two 1-D
zero-padded convolutions per channel, C1 = 0.01^2, C2 = 0.03^2, averaged over every pixel
(metrics.py:67-151), with target statistics computed once per fit (metrics.py:93-128).
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

C1 = 0.01 ** 2
C2 = 0.03 ** 2


def _to_bchw(x: torch.Tensor) -> torch.Tensor:
    return x.permute(2, 0, 1).unsqueeze(0) if x.dim() == 3 else x


def psnr_from_mse(mse_value: torch.Tensor, max_val: float = 1.0) -> torch.Tensor:
    return 10.0 * torch.log10((max_val * max_val) / mse_value.clamp_min(1e-12))


def psnr_tensor(pred: torch.Tensor, target: torch.Tensor, max_val: float = 1.0) -> torch.Tensor:
    return psnr_from_mse(F.mse_loss(pred, target), max_val)


def psnr(pred: torch.Tensor, target: torch.Tensor, max_val: float = 1.0) -> float:
    return float(psnr_tensor(pred, target, max_val))


def gaussian_window(win: int, sigma: float, device, dtype):
    """Separable window as (horizontal (3,1,1,win), vertical (3,1,win,1)) depthwise kernels."""
    x = torch.arange(win, device=device, dtype=dtype) - (win - 1) / 2.0
    g = torch.exp(-(x ** 2) / (2 * sigma ** 2))
    g = g / g.sum()
    return (g.view(1, 1, 1, win).expand(3, 1, 1, win).contiguous(),
            g.view(1, 1, win, 1).expand(3, 1, win, 1).contiguous())


def _blur(x, wh, wv, pad: int):
    x = F.conv2d(x, wh, padding=(0, pad), groups=3)
    return F.conv2d(x, wv, padding=(pad, 0), groups=3)


class SSIMTargetStats:
    """Blurred target statistics (mu_t, mu_t^2, sigma_t^2) for a fixed target."""

    def __init__(self, target: torch.Tensor, win: int = 11, sigma: float = 1.5):
        t = _to_bchw(target)
        self.win, self.sigma = win, sigma
        self.wh, self.wv = gaussian_window(win, sigma, t.device, t.dtype)
        pad = win // 2
        with torch.no_grad():
            self.t = t
            self.mu_t = _blur(t, self.wh, self.wv, pad)
            self.mu_t2 = self.mu_t * self.mu_t
            self.sig_t = _blur(t * t, self.wh, self.wv, pad) - self.mu_t2


def ssim(pred: torch.Tensor, target: torch.Tensor | None = None, win: int = 11,
         sigma: float = 1.5, stats: SSIMTargetStats | None = None) -> torch.Tensor:
    """Mean SSIM of pred against target (or against precomputed target statistics)."""
    p = _to_bchw(pred)
    if stats is None:
        stats = SSIMTargetStats(target, win, sigma)
    pad = stats.win // 2
    wh, wv = stats.wh, stats.wv
    mu_p = _blur(p, wh, wv, pad)
    mu_p2 = mu_p * mu_p
    mu_pt = mu_p * stats.mu_t
    sig_p = _blur(p * p, wh, wv, pad) - mu_p2
    sig_pt = _blur(p * stats.t, wh, wv, pad) - mu_pt
    s = ((2 * mu_pt + C1) * (2 * sig_pt + C2)) / ((mu_p2 + stats.mu_t2 + C1) *
                                                 (sig_p + stats.sig_t + C2))
    return s.mean()
