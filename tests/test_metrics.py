"""Pins structsplat_amd.metrics: SSIM equals a dense 2-D-window float64 restatement of
metrics.py:131-151, and PSNR follows metrics.py:40-50."""
from __future__ import annotations

import math

import numpy as np
import torch

from structsplat_amd import metrics as M


def _dense_ssim(p: np.ndarray, t: np.ndarray, win=11, sigma=1.5) -> float:
    """11x11 outer-product window, zero padding, mean over every pixel and channel."""
    x = np.arange(win) - (win - 1) / 2.0
    g = np.exp(-(x ** 2) / (2 * sigma ** 2))
    g /= g.sum()
    k = np.outer(g, g)
    pad = win // 2
    H, W, _ = p.shape

    def blur(img):
        z = np.pad(img, ((pad, pad), (pad, pad), (0, 0)))
        out = np.zeros_like(img)
        for dy in range(win):
            for dx in range(win):
                out += k[dy, dx] * z[dy:dy + H, dx:dx + W]
        return out

    mu_p, mu_t = blur(p), blur(t)
    sig_p = blur(p * p) - mu_p ** 2
    sig_t = blur(t * t) - mu_t ** 2
    sig_pt = blur(p * t) - mu_p * mu_t
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    s = ((2 * mu_p * mu_t + c1) * (2 * sig_pt + c2)) / ((mu_p ** 2 + mu_t ** 2 + c1) *
                                                       (sig_p + sig_t + c2))
    return float(s.mean())


def test_ssim_matches_dense_window_restatement():
    rng = np.random.default_rng(0)
    t = rng.random((23, 31, 3))
    p = np.clip(t + rng.normal(0, 0.1, t.shape), 0, 1)
    got = float(M.ssim(torch.tensor(p), torch.tensor(t)))
    assert abs(got - _dense_ssim(p, t)) < 1e-10


def test_ssim_identity_and_cached_stats():
    rng = np.random.default_rng(1)
    t = torch.tensor(rng.random((16, 20, 3)))
    assert abs(float(M.ssim(t, t)) - 1.0) < 1e-12
    p = (t + 0.05 * torch.randn_like(t)).clamp(0, 1)
    stats = M.SSIMTargetStats(t)
    assert float(M.ssim(p, stats=stats)) == float(M.ssim(p, t))


def test_psnr_closed_form():
    t = torch.zeros(4, 4, 3)
    p = torch.full((4, 4, 3), 0.1)
    assert abs(M.psnr(p, t) - 10 * math.log10(1 / 0.01)) < 1e-4
    assert M.psnr(t, t) == 10 * math.log10(1 / 1e-12)
