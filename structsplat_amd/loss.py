"""The fit step's image loss on the GPU: (1 - w) * L1 + w * (1 - SSIM) through ``libssplat_hip``.

Same definition as the reference's fit loss with ``pixel_loss="l1"`` (fit.py:800-823) and its
SSIM (metrics.py:88-151): L1 is the mean absolute difference over the 3 H W values, SSIM the mean
of the SSIM map with the 11-tap, sigma-1.5 Gaussian window applied per channel as two zero-padded
1-D passes.  The forward computes the loss and keeps the SSIM derivative maps; the backward turns
them into the image gradient, scaled on the device by the incoming gradient.  A batch's loss is
the sum of its images' losses.  ``metrics.ssim`` is the torch statement of the same SSIM.

A fit's target does not change, so its window moments (G*t and G*t^2 per channel) can be computed
once (``target_moments``, the role of the reference's ``SSIMTargetStats``, metrics.py:93-128) and
passed to every step's loss: the forward then blurs three maps of
the image instead of five, with the same loss and gradient bit for bit.
"""
from __future__ import annotations

import ctypes

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _native


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def loss_plan(sizes):
    """Completed host field table of the loss for images of `sizes` [(H, W)] and (units,
    pixels, workspace bytes): the loss kernels work in units, strips 16 columns wide and up to
    128 rows tall."""
    nf = len(sizes)
    table = (_native.Field * nf)()
    for i, (h, w) in enumerate(sizes):
        table[i].height = int(h)
        table[i].width = int(w)
    units, px, wsb = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_size_t(0)
    _native.check(_native.load().ssplat_loss_plan(ctypes.addressof(table), nf, ctypes.byref(units),
                                                   ctypes.byref(px), ctypes.byref(wsb)),
                  "ssplat_loss_plan")
    return table, int(units.value), int(px.value), int(wsb.value)


def loss_workspace(pixels: int, tiles: int, device) -> torch.Tensor:
    n = ctypes.c_size_t(0)
    _native.check(_native.load().ssplat_loss_workspace_bytes(int(pixels), int(tiles),
                                                               ctypes.byref(n)),
                  "ssplat_loss_workspace_bytes")
    return torch.empty(max(int(n.value), 256), dtype=torch.uint8, device=device)


def _check_pair(img, target, what):
    if not (img.is_cuda and target.is_cuda):
        raise RuntimeError(f"{what} requires CUDA tensors")
    if img.dtype != torch.float32 or target.dtype != torch.float32:
        raise ValueError(f"{what}: image and target must be float32")
    if img.shape != target.shape or img.shape[-1] != 3:
        raise ValueError(f"{what}: image and target must have the same (..., 3) shape")


def _check_weight(w: float) -> float:
    w = float(w)
    if not 0.0 <= w <= 1.0:
        raise ValueError(f"ssim_weight must be in [0, 1], got {w}")
    return w


def target_moments(target: torch.Tensor) -> torch.Tensor:
    """Window moments of an (H, W, 3) target: a (2, H, W, 3) tensor holding G*t and G*t^2."""
    if not target.is_cuda or target.dtype != torch.float32:
        raise ValueError("target_moments: target must be a float32 CUDA tensor")
    if target.dim() != 3 or target.shape[-1] != 3:
        raise ValueError("target_moments: target must be (H, W, 3)")
    target = target.contiguous()
    H, W = int(target.shape[0]), int(target.shape[1])
    out = torch.empty(2, H, W, 3, dtype=torch.float32, device=target.device)
    _native.check(_native.load().ssplat_loss_target_moments(
        _ptr(target), H, W, _ptr(out), _stream(target.device)), "ssplat_loss_target_moments")
    return out


def batch_target_moments(batch, target_flat: torch.Tensor) -> torch.Tensor:
    """Window moments of every field's target in a concatenated ``(total_pixels, 3)`` buffer: a
    ``(2, total_pixels, 3)`` tensor holding G*t and G*t^2."""
    if not target_flat.is_cuda or target_flat.dtype != torch.float32:
        raise ValueError("batch_target_moments: targets must be a float32 CUDA tensor")
    if target_flat.shape != (batch.total_pixels, 3):
        raise ValueError("batch_target_moments: targets must be (total_pixels, 3)")
    target_flat = target_flat.contiguous()
    out = torch.empty(2, batch.total_pixels, 3, dtype=torch.float32, device=target_flat.device)
    _native.check(_native.load().ssplat_batch_loss_target_moments(
        _ptr(batch.loss_fields_dev), batch.num_fields, batch.loss_total_tiles, batch.total_pixels,
        _ptr(target_flat), _ptr(out), _stream(target_flat.device)),
        "ssplat_batch_loss_target_moments")
    return out


def _check_moments(moments, shape):
    if moments is None:
        return None
    if (not moments.is_cuda or moments.dtype != torch.float32
            or tuple(moments.shape) != (2,) + tuple(shape)):
        raise ValueError("target moments must be a float32 CUDA tensor of shape "
                         f"{(2,) + tuple(shape)}")
    return moments.contiguous()


class _L1SSIM(Function):
    """Forward: the loss (and the SSIM derivative maps kept in a workspace); backward: the image
    gradient scaled by the incoming gradient, read on the device."""

    @staticmethod
    def forward(ctx, img, target, ssim_weight, batch, moments):
        img = img.contiguous()
        target = target.contiguous()
        dev = img.device
        loss = torch.empty((), dtype=torch.float32, device=dev)
        lib = _native.load()
        if batch is None:
            H, W = int(img.shape[0]), int(img.shape[1])
            _, units, _, _ = loss_plan([(H, W)])
            ws = loss_workspace(H * W, units, dev)
            _native.check(lib.ssplat_l1_ssim_loss_forward(
                _ptr(img), _ptr(target), _ptr(moments), H, W, float(ssim_weight), _ptr(loss),
                _ptr(ws), ws.numel(), _stream(dev)), "ssplat_l1_ssim_loss_forward")
        else:
            ws = loss_workspace(batch.total_pixels, batch.loss_total_tiles, dev)
            _native.check(lib.ssplat_batch_l1_ssim_loss_forward(
                _ptr(batch.loss_fields_dev), batch.num_fields, batch.loss_total_tiles,
                batch.total_pixels, _ptr(img), _ptr(target), _ptr(moments), float(ssim_weight),
                _ptr(loss), _ptr(ws), ws.numel(), _stream(dev)),
                "ssplat_batch_l1_ssim_loss_forward")
        ctx.save_for_backward(img, target, ws)
        ctx.ssim_weight = float(ssim_weight)
        ctx.batch = batch
        return loss

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_loss):
        img, target, ws = ctx.saved_tensors
        dev = img.device
        grad = torch.empty_like(img)
        g = grad_loss.to(dtype=torch.float32).contiguous()
        lib = _native.load()
        b = ctx.batch
        if b is None:
            _native.check(lib.ssplat_l1_ssim_loss_backward(
                _ptr(img), _ptr(target), int(img.shape[0]), int(img.shape[1]), ctx.ssim_weight,
                _ptr(g), _ptr(grad), _ptr(ws), ws.numel(), _stream(dev)),
                "ssplat_l1_ssim_loss_backward")
        else:
            _native.check(lib.ssplat_batch_l1_ssim_loss_backward(
                _ptr(b.loss_fields_dev), b.num_fields, b.loss_total_tiles, b.total_pixels,
                _ptr(img),
                _ptr(target), ctx.ssim_weight, _ptr(g), _ptr(grad), _ptr(ws), ws.numel(),
                _stream(dev)), "ssplat_batch_l1_ssim_loss_backward")
        return grad, None, None, None, None


def l1_ssim_loss(img: torch.Tensor, target: torch.Tensor, ssim_weight: float = 0.2,
                 moments: torch.Tensor | None = None):
    """(1 - w) * L1 + w * (1 - SSIM) of an (H, W, 3) image against its target (differentiable in
    ``img``).  ``moments``: ``target_moments(target)``, for a target used in many calls."""
    _check_pair(img, target, "l1_ssim_loss")
    if img.dim() != 3:
        raise ValueError("l1_ssim_loss: image must be (H, W, 3)")
    return _L1SSIM.apply(img, target, _check_weight(ssim_weight), None,
                         _check_moments(moments, target.shape))


def batch_l1_ssim_loss(batch, img_flat: torch.Tensor, target_flat: torch.Tensor,
                       ssim_weight: float = 0.2, moments: torch.Tensor | None = None):
    """Sum over the fields of ``batch`` (a ``batch.Batch``) of each field's loss, for concatenated
    ``(total_pixels, 3)`` images.  ``moments``: ``batch_target_moments(batch, target_flat)``."""
    _check_pair(img_flat, target_flat, "batch_l1_ssim_loss")
    if img_flat.shape != (batch.total_pixels, 3):
        raise ValueError("batch_l1_ssim_loss: images must be (total_pixels, 3)")
    return _L1SSIM.apply(img_flat, target_flat, _check_weight(ssim_weight), batch,
                         _check_moments(moments, target_flat.shape))
