"""Field layout and producer of the renderer inputs (the reference's ``GaussianField``).

A field stores means (N,2), log-scales (N,2), rotations (N,), colors (N,3) and optional opacity
logits (N,) (gaussians.py:15-29).  The renderer consumes what this class produces:
``conics()`` (inverse covariance entries (a, b, c)), ``radii()`` (int64 support half-widths) and
``opacity_values()`` (sigmoid of the logits).  The formulas follow gaussians.py:231-321 so the
integer radii, and therefore the tile lists, are identical to the reference's.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable


class GaussianField:
    """The reference's field (gaussians.py:15-29), with its constructor's argument order."""

    def __init__(self, means, log_scales, rotations, colors, opacities=None, scale_max=None,
                 color_grads=None, background_mask=None, filter_variance=None):
        self.means = means            # (N, 2)
        self.log_scales = log_scales  # (N, 2)
        self.rotations = rotations    # (N,)
        self.colors = colors          # base colors (N, 3)
        self.opacities = opacities    # optional logits (N,)
        self.scale_max = scale_max    # optional per-Gaussian scale cap (N, 2), kept as data
        self.color_grads = color_grads  # optional affine color coefficients (N, 2, 3)
        self.background_mask = background_mask  # optional bool rows, kept as data
        # Optional non-trainable isotropic covariance variance (px^2) per row.
        self.filter_variance = filter_variance

    @classmethod
    def from_numpy(cls, means, scales, angles, colors, opacities=None, scale_max=None,
                   device="cpu", dtype=torch.float32, color_grads=None,
                   background_mask=None, filter_variance=None):
        """Build a field from NumPy arrays; scales are clipped to >= 1e-3 (gaussians.py:31-55)."""
        def t(a):
            return torch.as_tensor(np.asarray(a), device=device, dtype=dtype).clone()

        scales = np.clip(np.asarray(scales), 1e-3, None)
        opacity_t = None if opacities is None else t(opacities).reshape(-1)
        mask = (None if background_mask is None else torch.as_tensor(
            np.asarray(background_mask), device=device, dtype=torch.bool).reshape(-1).clone())
        return cls(t(means), torch.log(t(scales)), t(angles).reshape(-1), t(colors), opacity_t,
                   None if scale_max is None else t(scale_max),
                   None if color_grads is None else t(color_grads), mask,
                   None if filter_variance is None else t(filter_variance).reshape(-1))

    @property
    def n(self) -> int:
        return self.means.shape[0]

    def trainable(self) -> "GaussianField":
        for p in self.parameters():
            p.requires_grad_(True)
        return self

    def detached(self) -> "GaussianField":
        def d(x):
            return None if x is None else x.detach().clone()

        return GaussianField(d(self.means), d(self.log_scales), d(self.rotations),
                             d(self.colors), d(self.opacities), d(self.scale_max),
                             d(self.color_grads), d(self.background_mask),
                             d(self.filter_variance))

    def parameters(self) -> list[torch.Tensor]:
        params = [self.means, self.log_scales, self.rotations, self.colors]
        if self.opacities is not None:
            params.append(self.opacities)
        if self.color_grads is not None:
            params.append(self.color_grads)
        return params

    def parameter_groups(self, lr_means, lr_scales, lr_rot, lr_color, lr_opacity=1e-2):
        """Adam parameter groups (gaussians.py:203-214)."""
        groups = [
            {"params": [self.means], "lr": lr_means},
            {"params": [self.log_scales], "lr": lr_scales},
            {"params": [self.rotations], "lr": lr_rot},
            {"params": [self.colors], "lr": lr_color},
        ]
        if self.opacities is not None:
            groups.append({"params": [self.opacities], "lr": lr_opacity})
        if self.color_grads is not None:
            groups.append({"params": [self.color_grads], "lr": lr_color})
        return groups

    def with_affine_colors(self) -> "GaussianField":
        """The field with zero affine color coefficients added (gaussians.py:216-229)."""
        if self.color_grads is not None:
            return self
        return GaussianField(self.means, self.log_scales, self.rotations, self.colors,
                             self.opacities, self.scale_max,
                             torch.zeros(self.n, 2, 3, device=self.colors.device,
                                         dtype=self.colors.dtype),
                             self.background_mask, self.filter_variance)

    def effective_scales(self, dilation: float = 0.0):
        """Scales after the isotropic covariance filter and anti-alias dilation
        (gaussians.py:234-246)."""
        if dilation < 0.0:
            raise ValueError(f"dilation must be >= 0, got {dilation}")
        scales = self.scales()
        if self.filter_variance is None and dilation == 0.0:
            return scales
        variance = scales.square() + float(dilation)
        if self.filter_variance is not None:
            variance = variance + self.filter_variance.to(device=scales.device,
                                                          dtype=scales.dtype).reshape(-1, 1)
        return torch.sqrt(variance.clamp_min(1e-12))

    def effective_log_scales(self, dilation: float = 0.0):
        """log of ``effective_scales`` (gaussians.py:248-252)."""
        if self.filter_variance is None and dilation == 0.0:
            return self.log_scales
        return torch.log(self.effective_scales(dilation))

    def materialize_covariance_filter(self) -> "GaussianField":
        """The filter variance folded into the scales (gaussians.py:254-268)."""
        if self.filter_variance is None:
            return self
        return GaussianField(self.means, self.effective_log_scales(), self.rotations, self.colors,
                             self.opacities, self.scale_max, self.color_grads,
                             self.background_mask, None)

    def scales(self):
        return torch.exp(self.log_scales)

    def opacity_values(self):
        """sigmoid(logits) or None (= opacity 1) (gaussians.py:270-273)."""
        if self.opacities is None:
            return None
        return torch.sigmoid(self.opacities)

    def conics(self, dilation: float = 0.0):
        """(a, b, c) of inv(Sigma) = [[a, b], [b, c]] with Sigma = R diag(sx^2, sy^2) R^T + d I
        (gaussians.py:275-300)."""
        if dilation < 0.0:
            raise ValueError(f"dilation must be >= 0, got {dilation}")
        s = self.scales()
        sx2 = s[:, 0] ** 2
        sy2 = s[:, 1] ** 2
        if self.filter_variance is not None:
            v = self.filter_variance.to(device=s.device, dtype=s.dtype).reshape(-1)
            sx2 = sx2 + v
            sy2 = sy2 + v
        inv_sx2 = 1.0 / (sx2 + dilation)
        inv_sy2 = 1.0 / (sy2 + dilation)
        c = torch.cos(self.rotations)
        sn = torch.sin(self.rotations)
        a = c * c * inv_sx2 + sn * sn * inv_sy2
        b = c * sn * (inv_sx2 - inv_sy2)
        cc = sn * sn * inv_sx2 + c * c * inv_sy2
        return torch.stack([a, b, cc], dim=1)

    def radii(self, sigma_cutoff: float, dilation: float = 0.0):
        """Per-axis support half-widths ceil(sigma_cutoff * sqrt(Sigma_xx)), ... clamped >= 1,
        int64, detached (gaussians.py:302-321)."""
        if dilation < 0.0:
            raise ValueError(f"dilation must be >= 0, got {dilation}")
        with torch.no_grad():
            s2 = self.scales() ** 2
            if self.filter_variance is not None:
                s2 = s2 + self.filter_variance.to(device=s2.device, dtype=s2.dtype).reshape(-1, 1)
            s2 = s2 + dilation
            c = torch.cos(self.rotations)
            sn = torch.sin(self.rotations)
            var_x = c * c * s2[:, 0] + sn * sn * s2[:, 1]
            var_y = sn * sn * s2[:, 0] + c * c * s2[:, 1]
            r = sigma_cutoff * torch.sqrt(torch.stack([var_x, var_y], dim=1))
            return torch.clamp(r.ceil().long(), min=1)

    def produce(self, sigma_cutoff: float, dilation: float = 0.0):
        """``(conics(dilation), radii(sigma_cutoff, dilation), opacity_values())`` of a CUDA field in
        one HIP launch (and one for the backward of the conics and opacities).  The radii equal
        ``radii()``'s bit for bit; the conics and opacities agree with the torch expressions to
        float32 rounding.  Opacities are ``None`` when the field has none."""
        if dilation < 0.0:
            raise ValueError(f"dilation must be >= 0, got {dilation}")
        if not self.means.is_cuda:
            raise RuntimeError("GaussianField.produce requires CUDA tensors")
        logits = self.opacities if self.opacities is not None else self.log_scales.new_empty(0)
        fv = self.filter_variance
        if fv is not None:
            fv = fv.to(device=self.log_scales.device, dtype=torch.float32).reshape(-1)
        conics, radii, op = _Produce.apply(self.log_scales, self.rotations, logits, fv,
                                           float(sigma_cutoff), float(dilation))
        return conics, radii, (op if self.opacities is not None else None)

    @torch.no_grad()
    def save(self, path: str) -> None:
        data = {
            "means": self.means.detach().cpu().numpy(),
            "log_scales": self.log_scales.detach().cpu().numpy(),
            "rotations": self.rotations.detach().cpu().numpy(),
            "colors": self.colors.detach().cpu().numpy(),
        }
        for name in ("opacities", "scale_max", "color_grads", "background_mask",
                     "filter_variance"):
            value = getattr(self, name)
            if value is not None:
                data[name] = value.detach().cpu().numpy()
        np.savez(path, **data)

    @classmethod
    def load(cls, path: str, device="cpu") -> "GaussianField":
        z = np.load(path)

        def t(k):
            return torch.as_tensor(z[k], device=device, dtype=torch.float32)

        def opt(k):
            return t(k) if k in z.files else None

        mask = (torch.as_tensor(z["background_mask"], device=device, dtype=torch.bool)
                if "background_mask" in z.files else None)
        return cls(t("means"), t("log_scales"), t("rotations"), t("colors"), opt("opacities"),
                   opt("scale_max"), opt("color_grads"), mask, opt("filter_variance"))


def _ptr(t):
    return None if t is None or t.numel() == 0 else ctypes.c_void_p(t.data_ptr())


class _Produce(Function):
    """Autograd node of ssplat_produce_field (conics and opacities differentiable, radii not)."""

    @staticmethod
    def forward(ctx, log_scales, rotations, logits, filter_variance, sigma_cutoff, dilation):
        from . import _native

        log_scales = log_scales.detach().contiguous()
        rotations = rotations.detach().contiguous()
        logits = logits.detach().contiguous()
        n = log_scales.shape[0]
        dev = log_scales.device
        conics = torch.empty(n, 3, dtype=torch.float32, device=dev)
        radii = torch.empty(n, 2, dtype=torch.int64, device=dev)
        op = torch.empty(n if logits.numel() else 0, dtype=torch.float32, device=dev)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _native.check(_native.load().ssplat_produce_field(
            _ptr(log_scales), _ptr(rotations), _ptr(logits), _ptr(filter_variance), n,
            dilation, sigma_cutoff, _ptr(conics), _ptr(radii), _ptr(op), stream),
            "ssplat_produce_field")
        ctx.mark_non_differentiable(radii)
        # Unused output gradients arrive as None rather than as zero-filled tensors (the radii's
        # would cost a fill kernel per step).
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(log_scales, rotations, op, filter_variance)
        ctx.dilation = dilation
        return conics, radii, op

    @staticmethod
    @once_differentiable
    def backward(ctx, g_conics, g_radii, g_op):
        from . import _native

        log_scales, rotations, op, fv = ctx.saved_tensors
        n = log_scales.shape[0]
        g_ls = torch.empty_like(log_scales)
        g_rot = torch.empty_like(rotations)
        has_op = op.numel() > 0
        g_logits = torch.empty_like(op) if has_op else None
        if g_conics is None:
            g_conics = torch.zeros(n, 3, dtype=torch.float32, device=log_scales.device)
        # Contiguous copies are bound to names so they stay alive until the launch is enqueued
        # (a temporary freed earlier could be reused by the next copy on the same stream).
        gq = g_conics.contiguous()
        go = g_op.contiguous() if (has_op and g_op is not None) else None
        stream = ctypes.c_void_p(torch.cuda.current_stream(log_scales.device).cuda_stream)
        _native.check(_native.load().ssplat_produce_field_backward(
            _ptr(log_scales), _ptr(rotations), _ptr(op) if has_op else None, _ptr(fv), n,
            ctx.dilation, _ptr(gq), _ptr(go), _ptr(g_ls), _ptr(g_rot), _ptr(g_logits), stream),
            "ssplat_produce_field_backward")
        return g_ls, g_rot, g_logits, None, None, None
