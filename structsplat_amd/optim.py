"""The fit step's optimizer: Adam over a field's parameter groups in one HIP launch.

``FusedAdam`` keeps ``torch.optim.Adam``'s update (the reference's optimizer, fit.py:672-691, with
its default betas and eps) and its ``param_groups`` / ``zero_grad`` / ``step`` surface for the fit
loops.  Its moment buffers are allocated up front and its step count lives on the device, so a
step can be captured in a HIP graph; ``set_clamp`` fuses the log-scale clamp of the reference's
step (fit.py:3033, :3359-3375) into the same launch.
"""
from __future__ import annotations

import ctypes

import torch

from . import _native


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class FusedAdam:
    def __init__(self, groups, betas=(0.9, 0.999), eps: float = 1e-8):
        self.param_groups = []
        for g in groups:
            params = list(g["params"])
            for p in params:
                if p.dtype != torch.float32 or not p.is_cuda or not p.is_contiguous():
                    raise ValueError("FusedAdam needs contiguous float32 CUDA parameters")
            self.param_groups.append({"params": params, "lr": float(g["lr"])})
        if sum(len(g["params"]) for g in self.param_groups) > 8:
            raise ValueError("FusedAdam handles at most 8 parameter tensors")
        self.betas = (float(betas[0]), float(betas[1]))
        self.eps = float(eps)
        dev = self.param_groups[0]["params"][0].device
        self.step_count = torch.zeros((), dtype=torch.float32, device=dev)
        self.state = {p: {"exp_avg": torch.zeros_like(p), "exp_avg_sq": torch.zeros_like(p)}
                      for g in self.param_groups for p in g["params"]}
        self._clamp = {}

    def set_clamp(self, param: torch.Tensor, lo: float, hi=None) -> None:
        """Clamp `param` to [lo, hi] after every step; `hi` is a number or a float32 tensor
        holding one upper bound per row of `param` (rows = the leading dimension) or one per
        element."""
        if param not in self.state:
            raise ValueError("set_clamp: not a parameter of this optimizer")
        if torch.is_tensor(hi):
            rows = hi.reshape(-1)
            if rows.dtype != torch.float32 or rows.numel() not in (param.shape[0], param.numel()):
                raise ValueError("set_clamp: need one float32 bound per row or per element")
            width = (1 if rows.numel() == param.numel()
                     else param.numel() // max(param.shape[0], 1))
            self._clamp[param] = (float(lo), 0.0, rows.contiguous(), width)
        else:
            self._clamp[param] = (float(lo), float(hi), None, 1)

    def zero_grad(self, set_to_none: bool = False) -> None:
        for g in self.param_groups:
            for p in g["params"]:
                if p.grad is None:
                    continue
                if set_to_none:
                    p.grad = None
                else:
                    p.grad.zero_()

    def state_tensors(self):
        """Every state tensor (moments and the step count), for snapshots."""
        out = [self.step_count]
        for g in self.param_groups:
            for p in g["params"]:
                out += [self.state[p]["exp_avg"], self.state[p]["exp_avg_sq"]]
        return out

    @torch.no_grad()
    def step(self) -> None:
        groups = []
        for g in self.param_groups:
            for p in g["params"]:
                if p.grad is None:
                    continue
                st = self.state[p]
                lo, hi, rows, width = self._clamp.get(p, (0.0, 0.0, None, 1))
                groups.append(_native.AdamGroup(
                    p.data_ptr(), p.grad.data_ptr(), st["exp_avg"].data_ptr(),
                    st["exp_avg_sq"].data_ptr(), p.numel(), g["lr"], int(p in self._clamp),
                    lo, hi, None if rows is None else rows.data_ptr(), width))
        if not groups:
            return
        arr = (_native.AdamGroup * len(groups))(*groups)
        dev = self.step_count.device
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _native.check(_native.load().ssplat_adam_step(arr, len(groups), _ptr(self.step_count),
                                                      self.betas[0], self.betas[1], self.eps,
                                                      stream), "ssplat_adam_step")
