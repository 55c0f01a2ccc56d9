"""GPU parity of the fused Adam step (with the log-scale clamp) against torch.optim.Adam."""
from __future__ import annotations

import math

import pytest
import torch

from structsplat_amd.optim import FusedAdam

pytestmark = pytest.mark.gpu


def _params(device, seed=0):
    g = torch.Generator(device="cpu").manual_seed(seed)
    shapes = [(500, 2), (500, 2), (500,), (500, 3), (500,)]
    return [torch.randn(*s, generator=g).to(device) for s in shapes]


@pytest.mark.parametrize("capturable", [False, True])
def test_fused_adam_matches_torch_adam(cuda, capturable):
    print(f"case: adam capturable={capturable}", flush=True)
    lrs = [1e-2, 5e-3, 5e-3, 5e-3, 1e-2]
    ref = [p.clone().requires_grad_(True) for p in _params(cuda)]
    mine = [p.clone().requires_grad_(True) for p in _params(cuda)]
    opt_ref = torch.optim.Adam([{"params": [p], "lr": lr} for p, lr in zip(ref, lrs)],
                               capturable=capturable)
    opt = FusedAdam([{"params": [p], "lr": lr} for p, lr in zip(mine, lrs)])
    g = torch.Generator(device="cpu").manual_seed(1)
    for _ in range(6):
        grads = [torch.randn(p.shape, generator=g).to(cuda) for p in ref]
        for p, q, gr in zip(ref, mine, grads):
            p.grad = gr.clone()
            q.grad = gr.clone()
        opt_ref.step()
        opt.step()
    torch.cuda.synchronize()
    assert float(opt.step_count) == 6.0
    for p, q in zip(ref, mine):
        assert torch.allclose(p, q, rtol=1e-5, atol=1e-6), float((p - q).abs().max())


def test_fused_clamp_scalar_and_per_row(cuda):
    ls = torch.zeros(4, 2, device=cuda, requires_grad=True)
    other = torch.zeros(3, device=cuda, requires_grad=True)
    opt = FusedAdam([{"params": [ls], "lr": 10.0}, {"params": [other], "lr": 10.0}])
    hi = torch.tensor([0.5, 1.0, 100.0, 100.0], device=cuda)
    opt.set_clamp(ls, math.log(0.35), hi)
    opt.set_clamp(other, -0.25, 0.25)
    ls.grad = torch.tensor([[-1.0, -1.0], [-1.0, 1.0], [1.0, -1.0], [1.0, 1.0]], device=cuda)
    other.grad = torch.tensor([-1.0, 1.0, 0.0], device=cuda)
    opt.step()
    torch.cuda.synchronize()
    # Adam's first step moves each element by lr * sign(g) (up to eps): +10 or -10, then clamps.
    lo = math.log(0.35)
    assert torch.allclose(ls.detach(), torch.tensor([[0.5, 0.5], [1.0, lo], [lo, 10.0],
                                                     [lo, lo]], device=cuda))
    assert torch.allclose(other.detach(), torch.tensor([0.25, -0.25, 0.0], device=cuda))
