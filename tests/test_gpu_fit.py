"""GPU tests of the fit step: graph replays match eager steps, capacity overflow rolls back."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from structsplat_amd import fit as F
from structsplat_amd import synth
from structsplat_amd.gaussians import GaussianField
from structsplat_amd.render import render_hip_affine

pytestmark = pytest.mark.gpu


def _problem(device, H=96, W=128, n=600, seed=3):
    target = torch.tensor(synth.fit_target(seed, H, W), device=device)
    init = synth.fit_init_field(synth.fit_target(seed, H, W), n, seed)
    field = GaussianField.from_numpy(init["means"], init["scales"], init["angles"],
                                     init["colors"], opacities=init["opacity"], device=device)
    return field, target


def test_graph_replays_match_eager_steps(cuda):
    """Each replay of the captured step equals one eager step from the same state (checked for
    several replays, with eager work and synchronizations between them)."""
    print("case: graph vs eager step, 96x128, n=600", flush=True)
    field, target = _problem(cuda)
    cfg = F.FitConfig(iters=0)
    H, W = target.shape[:2]
    field.trainable()
    opt = F.make_optimizer(field, cfg)
    plan = F.RenderPlan(field.n, H, W, 4 * F.pair_count(field, H, W, cfg) + 4096, cuda)
    losses = {}

    def step():
        opt.zero_grad(set_to_none=False)
        img = F.planned_render(field, plan, cfg)
        loss = F.loss_fn(img, target, cfg.ssim_weight)
        loss.backward()
        opt.step()
        losses["last"] = loss.detach()

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            step()
    torch.cuda.current_stream().wait_stream(side)
    graph = F._capture(step)
    graph_loss = losses["last"]  # the captured step writes its loss into this buffer

    for k in range(4):
        snap = F._Snapshot(field, opt)
        step()  # eager
        torch.cuda.synchronize()
        eager_params = [p.detach().clone() for p in field.parameters()]
        eager_loss = float(losses["last"])
        snap.restore(field, opt)
        _ = torch.randn(1000, device=cuda).sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert abs(float(graph_loss) - eager_loss) <= 1e-6 * max(1.0, abs(eager_loss)), k
        for p, e in zip(field.parameters(), eager_params):
            assert torch.allclose(p.detach(), e, rtol=1e-5, atol=1e-6), k
    assert int(plan.status.item()) == 0


def test_plain_and_graphed_fits_agree(cuda):
    print("case: plain vs graphed fit, 40 steps", flush=True)
    a, target = _problem(cuda, seed=5)
    b, _ = _problem(cuda, seed=5)
    cfg = F.FitConfig(iters=40, window=10)
    ra = F.fit(a, target, cfg, graph=False)
    rb = F.fit(b, target, cfg, graph=True)
    print(ra, rb)
    assert abs(ra["psnr"] - rb["psnr"]) < 0.01
    assert ra["psnr"] > 10.0


def test_capacity_overflow_rolls_back_and_matches(cuda):
    """A pair buffer too small for the field is detected by the sticky status word; the window is
    rolled back and rerun with a larger buffer, giving the same fit as an ample buffer."""
    print("case: overflow rollback", flush=True)
    a, target = _problem(cuda, seed=7)
    b, _ = _problem(cuda, seed=7)
    ok = F.fit(a, target, F.FitConfig(iters=12, window=4), graph=False)
    tight = F.fit(b, target, F.FitConfig(iters=12, window=4, capacity_slack=0.5,
                                         capacity_floor=1), graph=False)
    assert tight["regrows"] >= 1
    assert abs(ok["psnr"] - tight["psnr"]) < 0.01
    assert np.isfinite(tight["loss"])


def test_tile32_fit_matches_tile16(cuda):
    """The fit at 32-pixel tiles renders the same images as at 16 (gradients differ only in
    atomic summation order), so both fits reach the same PSNR."""
    print("case: fit tile 16 vs 32", flush=True)
    a, target = _problem(cuda, seed=9)
    b, _ = _problem(cuda, seed=9)
    ra = F.fit(a, target, F.FitConfig(iters=30, window=10, tile_size=16), graph=True)
    rb = F.fit(b, target, F.FitConfig(iters=30, window=10, tile_size=32), graph=True)
    print(ra, rb)
    assert (ra["tile_size"], rb["tile_size"]) == (16, 32)
    assert abs(ra["psnr"] - rb["psnr"]) < 0.01


def test_autotuned_tile_size(cuda):
    """The tuner times every candidate, picks the fastest, and the fit uses its choice."""
    from structsplat_amd import tune

    print("case: tile autotuner", flush=True)
    field, target = _problem(cuda, seed=11)
    H, W = target.shape[:2]
    cfg = F.FitConfig(iters=4, window=2)
    best, timings = tune.tune_field(field, H, W, cfg)
    assert set(timings) == set(tune.CANDIDATE_TILE_SIZES)
    assert all(t > 0 for t in timings.values())
    assert best == min(timings, key=timings.get)
    assert tune.tune_field(field, H, W, cfg) == (best, timings)  # cached per layout and mode
    r = F.fit(field, target, cfg, graph=False)
    assert r["tile_size"] == best
    with pytest.raises(ValueError):
        F.resolve_tile_size(F.FitConfig(tile_size=24), lambda: 16)


def _affine_problem(device, seed):
    field, target = _problem(device, seed=seed)
    g = torch.Generator(device=device).manual_seed(seed)
    field.color_grads = 0.05 * torch.randn(field.n, 2, 3, device=device, generator=g)
    return field, target


def test_planned_affine_step_matches_render_hip_affine(cuda):
    """The fit's planned affine render gives the image and all gradients of render_hip_affine
    with the same producer outputs (same tile lists, same deterministic slot order)."""
    print("case: planned affine step vs render_hip_affine, 96x128, n=600", flush=True)
    field, target = _affine_problem(cuda, 11)
    cfg = F.FitConfig(iters=0, color_basis="affine")
    H, W = target.shape[:2]
    grads = []
    for planned in (True, False):
        field.trainable()
        for p in field.parameters():
            p.grad = None
        if planned:
            plan = F.RenderPlan(field.n, H, W, 2 * F.pair_count(field, H, W, cfg) + 4096, cuda,
                                affine=True)
            img = F.planned_render(field, plan, cfg)
        else:
            conics, radii, op = field.produce(cfg.sigma_cutoff)
            img = render_hip_affine(field.means, conics, field.colors, radii, H, W,
                                    field.color_grads, field.effective_scales(),
                                    field.rotations, opacities=op)
        F.loss_fn(img, target, cfg.ssim_weight).backward()
        torch.cuda.synchronize()
        grads.append([img.detach().clone()] + [p.grad.clone() for p in field.parameters()])
    assert int(plan.status.item()) == 0
    assert torch.equal(grads[0][0], grads[1][0])
    for a, b in zip(grads[0][1:], grads[1][1:]):
        assert torch.allclose(a, b, rtol=1e-5, atol=1e-7)
    assert float(grads[0][-1].abs().max()) > 0.0  # color_grads receive gradient


def test_affine_plain_and_graphed_fits_agree(cuda):
    """color_basis="affine" adds zero coefficients to a constant field; the plain and graphed
    fits agree and move the coefficients off zero."""
    print("case: affine plain vs graphed fit, 30 steps", flush=True)
    a, target = _problem(cuda, seed=13)
    b, _ = _problem(cuda, seed=13)
    cfg = F.FitConfig(iters=30, window=10, color_basis="affine", color_grad_l2=1e-3)
    ra = F.fit(a, target, cfg, graph=False)
    rb = F.fit(b, target, cfg, graph=True)
    print(ra, rb)
    assert ra["tile_size"] == 16 and a.color_grads is not None
    assert abs(ra["psnr"] - rb["psnr"]) < 0.01 and ra["psnr"] > 10.0
    assert float(a.color_grads.abs().max()) > 0.0 and ra["color_grad_l2_term"] > 0.0
    assert torch.allclose(a.color_grads, b.color_grads, rtol=1e-3, atol=1e-5)


def test_fit_caps_log_scales_at_scale_max(cuda):
    """The reference's step caps log-scales at log(clamp(scale_max, 1e-3)) after the
    [ln 0.35, ln max(H, W)] clamp (fit.py:3359-3375): capped rows never exceed their cap, and a
    cap below ln 0.35 wins over the lower bound."""
    print("case: fit with scale_max caps", flush=True)
    field, target = _problem(cuda, H=64, W=80, n=300, seed=9)
    s0 = field.scales().detach()
    cap = s0 * 1.05
    cap[::3] = 0.1  # below exp(ln 0.35): the cap wins over the lower clamp
    field.scale_max = cap.clone()
    cfg = F.FitConfig(iters=30, window=10, lr_scales=0.05)
    F.fit(field, target, cfg, graph=False)
    torch.cuda.synchronize()
    limit = torch.log(torch.clamp(cap, min=1e-3))
    ls = field.log_scales.detach()
    assert bool((ls <= limit).all())
    assert torch.equal(ls[::3], limit[::3])
    lo = torch.tensor(math.log(0.35), dtype=torch.float32)
    free = limit >= lo
    assert bool((ls[free] >= lo).all())
