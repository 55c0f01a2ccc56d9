"""Fit step around the HIP renderer: render -> L1 + SSIM loss -> backward -> Adam -> log-scale clamp.

This is the reference's hot loop reduced to the renderer's neighbours (fit.py:3182-3375):
``_render`` (fit.py:1043-1056), ``_weighted_pixel_loss`` with ``pixel_loss="l1"`` (fit.py:800-823),
``loss = (1 - w) L1 + w (1 - SSIM)`` (fit.py:3219-3244), ``opt.step()`` with the parameter groups
of gaussians.py:203-214, and the log-scale clamp to [ln 0.35, ln max(H, W)] (fit.py:3033,
:3359-3375).  Densification, pruning and the other research events of the reference fitter are
not part of this path.

Two drivers run the same arithmetic:

* ``fit(..., graph=False)`` - the plain per-image loop (one eager step per iteration).
* ``fit(..., graph=True)`` - the whole step captured once as a HIP graph and replayed.

Both bin without host synchronization into a capacity-sized pair buffer (``RenderPlan``).  The
binning status word is sticky; it is read once per window of steps together with the last
binning's pair count (``tile_offsets[T]``, on the device), and a window that overflowed is
rolled back (parameters and Adam state restored) and replayed with a larger capacity, so the
result equals a run with unlimited capacity; a pair count near the capacity regrows the buffer
before it overflows.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import render as HR
from .gaussians import GaussianField
from .loss import l1_ssim_loss
from .optim import FusedAdam
from .tune import tune_field
from .metrics import psnr_tensor


@dataclass
class FitConfig:
    iters: int = 1000
    lr_means: float = 1e-2
    lr_scales: float = 5e-3
    lr_rot: float = 5e-3
    lr_color: float = 5e-3
    lr_opacity: float = 5e-3
    ssim_weight: float = 0.2       # loss = (1 - w) * L1 + w * (1 - SSIM)
    recompute_normalizer: bool = True  # backward recomputes den instead of storing it
    sigma_cutoff: float = 3.0
    support_fade: bool | float = False  # bool, or a fade strength in [0, 1]
    normalization_eps: float = 1e-8
    normalize: bool = True
    window: int = 25               # steps between capacity checks
    capacity_slack: float = 1.5    # pair-buffer capacity / current pair count
    capacity_floor: int = 4096     # minimum pair-buffer capacity
    tile_size: int | str = "auto"  # render tile edge (16 or 32), or "auto": the autotuner's choice
    color_basis: str = "constant"  # "constant", or "affine": per-Gaussian color gradients
    color_grad_l2: float = 1e-4    # weight of mean(color_grads^2) in the affine loss


class RenderPlan:
    """Preallocated pair buffer, offsets, sticky status, workspace and gradient slots for one
    field shape.  The slot buffer has the pair buffer's capacity (the slot count equals the pair
    count without ellipse cull); an overflow of either sets the status word."""

    def __init__(self, n: int, H: int, W: int, capacity: int, device, tile_size: int = 16,
                 recompute_normalizer: bool = True, affine: bool = False):
        if affine and int(tile_size) != HR.DEFAULT_TILE_SIZE:
            raise ValueError("affine-color plans use 16-pixel tiles")
        # tile_size: the edge, optionally with the TILE_PERSISTENT schedule flag.
        self.n, self.H, self.W, self.tile_size = int(n), int(H), int(W), int(tile_size)
        self.affine = bool(affine)
        # The affine backward always recomputes the normalizer.
        self.recompute_normalizer = bool(recompute_normalizer) or self.affine
        self.capacity = max(int(capacity), 1)
        T = HR.num_tiles(H, W, tile_size)
        self.gids = torch.empty(self.capacity, dtype=torch.int32, device=device)
        self.offsets = torch.empty(T + 1, dtype=torch.int32, device=device)
        self.status = torch.zeros(1, dtype=torch.int32, device=device)
        if self.affine:
            self.ws = HR.affine_workspace(n, H, W, device)
            self.slots = HR.affine_grad_slots(self.capacity, device)
        else:
            self.ws = HR.workspace(n, H, W, tile_size, device)
            self.slots = HR.grad_slots(self.capacity, device)
        self.den = (None if self.recompute_normalizer
                    else torch.empty(H, W, dtype=torch.float32, device=device))

    def buffers(self):
        return self.gids, self.offsets, self.status, self.ws


class _PlannedRender(Function):
    """Autograd node of one planned (synchronization-free) render; same gradients as
    ``render._ExactRenderHip``."""

    @staticmethod
    def forward(ctx, means, conics, colors, radii, opacities, plan, normalize, eps, fade,
                sigma_cutoff):
        means = means.contiguous()
        conics = conics.contiguous()
        colors = colors.contiguous()
        radii = radii.contiguous()
        opacities = opacities.contiguous()
        cull = HR._fade_alpha(fade) == 1.0
        gids, offsets, _ = HR.build_tile_index_capacity(
            means, conics, radii, plan.H, plan.W, plan.capacity, plan.tile_size, cull,
            sigma_cutoff, out=plan.buffers())
        out, den = HR.forward_tiled(means, conics, colors, radii, opacities, gids, offsets,
                                    plan.H, plan.W, normalize, eps, plan.tile_size, fade,
                                    sigma_cutoff, den=plan.den, ws=plan.ws)
        ctx.save_for_backward(means, conics, colors, radii, opacities, den, out)
        ctx.plan = plan
        ctx.args = (normalize, eps, fade, sigma_cutoff)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_out):
        means, conics, colors, radii, opacities, den, out = ctx.saved_tensors
        plan = ctx.plan
        normalize, eps, fade, sigma_cutoff = ctx.args
        g_m, g_q, g_c, g_o = HR.backward_tiled(
            grad_out.contiguous(), means, conics, colors, radii, opacities, den, out, plan.gids,
            plan.offsets, plan.H, plan.W, normalize, eps, plan.tile_size, fade, sigma_cutoff,
            ws=plan.ws, slots=plan.slots, slot_capacity=plan.capacity, status=plan.status,
            after_forward=True)
        if opacities.numel() == 0:
            g_o = None
        return g_m, g_q, g_c, None, g_o, None, None, None, None, None


class _PlannedAffineRender(Function):
    """Planned render with affine colors; same gradients as ``render._AffineRenderHip``."""

    @staticmethod
    def forward(ctx, means, conics, colors, radii, opacities, color_grads, scales, rotations,
                plan, normalize, eps, fade, sigma_cutoff):
        means, conics, colors = means.contiguous(), conics.contiguous(), colors.contiguous()
        radii, opacities = radii.contiguous(), opacities.contiguous()
        color_grads = color_grads.contiguous()
        scales, rotations = scales.contiguous(), rotations.contiguous()
        cull = HR._fade_alpha(fade) == 1.0
        gids, offsets, _ = HR.build_tile_index_capacity(
            means, conics, radii, plan.H, plan.W, plan.capacity, plan.tile_size, cull,
            sigma_cutoff, out=plan.buffers())
        out, _ = HR.forward_tiled_affine(means, conics, colors, radii, opacities, color_grads,
                                         scales, rotations, gids, offsets, plan.H, plan.W,
                                         normalize, eps, fade, sigma_cutoff, ws=plan.ws)
        ctx.save_for_backward(means, conics, colors, radii, opacities, color_grads, scales,
                              rotations, out)
        ctx.plan = plan
        ctx.args = (normalize, eps, fade, sigma_cutoff)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_out):
        means, conics, colors, radii, opacities, color_grads, scales, rotations, out = \
            ctx.saved_tensors
        plan = ctx.plan
        normalize, eps, fade, sigma_cutoff = ctx.args
        g = HR.backward_tiled_affine(
            grad_out.contiguous(), means, conics, colors, radii, opacities, color_grads, scales,
            rotations, None, out if normalize else None, plan.gids, plan.offsets, plan.H, plan.W,
            normalize, eps, fade, sigma_cutoff, plan.capacity, ws=plan.ws, slots=plan.slots,
            status=plan.status)
        g_m, g_q, g_c, g_o, g_cg, g_s, g_r = g
        if opacities.numel() == 0:
            g_o = None
        return g_m, g_q, g_c, None, g_o, g_cg, g_s, g_r, None, None, None, None, None


def planned_render(field: GaussianField, plan: RenderPlan, cfg: FitConfig) -> torch.Tensor:
    """Renders the field with the plan's buffers: conics, radii and opacities from the fused
    producer (gaussians.py:270-321), then binning, forward and (in backward) the gradients.
    An affine plan renders ``field.color_grads`` in each Gaussian's frame of
    ``effective_scales()`` and ``rotations`` (fit.py:1043-1056)."""
    conics, radii, op = field.produce(cfg.sigma_cutoff)
    if op is None:
        op = field.means.new_empty(0)
    args = (plan, cfg.normalize, float(cfg.normalization_eps), HR._fade_alpha(cfg.support_fade),
            float(cfg.sigma_cutoff))
    if plan.affine:
        return _PlannedAffineRender.apply(field.means, conics, field.colors, radii, op,
                                          field.color_grads, field.effective_scales(),
                                          field.rotations, *args)
    return _PlannedRender.apply(field.means, conics, field.colors, radii, op, *args)


def add_color_grad_l2(field: GaussianField, cfg: FitConfig) -> None:
    """Adds the gradient of ``color_grad_l2 * mean(color_grads^2)`` (fit.py:3256-3259) to
    ``color_grads.grad`` in place: one elementwise kernel, so a captured step holds no
    reduction."""
    cg = field.color_grads
    if cg is None or cfg.color_grad_l2 <= 0.0 or cg.grad is None or cg.numel() == 0:
        return
    with torch.no_grad():
        cg.grad.add_(cg, alpha=2.0 * float(cfg.color_grad_l2) / cg.numel())


def color_grad_l2_term(field: GaussianField, cfg: FitConfig) -> float:
    """The value ``color_grad_l2 * mean(color_grads^2)`` (0 without affine colors)."""
    cg = field.color_grads
    if cg is None or cfg.color_grad_l2 <= 0.0 or cg.numel() == 0:
        return 0.0
    with torch.no_grad():
        return float(cfg.color_grad_l2 * (cg * cg).mean())


def raw_weight_map_field(field: GaussianField, cfg: FitConfig, H: int, W: int,
                         support_fade_alpha: float | None = None,
                         detach_opacities: bool = False) -> torch.Tensor:
    """Differentiable (H*W,) raw opacity-weighted weight sum sum_i o_i G_i(p) of a field
    (``_raw_weight_map_field``, fit.py:338-367): the coverage behind the reference's mask and
    coverage-matching penalties.  ``support_fade_alpha`` overrides ``cfg.support_fade``;
    ``detach_opacities`` cuts only the opacity gradient path."""
    dilation = float(getattr(cfg, "aa_dilation", 0.0))
    fade = (HR._fade_alpha(cfg.support_fade) if support_fade_alpha is None
            else max(0.0, min(1.0, float(support_fade_alpha))))
    return HR.raw_weight_map(field.means, field.conics(dilation),
                             field.radii(cfg.sigma_cutoff, dilation), H, W,
                             opacities=field.opacity_values(), support_fade=fade,
                             sigma_cutoff=cfg.sigma_cutoff, detach_opacities=detach_opacities)


def loss_fn(img: torch.Tensor, target: torch.Tensor, w: float,
            moments: torch.Tensor | None = None) -> torch.Tensor:
    """(1 - w) * L1 + w * (1 - SSIM) (fit.py:3219-3244 with pixel_loss="l1"), computed with its
    gradient by the fused HIP loss kernels; ``moments``: the target's window moments
    (``loss.target_moments``), or None to blur them inline (the same bits)."""
    return l1_ssim_loss(img, target, w, moments)


def pair_count(field: GaussianField, H: int, W: int, cfg: FitConfig, tile_size: int = 16) -> int:
    """Exact (tile, Gaussian) pair count of the current field (one host readback)."""
    with torch.no_grad():
        _, offsets = HR.build_tile_index(field.means, field.conics(), field.radii(cfg.sigma_cutoff),
                                         H, W, tile_size,
                                         HR._fade_alpha(cfg.support_fade) == 1.0,
                                         cfg.sigma_cutoff)
    return int(offsets[-1].item())


def make_optimizer(field: GaussianField, cfg: FitConfig) -> FusedAdam:
    """Adam over the field's parameter groups (gaussians.py:203-214) in one HIP launch."""
    groups = field.parameter_groups(cfg.lr_means, cfg.lr_scales, cfg.lr_rot, cfg.lr_color,
                                    cfg.lr_opacity)
    return FusedAdam(groups)


class _Snapshot:
    """Parameters and optimizer state at the start of a window (rolled back on overflow)."""

    def __init__(self, field: GaussianField, opt: FusedAdam):
        self.params = [p.detach().clone() for p in field.parameters()]
        self.state = [t.clone() for t in opt.state_tensors()]

    def restore(self, field: GaussianField, opt: FusedAdam) -> None:
        with torch.no_grad():
            for p, saved in zip(field.parameters(), self.params):
                p.copy_(saved)
            for t, saved in zip(opt.state_tensors(), self.state):
                t.copy_(saved)


def log_scale_upper(field: GaussianField, H: int, W: int):
    """Upper bound of the log-scale clamp after each step (fit.py:3033, :3359-3375):
    ``ln max(H, W)``, and with a per-Gaussian ``scale_max`` also ``ln clamp(scale_max, 1e-3)``
    element by element.  Returns a number, or a float32 tensor of one bound per element.
    ``min(clamp(x, lo, hi), cap)`` (the reference's order) equals ``clamp(x, lo, min(hi, cap))``
    because ``lo < hi``."""
    hi = math.log(max(int(H), int(W)))
    if field.scale_max is None:
        return hi
    cap = torch.log(torch.clamp(field.scale_max.detach().to(device=field.log_scales.device,
                                                            dtype=torch.float32), min=1e-3))
    return torch.clamp(cap.reshape(field.log_scales.shape), max=hi).reshape(-1).contiguous()


def fit(field: GaussianField, target: torch.Tensor, cfg: FitConfig, graph: bool = False,
        verbose: bool = False) -> dict:
    """Fit `field` to `target` (H, W, 3) for cfg.iters Adam steps.

    With ``cfg.color_basis == "affine"`` (or a field that already has ``color_grads``) the field
    renders affine colors and its loss adds ``color_grad_l2 * mean(color_grads^2)``
    (fit.py:3256-3259); zero coefficients are added to the field in place when it has none.

    Returns the final PSNR, the last step's image loss, the pair-buffer regrows and capacity, the
    tile size, ``seconds``: the wall time of the step loop (window checks and regrows included;
    the tile tuning and first plan before it are not), and ``color_grad_l2_term``: the affine
    regularizer at the final parameters."""
    H, W = target.shape[:2]
    dev = target.device
    if cfg.color_basis not in ("constant", "affine"):
        raise ValueError(f"color_basis must be constant or affine, got {cfg.color_basis!r}")
    if cfg.color_basis == "affine" and field.color_grads is None:
        # _ensure_color_basis (fit.py:1219-1222), in place: zero coefficients.
        field.color_grads = torch.zeros(field.n, 2, 3, device=field.colors.device,
                                        dtype=field.colors.dtype)
    affine = field.color_grads is not None
    field.trainable()
    opt = make_optimizer(field, cfg)
    opt.set_clamp(field.log_scales, math.log(0.35), log_scale_upper(field, H, W))
    if affine:
        if cfg.tile_size not in ("auto", HR.DEFAULT_TILE_SIZE):
            raise ValueError("affine colors render with 16-pixel tiles")
        ts = HR.DEFAULT_TILE_SIZE
    else:
        ts = resolve_tile_size(cfg, lambda: tune_field(field, H, W, cfg)[0])

    def new_plan(capacity):
        return RenderPlan(field.n, H, W, capacity, dev, tile_size=ts,
                          recompute_normalizer=cfg.recompute_normalizer, affine=affine)

    plan = new_plan(_capacity(pair_count(field, H, W, cfg, ts), cfg))
    state = {"loss": None}
    target = target.contiguous()
    # The loss blurs the target's window moments inline (five maps a step) rather than reading
    # them precomputed (loss.target_moments): its stats pass is bound by memory traffic, and the
    # precomputed moments are 24 more bytes a pixel to read each step.  Both give the same bits.
    moments = None

    def step():
        # Gradients are assigned, not accumulated: no fill and add kernels per parameter.
        opt.zero_grad(set_to_none=True)
        img = planned_render(field, plan, cfg)
        loss = loss_fn(img, target, cfg.ssim_weight, moments)
        loss.backward(_seed(state, loss))
        add_color_grad_l2(field, cfg)
        opt.step()  # Adam and the log-scale clamp
        state["loss"] = loss.detach()

    done = 0
    graph_obj = None
    regrows = 0
    replay = {"steps": 0, "seconds": 0.0}  # graph replays (or eager steps) alone
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    while done < cfg.iters:
        n = min(cfg.window, cfg.iters - done)
        snap = _Snapshot(field, opt)
        plan.status.zero_()
        if graph and done >= 2:
            if graph_obj is None:
                graph_obj = _capture(step)
            _timed_steps(replay, n, graph_obj.replay, dev)
        elif graph:
            # The graphed driver's first two steps run eagerly on a side stream (torch's capture
            # recipe): they create Adam's state and the gradient buffers the graph reuses.
            n = min(n, 2 - done)
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(side):
                for _ in range(n):
                    step()
            torch.cuda.current_stream(dev).wait_stream(side)
        else:
            _timed_steps(replay, n, step, dev)
        # One readback per window: the sticky status and the last binning's pair count
        # (offsets[T], written on the device by every step's binning).
        status, p = window_readback(plan)
        if status != 0:
            snap.restore(field, opt)
            plan = new_plan(max(2 * plan.capacity, _capacity(p, cfg)))
            graph_obj = None
            regrows += 1
            continue
        done += n
        # Keep headroom: regrow (and re-capture) when the pair count approaches capacity.
        if p * 1.15 > plan.capacity:
            plan = new_plan(_headroom(p, cfg))
            graph_obj = None
            regrows += 1
        if verbose:
            print(f"step {done}: loss {float(state['loss']):.6f} pairs {p}", flush=True)
    torch.cuda.synchronize(dev)
    seconds = time.perf_counter() - t0
    with torch.no_grad():
        kw = dict(opacities=field.opacity_values(), normalize=cfg.normalize,
                  eps=cfg.normalization_eps, support_fade=cfg.support_fade,
                  sigma_cutoff=cfg.sigma_cutoff)
        if affine:
            img = HR.render_hip_affine(field.means, field.conics(), field.colors,
                                       field.radii(cfg.sigma_cutoff), H, W, field.color_grads,
                                       field.effective_scales(), field.rotations, **kw)
        else:
            img = HR.render_hip_exact(field.means, field.conics(), field.colors,
                                      field.radii(cfg.sigma_cutoff), H, W, **kw)
        final_psnr = float(psnr_tensor(img, target))
    return {"psnr": final_psnr, "loss": float(state["loss"]), "regrows": regrows,
            "capacity": plan.capacity, "tile_size": ts, "seconds": seconds,
            "step_seconds": replay["seconds"], "timed_steps": replay["steps"],
            "color_grad_l2_term": color_grad_l2_term(field, cfg)}


def _timed_steps(acc: dict, n: int, fn, dev) -> None:
    """Runs `fn` n times between two device synchronizations and adds the wall time to `acc`
    (the steps alone: no capture, snapshot or window readback)."""
    torch.cuda.synchronize(dev)
    t = time.perf_counter()
    for _ in range(n):
        fn()
    torch.cuda.synchronize(dev)
    acc["seconds"] += time.perf_counter() - t
    acc["steps"] += n


def window_readback(plan) -> tuple[int, int]:
    """(sticky binning status, pair count of the plan's last binning) in one device-to-host
    copy.  The pair count is tile_offsets[T], which every binning writes on the device (int32,
    saturated at 2^31 - 1), so a window needs no recount on the host."""
    vals = torch.cat([plan.status.reshape(-1), plan.offsets[-1:]]).to("cpu")
    return int(vals[0]), int(vals[1])


def _seed(state: dict, loss: torch.Tensor) -> torch.Tensor:
    """dL/dL = 1, allocated once (the first, eager step) and reused, so a step holds no fill
    kernel for it."""
    seed = state.get("seed")
    if seed is None or seed.shape != loss.shape or seed.device != loss.device:
        seed = state["seed"] = torch.ones_like(loss)
    return seed


def resolve_tile_size(cfg: FitConfig, tune) -> int:
    """cfg.tile_size, or the autotuner's choice (`tune()`) for "auto".  An explicit size is 16
    or 32, optionally OR-ed with ``render.TILE_PERSISTENT`` (the persistent schedule)."""
    if cfg.tile_size == "auto":
        return int(tune())
    ts = int(cfg.tile_size)
    if HR.tile_edge(ts) not in (16, 32) or ts & ~(HR.TILE_PERSISTENT | 0xff):
        raise ValueError(f"tile_size must be 16, 32 or 'auto', got {cfg.tile_size!r}")
    return ts


def _capacity(pairs: int, cfg: FitConfig) -> int:
    return max(int(pairs * cfg.capacity_slack), cfg.capacity_floor, 1)


def _headroom(pairs: int, cfg: FitConfig) -> int:
    """Capacity after a regrow: the configured slack, but never less than 1.3x the pairs."""
    return max(_capacity(pairs, cfg), int(pairs * 1.3) + 1)


def _capture(step) -> torch.cuda.CUDAGraph:
    """Capture one full step (render, loss, backward, Adam, clamp) as a HIP graph.

    The caller has already run two eager steps on a side stream; they are real optimization
    steps (counted in the iteration budget) and they create the Adam state and gradient buffers
    that the captured step updates in place.
    """
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        step()
    torch.cuda.synchronize()
    return g
