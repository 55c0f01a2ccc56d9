"""Ragged batches: many independent fields fitted by one launch per kernel.

The reference fits many images with a process farm, one ``fit`` per image (batch.py:83-245).  Here
all fields share one set of kernel launches: their Gaussians are concatenated (rows in field
order), their images are concatenated (pixels in field order), and a device field table tells every
kernel which image a tile or a Gaussian belongs to (``ssplat_batch_*`` in include/ssplat.h).

* ``Batch`` - planned field table, workspace and launchers (binning, forward, backward).
* ``fit_batch(fields, targets, cfg, graph=True)`` - one Adam fit over all fields at once, the
  whole step captured as a HIP graph; per-field losses are summed, so every field gets exactly
  the gradient of its own loss and Adam (elementwise) treats every field's rows as its own.
* ``fit_each(fields, targets, cfg)`` - the plain per-image loop over the same fields (one field per
  launch, no graph), the baseline the batched fit is compared with.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import time
from dataclasses import dataclass

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _native
from . import fit as F
from .gaussians import GaussianField
from .loss import batch_l1_ssim_loss, loss_plan
from .metrics import psnr_tensor
from .tune import tune_field, tune_tile_size
from .render import _fade_alpha, _ptr, _stream, grad_slots, tile_edge


def _plan(sizes, counts, tile_size):
    """Completed host field table and (gaussians, tiles, pixels, workspace bytes)."""
    lib = _native.load()
    nf = len(sizes)
    table = (_native.Field * nf)()
    for i, ((h, w), n) in enumerate(zip(sizes, counts)):
        table[i].height = int(h)
        table[i].width = int(w)
        table[i].num_gaussians = int(n)
    g, t, px = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    wsb = ctypes.c_size_t(0)
    _native.check(lib.ssplat_batch_plan(ctypes.addressof(table), nf, int(tile_size),
                                        ctypes.byref(g), ctypes.byref(t), ctypes.byref(px),
                                        ctypes.byref(wsb)), "ssplat_batch_plan")
    return table, int(g.value), int(t.value), int(px.value), int(wsb.value)


def _device_table(table, device):
    raw = bytes(ctypes.string_at(ctypes.addressof(table), ctypes.sizeof(table)))
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)


class Batch:
    """Planned layout of a batch of fields of sizes `sizes` [(H, W)] with `counts` Gaussians,
    rendered with `tile_size`-pixel tiles (16 or 32, optionally with the TILE_PERSISTENT
    schedule flag: every field's tiles in one persistent launch).  The loss has its own table
    (``loss_fields_dev``, ``loss_total_tiles`` units of ``loss.loss_plan``)."""

    def __init__(self, sizes, counts, device, tile_size: int = 16):
        nf = len(sizes)
        if nf == 0 or len(counts) != nf:
            raise ValueError("a batch needs one Gaussian count per field size")
        table, g, t, px, wsb = _plan(sizes, counts, tile_size)
        self.device = torch.device(device)
        self.tile_config = int(tile_size)       # edge and schedule flag, passed to the kernels
        self.tile_size = tile_edge(tile_size)   # the tile edge
        self.num_fields = nf
        self.table = table
        self.total_gaussians = g
        self.total_tiles = t
        self.total_pixels = px
        self.fields_dev = _device_table(table, self.device)
        self.workspace = torch.empty(max(wsb, 256), dtype=torch.uint8, device=self.device)
        loss_table, self.loss_total_tiles, _, _ = loss_plan(sizes)
        self.loss_fields_dev = _device_table(loss_table, self.device)
        self.sizes = [(int(h), int(w)) for h, w in sizes]
        self.gauss_slices = [slice(table[i].first_gaussian,
                                   table[i].first_gaussian + table[i].num_gaussians)
                             for i in range(nf)]
        self.pixel_slices = [slice(table[i].first_pixel,
                                   table[i].first_pixel + table[i].height * table[i].width)
                             for i in range(nf)]

    def _head(self):
        return (_ptr(self.fields_dev), self.num_fields, self.total_gaussians, self.total_tiles,
                self.tile_config)

    def images(self, flat: torch.Tensor):
        """Per-field (H, W, 3) views of a concatenated (total_pixels, 3) image buffer."""
        return [flat[s].view(h, w, 3) for s, (h, w) in zip(self.pixel_slices, self.sizes)]

    def build_tile_index(self, means, conics, radii, capacity: int, cull: bool = False,
                         sigma_cutoff: float = 3.0, out=None):
        lib = _native.load()
        if out is None:
            gids = torch.empty(max(int(capacity), 1), dtype=torch.int32, device=self.device)
            offsets = torch.empty(self.total_tiles + 1, dtype=torch.int32, device=self.device)
            status = torch.zeros(1, dtype=torch.int32, device=self.device)
        else:
            gids, offsets, status = out
        _native.check(lib.ssplat_batch_build_tile_index(
            *self._head(), _ptr(means), _ptr(conics), _ptr(radii), int(bool(cull)),
            float(sigma_cutoff), _ptr(offsets), _ptr(gids), int(min(capacity, gids.numel())),
            _ptr(status), _ptr(self.workspace), self.workspace.numel(), _stream(self.device)),
            "ssplat_batch_build_tile_index")
        return gids, offsets, status

    def pair_count(self, means, conics, radii, cull: bool = False, sigma_cutoff: float = 3.0):
        """Exact pair count (binning into a large buffer, one host readback)."""
        gids, offsets, status = self.build_tile_index(means, conics, radii, 1, cull,
                                                      sigma_cutoff)
        return int(offsets[-1].item())

    def forward(self, means, conics, colors, radii, opacities, gids, offsets, normalize=True,
                eps=1e-8, support_fade=False, sigma_cutoff=3.0, out=None, den=None):
        lib = _native.load()
        if out is None:
            out = torch.empty(self.total_pixels, 3, dtype=torch.float32, device=self.device)
        _native.check(lib.ssplat_batch_forward_tiled(
            *self._head(), _ptr(means), _ptr(conics), _ptr(colors), _ptr(radii),
            _ptr(opacities), _ptr(gids), _ptr(offsets), gids.numel(), int(bool(normalize)),
            float(eps), _fade_alpha(support_fade), float(sigma_cutoff), _ptr(out), _ptr(den),
            _ptr(self.workspace), self.workspace.numel(), _stream(self.device)),
            "ssplat_batch_forward_tiled")
        return out

    def backward(self, grad_out, means, conics, colors, radii, opacities, den, out, gids,
                 offsets, normalize=True, eps=1e-8, support_fade=False, sigma_cutoff=3.0,
                 slots=None, slot_capacity=None, status=None, after_forward=False):
        """Batched deterministic backward; ``den``/``out`` are the batched forward's outputs
        (``None`` allowed for additive renders).  ``slots`` holds ``slot_capacity`` gradient
        slots and ``status`` records an overflow (both from a ``BatchPlan`` in the fit loop).
        Without them, the slots get the pair buffer's length and an overflow raises (one host
        synchronization).  ``after_forward``: the batch's workspace still holds the records of
        ``forward`` on the same inputs (``ssplat_batch_backward_tiled_after_forward``)."""
        lib = _native.load()
        has_op = opacities is not None and opacities.numel() > 0
        g_m = torch.empty_like(means)
        g_q = torch.empty_like(conics)
        g_c = torch.empty_like(colors)
        g_o = torch.empty_like(opacities) if has_op else None
        check_here = slots is None
        if check_here:
            slot_capacity = gids.numel()
            slots = grad_slots(slot_capacity, self.device)
            status = torch.zeros(1, dtype=torch.int32, device=self.device)
        fn, name = ((lib.ssplat_batch_backward_tiled_after_forward,
                     "ssplat_batch_backward_tiled_after_forward") if after_forward
                    else (lib.ssplat_batch_backward_tiled, "ssplat_batch_backward_tiled"))
        _native.check(fn(
            *self._head(), _ptr(grad_out), _ptr(means), _ptr(conics), _ptr(colors), _ptr(radii),
            _ptr(opacities), _ptr(den), _ptr(out), _ptr(gids), _ptr(offsets), gids.numel(),
            int(bool(normalize)), float(eps), _fade_alpha(support_fade), float(sigma_cutoff),
            _ptr(g_m), _ptr(g_q), _ptr(g_c), _ptr(g_o), _ptr(slots), int(slot_capacity),
            _ptr(status), _ptr(self.workspace), self.workspace.numel(), _stream(self.device)),
            name)
        if check_here and int(status.item()) != 0:
            raise RuntimeError("batched backward: gradient slot capacity exceeded")
        return g_m, g_q, g_c, g_o


class BatchPlan:
    """Pair buffer, offsets, sticky status and gradient slots of a batch (capturable binning
    and backward); the slot buffer has the pair buffer's capacity."""

    def __init__(self, batch: Batch, capacity: int, recompute_normalizer: bool = True):
        self.batch = batch
        self.capacity = max(int(capacity), 1)
        self.recompute_normalizer = bool(recompute_normalizer)
        dev = batch.device
        self.gids = torch.empty(self.capacity, dtype=torch.int32, device=dev)
        self.offsets = torch.empty(batch.total_tiles + 1, dtype=torch.int32, device=dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.slots = grad_slots(self.capacity, dev)
        self.den = (None if self.recompute_normalizer
                    else torch.empty(batch.total_pixels, dtype=torch.float32, device=dev))


class _BatchRender(Function):
    @staticmethod
    def forward(ctx, means, conics, colors, radii, opacities, plan, normalize, eps, fade,
                sigma_cutoff):
        b = plan.batch
        means, conics, colors = means.contiguous(), conics.contiguous(), colors.contiguous()
        radii, opacities = radii.contiguous(), opacities.contiguous()
        b.build_tile_index(means, conics, radii, plan.capacity, _fade_alpha(fade) == 1.0,
                           sigma_cutoff,
                           out=(plan.gids, plan.offsets, plan.status))
        out = b.forward(means, conics, colors, radii, opacities, plan.gids, plan.offsets,
                        normalize, eps, fade, sigma_cutoff, den=plan.den)
        ctx.save_for_backward(means, conics, colors, radii, opacities, plan.den, out)
        ctx.plan = plan
        ctx.args = (normalize, eps, fade, sigma_cutoff)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_out):
        means, conics, colors, radii, opacities, den, out = ctx.saved_tensors
        plan = ctx.plan
        g_m, g_q, g_c, g_o = plan.batch.backward(grad_out.contiguous(), means, conics, colors,
                                                 radii, opacities, den, out, plan.gids,
                                                 plan.offsets, *ctx.args, slots=plan.slots,
                                                 slot_capacity=plan.capacity,
                                                 status=plan.status, after_forward=True)
        if opacities.numel() == 0:
            g_o = None
        return g_m, g_q, g_c, None, g_o, None, None, None, None, None


def concat_fields(fields: list[GaussianField]) -> GaussianField:
    """One field holding every field's rows in order (opacity logits required on all or none).

    The covariance filter variance and the scale caps travel with the rows.  A field without
    them gets the neutral value where another field has them: variance 0 (adding 0 leaves the
    producer's float32 arithmetic unchanged) and cap +inf (no cap)."""
    has_op = [f.opacities is not None for f in fields]
    if any(has_op) and not all(has_op):
        raise ValueError("either every field has opacities or none has")

    def cat(name):
        return torch.cat([getattr(f, name).detach() for f in fields], 0).clone()

    def cat_optional(name, shape_of, fill):
        if all(getattr(f, name) is None for f in fields):
            return None
        parts = []
        for f in fields:
            v = getattr(f, name)
            ref = f.log_scales.detach()
            parts.append(torch.full(shape_of(f), fill, dtype=torch.float32, device=ref.device)
                         if v is None else v.detach().to(device=ref.device,
                                                        dtype=torch.float32).reshape(shape_of(f)))
        return torch.cat(parts, 0).clone()

    return GaussianField(cat("means"), cat("log_scales"), cat("rotations"), cat("colors"),
                         cat("opacities") if all(has_op) else None,
                         scale_max=cat_optional("scale_max", lambda f: (f.n, 2), math.inf),
                         filter_variance=cat_optional("filter_variance", lambda f: (f.n,), 0.0))


@dataclass
class BatchResult:
    psnr: list
    seconds: float
    steps: int
    tile_size: int | None = None  # the batched fit's render tile (per field for the plain loop)
    step_seconds: float = 0.0     # the timed steps alone (graph replays or eager steps)
    timed_steps: int = 0
    tune_ms: dict | None = None   # the autotuner's fwd+bwd time per tile configuration


def fit_batch(fields: list[GaussianField], targets: list[torch.Tensor], cfg: F.FitConfig,
              graph: bool = True) -> BatchResult:
    """Fit every field to its target in one batched step per iteration.  The fitted rows are
    copied back into `fields`."""
    if cfg.color_basis != "constant" or any(f.color_grads is not None for f in fields):
        raise ValueError("fit_batch renders constant colors; fit affine-color fields with fit()")
    dev = targets[0].device
    sizes = [tuple(t.shape[:2]) for t in targets]
    counts = [f.n for f in fields]
    big = concat_fields(fields).trainable()

    tuned = {}

    def tune():
        conics, radii, op = big.produce(cfg.sigma_cutoff)
        best, tuned["ms"] = tune_tile_size(
            sizes, counts, big.means.detach(), conics.detach(), big.colors.detach(), radii,
            None if op is None else op.detach(), normalize=cfg.normalize,
            eps=cfg.normalization_eps, support_fade=cfg.support_fade,
            sigma_cutoff=cfg.sigma_cutoff, recompute_normalizer=cfg.recompute_normalizer)
        return best

    batch = Batch(sizes, counts, dev, tile_size=F.resolve_tile_size(cfg, tune))
    targets_flat = torch.cat([t.reshape(-1, 3) for t in targets], 0).contiguous()
    moments = None  # blurred inline by the loss, as in fit.fit (the same bits, less traffic)
    opt = F.make_optimizer(big, cfg)
    # Per-field log-scale upper bounds (one per element when a field caps its scales).
    uppers = [F.log_scale_upper(f, h, w) for f, (h, w) in zip(fields, sizes)]
    hi = torch.cat([u.to(dev) if torch.is_tensor(u)
                    else torch.full((2 * f.n,), u, dtype=torch.float32, device=dev)
                    for f, u in zip(fields, uppers)], 0)
    opt.set_clamp(big.log_scales, math.log(0.35), hi)
    w = cfg.ssim_weight

    def count():
        with torch.no_grad():
            return batch.pair_count(big.means, big.conics(), big.radii(cfg.sigma_cutoff),
                                    _fade_alpha(cfg.support_fade) == 1.0, cfg.sigma_cutoff)

    plan = BatchPlan(batch, F._capacity(count(), cfg), cfg.recompute_normalizer)

    seeds = {}

    def step():
        opt.zero_grad(set_to_none=True)  # gradients assigned, not accumulated
        conics, radii, op = big.produce(cfg.sigma_cutoff)
        op = big.means.new_empty(0) if op is None else op
        img = _BatchRender.apply(big.means, conics, big.colors, radii, op, plan, cfg.normalize,
                                 float(cfg.normalization_eps), _fade_alpha(cfg.support_fade),
                                 float(cfg.sigma_cutoff))
        loss = batch_l1_ssim_loss(batch, img, targets_flat, w, moments)
        loss.backward(F._seed(seeds, loss))
        opt.step()  # Adam and the per-field log-scale clamp

    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    done = 0
    g = None
    replay = {"steps": 0, "seconds": 0.0}
    while done < cfg.iters:
        n = min(cfg.window, cfg.iters - done)
        snap = F._Snapshot(big, opt)
        plan.status.zero_()
        if graph and done >= 2:
            if g is None:
                g = F._capture(step)
            F._timed_steps(replay, n, g.replay, dev)
        elif graph:
            n = min(n, 2 - done)
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(side):
                for _ in range(n):
                    step()
            torch.cuda.current_stream(dev).wait_stream(side)
        else:
            F._timed_steps(replay, n, step, dev)
        status, p = F.window_readback(plan)
        if status != 0:
            snap.restore(big, opt)
            plan = BatchPlan(batch, max(2 * plan.capacity, F._capacity(p, cfg)),
                             cfg.recompute_normalizer)
            g = None
            continue
        done += n
        if p * 1.15 > plan.capacity:
            plan = BatchPlan(batch, F._headroom(p, cfg), cfg.recompute_normalizer)
            g = None
    torch.cuda.synchronize(dev)
    seconds = time.perf_counter() - t0
    with torch.no_grad():
        op = big.opacity_values()
        op = big.means.new_empty(0) if op is None else op
        gids, offsets, _ = batch.build_tile_index(big.means, big.conics(),
                                                  big.radii(cfg.sigma_cutoff), plan.capacity,
                                                  _fade_alpha(cfg.support_fade) == 1.0,
                                                  cfg.sigma_cutoff)
        img = batch.forward(big.means, big.conics(), big.colors, big.radii(cfg.sigma_cutoff), op,
                            gids, offsets, cfg.normalize, cfg.normalization_eps,
                            cfg.support_fade, cfg.sigma_cutoff)
        psnrs = [float(psnr_tensor(im, t)) for im, t in zip(batch.images(img), targets)]
        for f, sl in zip(fields, batch.gauss_slices):
            for name in ("means", "log_scales", "rotations", "colors", "opacities"):
                src = getattr(big, name)
                if src is not None:
                    setattr(f, name, src.detach()[sl].clone())
    return BatchResult(psnrs, seconds, cfg.iters, batch.tile_config, replay["seconds"],
                       replay["steps"], tuned.get("ms"))


def fit_each(fields: list[GaussianField], targets: list[torch.Tensor], cfg: F.FitConfig
             ) -> BatchResult:
    """The plain per-image loop: each field fitted on its own, eagerly, one after the other."""
    dev = targets[0].device
    # Tile sizes are chosen before the clock starts, as the batched fit's is.
    cfgs = [dataclasses.replace(cfg, tile_size=F.resolve_tile_size(
        cfg, lambda f=f, t=t: tune_field(f, int(t.shape[0]), int(t.shape[1]), cfg)[0]))
        for f, t in zip(fields, targets)]
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    psnrs, step_seconds, timed = [], 0.0, 0
    for f, t, c in zip(fields, targets, cfgs):
        r = F.fit(f, t, c, graph=False)
        psnrs.append(r["psnr"])
        step_seconds += r["step_seconds"]
        timed += r["timed_steps"]
    torch.cuda.synchronize(dev)
    return BatchResult(psnrs, time.perf_counter() - t0, cfg.iters * len(fields), None,
                       step_seconds, timed)
