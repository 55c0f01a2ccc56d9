"""Render entry points of the HIP rasterizer: the drop-in for StructSplat's render path.

The reference renders through ``render_field`` (render.py:319-414), whose CUDA modes go through
``render_cuda_exact`` -> ``_ExactRenderCuda`` -> a pybind11 extension (cuda_render.py:52-294,
render_ext.cpp:139-317).  This module keeps those Python names, arguments, mode strings and error
phrases and routes every mode to the gfx950 kernels of ``libssplat_hip.so`` through the C ABI:

* ``render_field(...)`` - mode dispatcher.  ``normalized`` / ``additive`` and all of the
  reference's CUDA mode names (plus ``hip``, ``hip_additive``) run the HIP kernels on GPU tensors.
  There is no CPU path: CPU tensors raise the reference's "requires CUDA tensors" error.
* ``render_hip_exact(...)`` - the ``render_cuda_exact`` signature (cuda_render.py:256-294).
* ``build_tile_index(...)`` - the native ``build_tile_index`` (render_ext.cpp:139-167).
* ``render_hip_exact_with_coverage(...)`` - ``render_cuda_exact_with_coverage``
  (cuda_render.py:229-253): normalized RGB plus raw coverage, no autograd.

By default the backward recomputes the per-pixel normalizer ``den`` with a second sweep of the
tile lists (bit-identical to the forward's) instead of storing it, and reads only the saved
image; ``recompute_normalizer=False`` saves ``den`` and ``out`` as the reference does
(render_ext.cu:972-1008), which skips that sweep for 4 bytes a pixel of memory.  Other
differences by design: tile lists are built without a radix sort
(ascending Gaussian id per tile, like the reference's stable sort); the forward is bitwise
deterministic.
"""
from __future__ import annotations

import math

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _native

_EPS = 1e-8
DEFAULT_TILE_SIZE = 16
_INT32_SATURATED = 2**31 - 1
# Render schedule flag OR-ed into a tile size (SSPLAT_TILE_PERSISTENT): the forward and backward
# kernels run as one persistent launch over per-XCD tile queues instead of one workgroup a tile.
# The binning ignores it; results are bit-identical either way.
TILE_PERSISTENT = 0x100


def tile_edge(tile_size: int) -> int:
    """The tile edge of a tile size that may carry the TILE_PERSISTENT schedule flag."""
    return int(tile_size) & ~TILE_PERSISTENT

_NORMALIZED_MODES = (
    "normalized", "hip", "hip_normalized", "hip_tiled", "hip_tiled_normalized",
    "cuda", "cuda_normalized", "cuda_block_reduce", "cuda_tiled", "cuda_tiled_normalized",
)
_ADDITIVE_MODES = (
    "additive", "hip_additive", "hip_tiled_additive", "cuda_additive", "cuda_tiled_additive",
)

_DEVICE_MESSAGE = (
    "StructSplat HIP renderer requires CUDA tensors (torch's device type for ROCm GPUs); "
    "pass device='cuda'."
)


def _validate_normalization_eps(value: float) -> float:
    """eps must be finite and positive (render.py:33-37)."""
    eps = float(value)
    if not math.isfinite(eps) or eps <= 0.0:
        raise ValueError(f"normalization_eps must be finite and > 0, got {value}")
    return eps


def _resolve_support_fade_alpha(support_fade: bool, support_fade_alpha: float | None = None
                                ) -> float:
    """Fade strength in [0, 1] (render.py:94-98)."""
    if support_fade_alpha is None:
        return 1.0 if support_fade else 0.0
    return max(0.0, min(1.0, float(support_fade_alpha)))


def _stream(device: torch.device) -> int:
    return int(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None or t.numel() == 0:
        return None
    return int(t.data_ptr())


def _check_field(means, conics, colors, radii, opacities, need_colors: bool = True) -> None:
    """Tensor contract of the reference extension (render_ext.cpp:94-125, cuda_render.py:274-277)."""
    if not means.is_cuda:
        raise RuntimeError(_DEVICE_MESSAGE)
    tensors = (means, conics, colors) if need_colors else (means, conics)
    if any(t.dtype != torch.float32 for t in tensors):
        raise RuntimeError("StructSplat HIP renderer currently supports float32 tensors only.")
    n = means.shape[0]
    if means.dim() != 2 or means.shape[1] != 2:
        raise ValueError("means must be (N, 2)")
    if conics.dim() != 2 or conics.shape != (n, 3):
        raise ValueError("conics must be (N, 3)")
    if need_colors and (colors.dim() != 2 or colors.shape != (n, 3)):
        raise ValueError("colors must be (N, 3)")
    if radii.dim() != 2 or radii.shape != (n, 2):
        raise ValueError("radii must be (N, 2)")
    if radii.dtype not in (torch.int64, torch.int32):
        raise ValueError("radii must be an integer tensor (int64)")
    if opacities is not None and opacities.numel() not in (0, n):
        raise ValueError("opacities must be empty or length N")
    if opacities is not None and opacities.numel() and opacities.dtype != torch.float32:
        raise RuntimeError("opacities must be empty or float32")
    for t in (conics, colors, radii) + ((opacities,) if opacities is not None else ()):
        if t.numel() and t.device != means.device:
            raise RuntimeError("all field tensors must be on the same device")


def _check_image(height: int, width: int, tile_size: int) -> None:
    if int(height) <= 0 or int(width) <= 0:
        raise ValueError("height and width must be positive")
    ts = tile_edge(tile_size)
    if ts <= 0 or ts ** 2 > 1024:
        raise ValueError("tile_size must be positive with tile_size^2 <= 1024")


def _fade_alpha(support_fade) -> float:
    """Fade strength for the C ABI: a bool maps to 0/1, a number is clamped to [0, 1]."""
    if isinstance(support_fade, bool):
        return 1.0 if support_fade else 0.0
    return max(0.0, min(1.0, float(support_fade)))


def _radii64(radii: torch.Tensor) -> torch.Tensor:
    return radii.detach().to(torch.int64).contiguous()


def num_tiles(height: int, width: int, tile_size: int = DEFAULT_TILE_SIZE) -> int:
    ts = tile_edge(tile_size)
    return ((int(width) + ts - 1) // ts) * ((int(height) + ts - 1) // ts)


def workspace(n: int, height: int, width: int, tile_size: int, device) -> torch.Tensor:
    """Caller-owned scratch for one call sequence (the library never allocates)."""
    nbytes = _native.workspace_bytes(n, height, width, tile_size)
    return torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)


def affine_workspace(n: int, height: int, width: int, device) -> torch.Tensor:
    """Scratch of the affine-color entry points (tile size 16)."""
    nbytes = _native.affine_workspace_bytes(n, height, width, DEFAULT_TILE_SIZE)
    return torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)


def affine_grad_slots(slot_capacity: int, device) -> torch.Tensor:
    """Zero-filled gradient slot buffer of the affine-color backward."""
    nbytes = _native.affine_grad_slots_bytes(max(int(slot_capacity), 0))
    return torch.zeros(nbytes, dtype=torch.uint8, device=device)


def grad_slots(slot_capacity: int, device) -> torch.Tensor:
    """Zero-filled gradient slot buffer for deterministic backwards with up to `slot_capacity`
    slots (one per support-rectangle tile of each Gaussian); reusable across calls."""
    nbytes = _native.grad_slots_bytes(max(int(slot_capacity), 0))
    return torch.zeros(nbytes, dtype=torch.uint8, device=device)


@torch.no_grad()
def rect_pair_count(means, radii, H: int, W: int, tile_size: int, ws=None) -> int:
    """Total slot count of a field: support-rectangle tiles summed over the Gaussians (equal to
    the pair count of lists built without ellipse cull).  One host synchronization."""
    dev = means.device
    means = means.detach().contiguous()
    radii = _radii64(radii)
    n = means.shape[0]
    if ws is None:
        ws = workspace(n, H, W, tile_size, dev)
    total = torch.empty(1, dtype=torch.int64, device=dev)
    _native.check(_native.load().ssplat_rect_pair_count(
        _ptr(means), _ptr(radii), n, int(H), int(W), int(tile_size), _ptr(total), _ptr(ws),
        ws.numel(), _stream(dev)), "ssplat_rect_pair_count")
    value = int(total.item())
    if value < 0:
        raise RuntimeError("gradient slot count exceeds int32; reduce N or image size")
    return value


@torch.no_grad()
def build_tile_index(means, conics, radii, H: int, W: int, tile_size: int = DEFAULT_TILE_SIZE,
                     ellipse_cull: bool = False, sigma_cutoff: float = 3.0):
    """Per-tile Gaussian lists: ``(gids int32 (P,), offsets int32 (T+1,))``.

    Mirrors the reference's native ``build_tile_index`` (render_ext.cpp:139-167,
    render_ext.cu:858-952): tile ``t`` owns ``gids[offsets[t]:offsets[t+1]]`` in ascending Gaussian
    id.  Culled pairs are never written, so ``gids`` has exactly ``offsets[-1]`` entries (the
    reference keeps culled pairs as a sorted-out tail beyond ``offsets[-1]``).  One host
    synchronization sizes ``gids`` exactly, like the reference's readback (render_ext.cu:910);
    ``build_tile_index_capacity`` is the synchronization-free variant.
    """
    _check_field(means, conics, conics, radii, None, need_colors=False)
    _check_image(H, W, tile_size)
    lib = _native.load()
    dev = means.device
    means = means.detach().contiguous()
    conics = conics.detach().contiguous()
    radii = _radii64(radii)
    n = means.shape[0]
    T = num_tiles(H, W, tile_size)
    offsets = torch.empty(T + 1, dtype=torch.int32, device=dev)
    ws = workspace(n, H, W, tile_size, dev)
    stream = _stream(dev)
    args = (_ptr(means), _ptr(conics), _ptr(radii), n, int(H), int(W), int(tile_size),
            int(bool(ellipse_cull)), float(sigma_cutoff))
    _native.check(lib.ssplat_tile_index_count(*args, _ptr(offsets), _ptr(ws), ws.numel(), stream),
                  "ssplat_tile_index_count")
    total = int(offsets[T].item())
    if total >= _INT32_SATURATED or total < 0:
        raise RuntimeError("tile index pair count exceeds int32; reduce N or image size")
    gids = torch.empty(total, dtype=torch.int32, device=dev)
    _native.check(lib.ssplat_tile_index_fill(*args, _ptr(offsets), _ptr(gids), total, None,
                                             _ptr(ws), ws.numel(), stream),
                  "ssplat_tile_index_fill")
    return gids, offsets


@torch.no_grad()
def build_tile_index_capacity(means, conics, radii, H: int, W: int, capacity: int,
                              tile_size: int = DEFAULT_TILE_SIZE, ellipse_cull: bool = False,
                              sigma_cutoff: float = 3.0, out=None):
    """Synchronization-free binning into a caller-sized gid buffer (HIP-graph capturable).

    Returns ``(gids (capacity,), offsets (T+1,), status (1,))``; ``status`` is a device int32 that
    reads 0 when every pair fit and 2 (capacity exceeded) otherwise.  ``out`` may pass
    preallocated ``(gids, offsets, status, workspace)`` tensors to reuse; the status word is
    sticky (an overflow stays recorded until the caller clears it).
    """
    _check_field(means, conics, conics, radii, None, need_colors=False)
    _check_image(H, W, tile_size)
    lib = _native.load()
    dev = means.device
    n = means.shape[0]
    T = num_tiles(H, W, tile_size)
    if out is None:
        gids = torch.empty(max(int(capacity), 1), dtype=torch.int32, device=dev)
        offsets = torch.empty(T + 1, dtype=torch.int32, device=dev)
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        ws = workspace(n, H, W, tile_size, dev)
    else:
        gids, offsets, status, ws = out
    means = means.detach().contiguous()
    conics = conics.detach().contiguous()
    radii = _radii64(radii)
    _native.check(lib.ssplat_build_tile_index(
        _ptr(means), _ptr(conics), _ptr(radii), n, int(H), int(W), int(tile_size),
        int(bool(ellipse_cull)), float(sigma_cutoff), _ptr(offsets), _ptr(gids),
        int(min(capacity, gids.numel())), _ptr(status), _ptr(ws), ws.numel(), _stream(dev)),
        "ssplat_build_tile_index")
    return gids, offsets, status


def forward_tiled(means, conics, colors, radii, opacities, gids, offsets, H: int, W: int,
                  normalize: bool, eps: float, tile_size: int, support_fade: bool,
                  sigma_cutoff: float, out=None, den=None, ws=None):
    """One tiled forward through the C ABI; returns ``(out (H,W,3), den (H,W) or None)``."""
    lib = _native.load()
    dev = means.device
    n = means.shape[0]
    if out is None:
        out = torch.empty(int(H), int(W), 3, dtype=torch.float32, device=dev)
    if ws is None:
        ws = workspace(n, H, W, tile_size, dev)
    _native.check(lib.ssplat_forward_tiled(
        _ptr(means), _ptr(conics), _ptr(colors), _ptr(radii), _ptr(opacities), n, _ptr(gids),
        _ptr(offsets), gids.numel(), int(H), int(W), int(tile_size), int(bool(normalize)),
        float(eps), _fade_alpha(support_fade), float(sigma_cutoff), _ptr(out), _ptr(den),
        _ptr(ws), ws.numel(), _stream(dev)), "ssplat_forward_tiled")
    return out, den


def backward_tiled(grad_out, means, conics, colors, radii, opacities, den, out, gids, offsets,
                   H: int, W: int, normalize: bool, eps: float, tile_size: int,
                   support_fade: bool, sigma_cutoff: float, grads=None, ws=None, slots=None,
                   slot_capacity=None, status=None, after_forward: bool = False):
    """One tiled backward; returns ``(g_means, g_conics, g_colors, g_opacities or None)``.

    ``den`` and ``out`` are the forward's coverage and image for the same inputs (the
    reference's saved tensors, cuda_render.py:111-113); additive renders may pass ``None``.
    The gradients are deterministic.  ``slots`` (from ``grad_slots``) holds ``slot_capacity``
    gradient slots; without it a buffer is made for the field's exact slot count.  ``status``
    (device int32) is set to 2 if the slot capacity was too small.  ``after_forward``: ``ws``
    holds the records of ``forward_tiled`` on the same inputs, which are used as they are
    (``ssplat_backward_tiled_after_forward``).
    """
    lib = _native.load()
    dev = means.device
    n = means.shape[0]
    has_opacity = opacities is not None and opacities.numel() > 0
    if grads is None:
        grads = (torch.empty_like(means), torch.empty_like(conics), torch.empty_like(colors),
                 torch.empty_like(opacities) if has_opacity else None)
    g_means, g_conics, g_colors, g_opac = grads
    if ws is None:
        ws = workspace(n, H, W, tile_size, dev)
    if slots is None:
        if slot_capacity is None:
            slot_capacity = rect_pair_count(means, radii, H, W, tile_size, ws)
        slots = grad_slots(slot_capacity, dev)
    elif slot_capacity is None:
        raise ValueError("slot_capacity is required with a slot buffer")
    if after_forward and ws is None:
        raise ValueError("after_forward needs the forward's workspace")
    fn, name = ((lib.ssplat_backward_tiled_after_forward, "ssplat_backward_tiled_after_forward")
                if after_forward else (lib.ssplat_backward_tiled, "ssplat_backward_tiled"))
    _native.check(fn(
        _ptr(grad_out), _ptr(means), _ptr(conics), _ptr(colors), _ptr(radii), _ptr(opacities),
        _ptr(den), _ptr(out), n, _ptr(gids), _ptr(offsets), gids.numel(), int(H), int(W),
        int(tile_size), int(bool(normalize)), float(eps), _fade_alpha(support_fade),
        float(sigma_cutoff), _ptr(g_means), _ptr(g_conics), _ptr(g_colors), _ptr(g_opac),
        _ptr(slots), int(slot_capacity), _ptr(status), _ptr(ws), ws.numel(), _stream(dev)), name)
    return g_means, g_conics, g_colors, g_opac


def forward_tiled_affine(means, conics, colors, radii, opacities, color_grads, scales, rotations,
                         gids, offsets, H: int, W: int, normalize: bool, eps: float,
                         support_fade: bool, sigma_cutoff: float, den=None, ws=None):
    """Tiled forward with affine colors (tile size 16); returns ``(out (H,W,3), den or None)``.
    ``ws`` comes from ``affine_workspace``."""
    lib = _native.load()
    dev = means.device
    n = means.shape[0]
    out = torch.empty(int(H), int(W), 3, dtype=torch.float32, device=dev)
    if ws is None:
        ws = affine_workspace(n, H, W, dev)
    _native.check(lib.ssplat_forward_tiled_affine(
        _ptr(means), _ptr(conics), _ptr(colors), _ptr(radii), _ptr(opacities), _ptr(color_grads),
        _ptr(scales), _ptr(rotations), n, _ptr(gids), _ptr(offsets), gids.numel(), int(H),
        int(W), DEFAULT_TILE_SIZE, int(bool(normalize)), float(eps), _fade_alpha(support_fade),
        float(sigma_cutoff), _ptr(out), _ptr(den), _ptr(ws), ws.numel(), _stream(dev)),
        "ssplat_forward_tiled_affine")
    return out, den


def backward_tiled_affine(grad_out, means, conics, colors, radii, opacities, color_grads, scales,
                          rotations, den, out, gids, offsets, H: int, W: int, normalize: bool,
                          eps: float, support_fade: bool, sigma_cutoff: float, slot_capacity: int,
                          ws=None, slots=None, status=None):
    """Backward of ``forward_tiled_affine``; returns the gradients of means, conics, colors,
    opacities (or None), color_grads, scales and rotations.  Deterministic."""
    lib = _native.load()
    dev = means.device
    n = means.shape[0]
    has_opacity = opacities is not None and opacities.numel() > 0
    grads = (torch.empty_like(means), torch.empty_like(conics), torch.empty_like(colors),
             torch.empty_like(opacities) if has_opacity else None, torch.empty_like(color_grads),
             torch.empty_like(scales), torch.empty_like(rotations))
    if ws is None:
        ws = affine_workspace(n, H, W, dev)
    if slots is None:
        slots = affine_grad_slots(slot_capacity, dev)
    _native.check(lib.ssplat_backward_tiled_affine(
        _ptr(grad_out), _ptr(means), _ptr(conics), _ptr(colors), _ptr(radii), _ptr(opacities),
        _ptr(color_grads), _ptr(scales), _ptr(rotations), _ptr(den), _ptr(out), n, _ptr(gids),
        _ptr(offsets), gids.numel(), int(H), int(W), DEFAULT_TILE_SIZE, int(bool(normalize)),
        float(eps), _fade_alpha(support_fade), float(sigma_cutoff), *(_ptr(g) for g in grads),
        _ptr(slots), int(slot_capacity), _ptr(status), _ptr(ws), ws.numel(), _stream(dev)),
        "ssplat_backward_tiled_affine")
    return grads


class _ExactRenderHip(Function):
    """Autograd boundary of the HIP renderer (the role of ``_ExactRenderCuda``,
    cuda_render.py:52-174): gradients for means, conics, colors and opacities only.

    With ``recompute_normalizer`` (the default) the per-pixel normalizer ``den`` is not stored:
    the backward recomputes it from the tile lists (bit-identical to the forward's) and reads
    only the saved image.  Otherwise ``den`` and ``out`` are saved, as the reference does."""

    @staticmethod
    def forward(ctx, means, conics, colors, radii, opacities, height, width, normalize, eps,
                tile_size, support_fade, sigma_cutoff, tile_ellipse_cull, recompute_normalizer):
        means = means.contiguous()
        conics = conics.contiguous()
        colors = colors.contiguous()
        radii = _radii64(radii)
        opacities = opacities.contiguous()
        # Culling is exact only under full support fade (cuda_render.py:80-85): with alpha = 1
        # the faded weight is exactly 0 beyond sigma_cutoff.
        alpha = _fade_alpha(support_fade)
        cull = bool(tile_ellipse_cull) and alpha == 1.0
        gids, offsets = build_tile_index(means, conics, radii, int(height), int(width),
                                         int(tile_size), cull, float(sigma_cutoff))
        ws = workspace(means.shape[0], height, width, tile_size, means.device)
        # Gradient slots: one per support-rectangle tile, i.e. the pair count unless culled.
        ctx.slot_capacity = (rect_pair_count(means, radii, int(height), int(width),
                                             int(tile_size), ws) if cull else gids.numel())
        keep_den = bool(normalize) and not recompute_normalizer
        den = (torch.empty(int(height), int(width), dtype=torch.float32, device=means.device)
               if keep_den else None)
        out, _ = forward_tiled(means, conics, colors, radii, opacities, gids, offsets,
                               int(height), int(width), bool(normalize), float(eps),
                               int(tile_size), alpha, float(sigma_cutoff), den=den, ws=ws)
        ctx.save_for_backward(means, conics, colors, radii, opacities, den, out, gids, offsets)
        ctx.ws = ws  # holds the forward's packed records, which the backward reuses
        ctx.height = int(height)
        ctx.width = int(width)
        ctx.normalize = bool(normalize)
        ctx.eps = float(eps)
        ctx.tile_size = int(tile_size)
        ctx.support_fade = alpha
        ctx.sigma_cutoff = float(sigma_cutoff)
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_out):
        means, conics, colors, radii, opacities, den, out, gids, offsets = ctx.saved_tensors
        need_means, need_conics, need_colors, _, need_opac = ctx.needs_input_grad[:5]
        none_tail = (None,) * 9
        if not (need_means or need_conics or need_colors or need_opac):
            return (None,) * 5 + none_tail
        g_means, g_conics, g_colors, g_opac = backward_tiled(
            grad_out.contiguous(), means, conics, colors, radii, opacities, den, out, gids,
            offsets, ctx.height, ctx.width, ctx.normalize, ctx.eps, ctx.tile_size, ctx.support_fade,
            ctx.sigma_cutoff, ws=ctx.ws, slot_capacity=ctx.slot_capacity, after_forward=True)
        if opacities.numel() == 0 or not need_opac:
            g_opac = None
        return (
            g_means if need_means else None,
            g_conics if need_conics else None,
            g_colors if need_colors else None,
            None,
            g_opac,
        ) + none_tail


class _AffineRenderHip(Function):
    """Autograd boundary of the affine-color render (render.py:110-123 inside the reference's
    torch renderer): gradients for means, conics, colors, opacities, color_grads, scales and
    rotations.  The normalizer is recomputed in the backward like ``_ExactRenderHip``'s."""

    @staticmethod
    def forward(ctx, means, conics, colors, radii, opacities, color_grads, scales, rotations,
                height, width, normalize, eps, support_fade, sigma_cutoff):
        means, conics, colors = means.contiguous(), conics.contiguous(), colors.contiguous()
        color_grads = color_grads.contiguous()
        scales, rotations = scales.contiguous(), rotations.contiguous()
        radii = _radii64(radii)
        opacities = opacities.contiguous()
        alpha = _fade_alpha(support_fade)
        cull = alpha == 1.0
        H, W = int(height), int(width)
        gids, offsets = build_tile_index(means, conics, radii, H, W, DEFAULT_TILE_SIZE, cull,
                                         float(sigma_cutoff))
        ws = affine_workspace(means.shape[0], H, W, means.device)
        ctx.slot_capacity = (rect_pair_count(means, radii, H, W, DEFAULT_TILE_SIZE, ws)
                             if cull else gids.numel())
        out, _ = forward_tiled_affine(means, conics, colors, radii, opacities, color_grads,
                                      scales, rotations, gids, offsets, H, W, bool(normalize),
                                      float(eps), alpha, float(sigma_cutoff), ws=ws)
        ctx.save_for_backward(means, conics, colors, radii, opacities, color_grads, scales,
                              rotations, out, gids, offsets)
        ctx.args = (H, W, bool(normalize), float(eps), alpha, float(sigma_cutoff))
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_out):
        (means, conics, colors, radii, opacities, color_grads, scales, rotations, out, gids,
         offsets) = ctx.saved_tensors
        needs = ctx.needs_input_grad[:8]
        none_tail = (None,) * 6
        if not any(needs):
            return (None,) * 8 + none_tail
        H, W, normalize, eps, alpha, sigma = ctx.args
        g = backward_tiled_affine(grad_out.contiguous(), means, conics, colors, radii, opacities,
                                  color_grads, scales, rotations, None,
                                  out if normalize else None, gids, offsets, H, W, normalize,
                                  eps, alpha, sigma, ctx.slot_capacity)
        g_means, g_conics, g_colors, g_opac, g_cg, g_scales, g_rot = g
        if opacities.numel() == 0:
            g_opac = None
        grads = (g_means, g_conics, g_colors, None, g_opac, g_cg, g_scales, g_rot)
        return tuple(x if need else None for x, need in zip(grads, needs)) + none_tail


def render_hip_affine(means, conics, colors, radii, H: int, W: int, color_grads, scales,
                      rotations, opacities=None, normalize: bool = True, eps: float = 1e-8,
                      support_fade: bool | float = False, sigma_cutoff: float = 3.0):
    """Normalized or additive render with affine colors c_i + G_i0 x + G_i1 y in each Gaussian's
    scaled, rotated frame (render.py:110-123); ``scales`` (N,2) are linear scales, ``rotations``
    (N,) angles, ``color_grads`` (N,2,3)."""
    _check_field(means, conics, colors, radii, opacities)
    if scales is None or rotations is None:
        raise ValueError("affine color rendering requires scales and rotations")
    n = means.shape[0]
    if color_grads.shape != (n, 2, 3) or scales.shape != (n, 2) or rotations.shape != (n,):
        raise ValueError("color_grads must be (N, 2, 3), scales (N, 2) and rotations (N,)")
    if any(t.dtype != torch.float32 or t.device != means.device
           for t in (color_grads, scales, rotations)):
        raise RuntimeError("color_grads, scales and rotations must be float32 on the means' device")
    _check_image(H, W, DEFAULT_TILE_SIZE)
    if normalize:
        eps = _validate_normalization_eps(eps)
    if opacities is None:
        opacities = means.new_empty(0)
    return _AffineRenderHip.apply(means, conics, colors, radii, opacities, color_grads, scales,
                                  rotations, int(H), int(W), bool(normalize), float(eps),
                                  _fade_alpha(support_fade), float(sigma_cutoff))


RENDER_TILE_SIZES = (16, 32)


def render_tile_size(tile_size: int) -> int:
    """Tile edge the render kernels run for a requested ``tile_size`` (any edge with
    ``tile_size^2 <= 1024``, the reference's contract, render_ext.cpp:162-163).

    The kernels are built for 16- and 32-pixel tiles.  Every pixel sums exactly the Gaussians
    whose support rectangle holds it, in ascending id, whatever the tiling, so the image does
    not depend on the tile edge; another requested edge renders at 16, the reference default.
    """
    ts = int(tile_size)
    if ts <= 0 or ts * ts > 1024:
        raise ValueError("tile_size must be positive with tile_size^2 <= 1024")
    return ts if ts in RENDER_TILE_SIZES else DEFAULT_TILE_SIZE


def _check_exact_args(H: int, W: int, tiled: bool, tile_size: int, backward_variant: str,
                      tile_index_backend: str) -> int:
    """Argument checks of ``render_cuda_exact`` after its device and dtype checks, in the
    reference's order (cuda_render.py:278-288); returns the tile edge the kernels run."""
    if backward_variant not in ("baseline", "block_reduce"):
        raise ValueError(
            f"backward_variant must be baseline or block_reduce, got {backward_variant!r}"
        )
    if tiled and backward_variant != "baseline":
        raise ValueError("block_reduce backward is experimental and untiled-only")
    if tile_index_backend not in ("hip", "cuda", "torch"):
        raise ValueError(
            f"tile_index_backend must be hip, cuda or torch, got {tile_index_backend!r}")
    _check_image(H, W, tile_size)
    return render_tile_size(tile_size) if tiled else DEFAULT_TILE_SIZE


def render_hip_exact(means, conics, colors, radii, H: int, W: int, opacities=None,
                     normalize: bool = True, eps: float = 1e-8, tiled: bool = False,
                     tile_size: int = DEFAULT_TILE_SIZE, support_fade: bool | float = False,
                     sigma_cutoff: float = 3.0, backward_variant: str = "baseline",
                     tile_index_backend: str = "hip", tile_ellipse_cull: bool = True,
                     recompute_normalizer: bool = True):
    """Exact normalized/additive render on the GPU (signature and checks of
    ``render_cuda_exact``, cuda_render.py:256-294, plus ``recompute_normalizer``).

    ``support_fade`` is the reference's bool (full fade) or a fade strength in [0, 1]
    (render.py:94-107).

    The HIP path is always tiled and has one deterministic backward.  ``tiled`` keeps the
    reference's meaning for argument checking: ``tiled=True`` with ``backward_variant=
    "block_reduce"`` raises as in cuda_render.py:283-284; with ``tiled=False`` (the reference
    default, whose untiled kernels compute the same sums) ``tile_size`` is ignored as the
    reference ignores it.  A tiled ``tile_size`` may be any edge with ``tile_size^2 <= 1024``
    (render_ext.cpp:162-163); see ``render_tile_size`` for the edge the kernels run.
    ``recompute_normalizer`` (default) keeps only the image for the backward and recomputes the
    per-pixel normalizer there; ``False`` saves it too, as the reference does.  ``tile_index_backend`` accepts
    ``"hip"`` and the reference's ``"cuda"`` and ``"torch"`` (cuda_render.py:285-288); all three
    bin on the GPU, since the torch builder's lists hold the same Gaussians per tile
    (test_render.py:401-409) and the ascending order is one of its orders.
    """
    _check_field(means, conics, colors, radii, opacities)
    ts = _check_exact_args(H, W, tiled, tile_size, backward_variant, tile_index_backend)
    if normalize:
        eps = _validate_normalization_eps(eps)
    if opacities is None:
        opacities = means.new_empty(0)
    return _ExactRenderHip.apply(means, conics, colors, radii, opacities, int(H), int(W),
                                 bool(normalize), float(eps), ts,
                                 _fade_alpha(support_fade), float(sigma_cutoff),
                                 bool(tile_ellipse_cull), bool(recompute_normalizer))


@torch.no_grad()
def render_hip_exact_with_coverage(means, conics, colors, radii, H: int, W: int,
                                   opacities=None, eps: float = 1e-8,
                                   support_fade: bool | float = False,
                                   sigma_cutoff: float = 3.0):
    """Normalized RGB and raw coverage ``den = sum o G`` (no eps) from one forward, no autograd
    (``render_cuda_exact_with_coverage``, cuda_render.py:229-253)."""
    if not means.is_cuda:
        raise RuntimeError("joint HIP rendering requires CUDA tensors")
    if any(t.dtype != torch.float32 for t in (means, conics, colors)):
        raise RuntimeError("joint HIP rendering requires float32 tensors")
    eps = _validate_normalization_eps(eps)
    _check_field(means, conics, colors, radii, opacities)
    if opacities is None:
        opacities = means.new_empty(0)
    means = means.contiguous()
    conics = conics.contiguous()
    colors = colors.contiguous()
    radii = _radii64(radii)
    opacities = opacities.contiguous()
    gids, offsets = build_tile_index(means, conics, radii, int(H), int(W), DEFAULT_TILE_SIZE,
                                     False, float(sigma_cutoff))
    den = torch.empty(int(H), int(W), dtype=torch.float32, device=means.device)
    out, den = forward_tiled(means, conics, colors, radii, opacities, gids, offsets, int(H),
                             int(W), True, eps, DEFAULT_TILE_SIZE, _fade_alpha(support_fade),
                             float(sigma_cutoff), den=den)
    return out, den


@torch.no_grad()
def gaussian_activity(means, conics, radii, H: int, W: int, chunk: int = 4096,
                      support_fade: bool = False, sigma_cutoff: float = 3.0,
                      support_fade_alpha: float | None = None) -> torch.Tensor:
    """Each Gaussian's summed unnormalized, opacity-free weight over the image (the reference's
    pruning statistic, render.py:417-442), from one HIP launch.  ``chunk`` bounds the reference
    loop's memory and has no effect here.  Each sum is formed in a fixed order."""
    del chunk
    if not means.is_cuda:
        raise RuntimeError(_DEVICE_MESSAGE)
    if means.dtype != torch.float32 or conics.dtype != torch.float32:
        raise RuntimeError("gaussian_activity requires float32 means and conics")
    fade_alpha = _resolve_support_fade_alpha(support_fade, support_fade_alpha)
    means = means.detach().contiguous()
    conics = conics.detach().contiguous()
    radii = _radii64(radii)
    n = means.shape[0]
    activity = torch.empty(n, dtype=torch.float32, device=means.device)
    _native.check(_native.load().ssplat_gaussian_activity(
        _ptr(means), _ptr(conics), _ptr(radii), n, int(H), int(W), float(fade_alpha),
        float(sigma_cutoff), _ptr(activity), _stream(means.device)), "ssplat_gaussian_activity")
    return activity


def raw_weight_map(means, conics, radii, H: int, W: int, opacities=None,
                   support_fade: bool | float = False, sigma_cutoff: float = 3.0,
                   detach_opacities: bool = False) -> torch.Tensor:
    """Differentiable (H*W,) raw coverage sum_i o_i G_i(p), the accumulation of the reference's
    ``_raw_weight_map_field`` (fit.py:338-367), on the HIP kernels: an additive render of unit
    colors, whose first channel is the coverage bit for bit (w * 1 is exact).  Gradients reach
    the means, conics and (unless ``detach_opacities``) the opacities."""
    _check_field(means, conics, conics, radii, opacities, need_colors=False)
    if opacities is not None and detach_opacities:
        opacities = opacities.detach()
    ones = torch.ones(means.shape[0], 3, dtype=torch.float32, device=means.device)
    out = render_hip_exact(means, conics, ones, radii, H, W, opacities=opacities,
                           normalize=False, support_fade=support_fade, sigma_cutoff=sigma_cutoff)
    return out[..., 0].reshape(-1)


def render(means, conics, colors, radii, H: int, W: int, chunk: int = 4096, opacities=None,
           support_fade: bool = False, sigma_cutoff: float = 3.0, color_grads=None,
           scales=None, rotations=None, support_fade_alpha: float | None = None,
           checkpoint_chunks: bool = False, normalization_eps: float = _EPS):
    """The normalized weighted-sum renderer under the reference's name and arguments
    (render.py:238-249), on the HIP kernels (affine colors included)."""
    return render_field(means, conics, colors, radii, H, W, chunk, "normalized", opacities,
                        scales, rotations, support_fade, sigma_cutoff, color_grads,
                        support_fade_alpha, checkpoint_chunks, normalization_eps)


def render_additive(means, conics, colors, radii, H: int, W: int, chunk: int = 4096,
                    opacities=None, support_fade: bool = False, sigma_cutoff: float = 3.0,
                    color_grads=None, scales=None, rotations=None,
                    support_fade_alpha: float | None = None,
                    checkpoint_chunks: bool = False):
    """The additive renderer under the reference's name and arguments (render.py:252-263)."""
    return render_field(means, conics, colors, radii, H, W, chunk, "additive", opacities,
                        scales, rotations, support_fade, sigma_cutoff, color_grads,
                        support_fade_alpha, checkpoint_chunks)


# The reference's CUDA entry names (cuda_render.py:229-294) for callers that import them.
render_cuda_exact = render_hip_exact
render_cuda_exact_with_coverage = render_hip_exact_with_coverage


def render_field(means, conics, colors, radii, H: int, W: int,
                 chunk: int = 4096, mode: str = "normalized", opacities=None,
                 scales=None, rotations=None, support_fade: bool = False,
                 sigma_cutoff: float = 3.0, color_grads=None,
                 support_fade_alpha: float | None = None,
                 checkpoint_chunks: bool = False,
                 normalization_eps: float = _EPS):
    """Mode dispatcher with the reference's signature (render.py:319-414).

    ``chunk`` and ``checkpoint_chunks`` bound the reference renderer's memory and have no effect
    here.  Fractional fade strengths run in the HIP kernels (the reference falls back to torch for
    them, render.py:341-358), and so do affine colors (``color_grads`` with ``scales`` and
    ``rotations``), which the reference renders with its torch loop for the normalized modes and
    ``additive`` (render.py:359-374); the fixed-fade additive kernel modes reject them as the
    reference does.
    """
    del chunk, checkpoint_chunks
    fade_alpha = _resolve_support_fade_alpha(support_fade, support_fade_alpha)
    if mode in ("gsplat", "cuda_gsplat"):
        if scales is None or rotations is None:
            raise ValueError("renderer='gsplat' requires scales and rotations")
        raise ValueError(
            "renderer='gsplat' (the GaussianImage++ alpha/sum comparator) is not part of the "
            "HIP renderer"
        )
    if mode not in _NORMALIZED_MODES and mode not in _ADDITIVE_MODES:
        raise ValueError(
            f"unknown renderer {mode!r}; expected normalized, additive, hip, hip_additive, cuda, "
            "cuda_block_reduce, cuda_additive, cuda_tiled, or cuda_tiled_additive"
        )
    if not means.is_cuda:
        raise RuntimeError(_DEVICE_MESSAGE)
    if color_grads is not None:
        if mode in _NORMALIZED_MODES:
            eps = _validate_normalization_eps(normalization_eps)
            return render_hip_affine(means, conics, colors, radii, H, W, color_grads, scales,
                                     rotations, opacities, True, eps, fade_alpha, sigma_cutoff)
        if mode == "additive" or fade_alpha not in (0.0, 1.0):
            return render_hip_affine(means, conics, colors, radii, H, W, color_grads, scales,
                                     rotations, opacities, False, _EPS, fade_alpha, sigma_cutoff)
        raise ValueError(
            "affine color rendering is currently supported by the reference normalized/additive "
            f"renderers only; got renderer={mode!r}"
        )
    if mode in _NORMALIZED_MODES:
        eps = _validate_normalization_eps(normalization_eps)
        return render_hip_exact(means, conics, colors, radii, H, W, opacities=opacities,
                                normalize=True, eps=eps, support_fade=fade_alpha,
                                sigma_cutoff=sigma_cutoff)
    return render_hip_exact(means, conics, colors, radii, H, W, opacities=opacities,
                            normalize=False, eps=_EPS, support_fade=fade_alpha,
                            sigma_cutoff=sigma_cutoff)
