"""StructSplat's sorting-free 2D Gaussian rasterizer, rebuilt for MI355X (gfx950).

Public entry points mirror the reference's render path:

* :func:`render_field` - mode dispatcher (render.py:319-414).
* :func:`render_hip_exact` - the ``render_cuda_exact`` entry (cuda_render.py:256-294).
* ``structsplat_amd.render`` also holds ``render``, ``render_additive`` (render.py:238-263),
  ``render_cuda_exact`` and ``render_cuda_exact_with_coverage`` under the reference's names; the
  package keeps ``render`` for that module, as the reference's package does.
* :func:`build_tile_index` - per-tile Gaussian lists (render_ext.cpp:139-167).
* :func:`gaussian_activity` - per-Gaussian summed weight, the pruning statistic (render.py:417-442).
* :func:`raw_weight_map` - differentiable raw coverage sum_i o_i G_i (fit.py:338-367).
* :func:`render_hip_affine` - affine-color render (render.py:110-123).
* :class:`GaussianField` - field layout and conic/radii/opacity producer (gaussians.py).
* :mod:`structsplat_amd.codec` - SSPL1 decode and decode-and-render (codec.py:333-500).
* :mod:`structsplat_amd.shard` - one row band of a large image rendered from its own Gaussians.

The compute runs in hand-written HIP kernels in ``libssplat_hip.so`` behind a C ABI
(``include/ssplat.h``); build it with ``python -m structsplat_amd._build``.
"""
from .gaussians import GaussianField
from .render import (
    build_tile_index,
    build_tile_index_capacity,
    gaussian_activity,
    raw_weight_map,
    render_field,
    render_hip_affine,
    render_hip_exact,
    render_hip_exact_with_coverage,
)

__all__ = [
    "GaussianField",
    "build_tile_index",
    "build_tile_index_capacity",
    "gaussian_activity",
    "raw_weight_map",
    "render_field",
    "render_hip_affine",
    "render_hip_exact",
    "render_hip_exact_with_coverage",
]
