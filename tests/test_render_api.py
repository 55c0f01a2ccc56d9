"""Host-side contract of render_field / render_hip_exact (no GPU needed): the reference's error
phrases and argument validation (test_render.py:136-164, test_normalization_epsilon.py:33-67)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from structsplat_amd import GaussianField, render_field, render_hip_exact
from structsplat_amd.render import _resolve_support_fade_alpha, _validate_normalization_eps


def _field():
    g = GaussianField.from_numpy(np.array([[3.0, 3.0]]), np.full((1, 2), 1.5), np.zeros(1),
                                 np.array([[1.0, 0, 0]]))
    return g.means, g.conics(), g.colors, g.radii(3.0)


@pytest.mark.parametrize("mode", ["hip", "normalized", "additive", "cuda", "cuda_tiled",
                                  "cuda_block_reduce", "cuda_additive", "cuda_tiled_additive"])
def test_cpu_tensors_are_rejected_with_the_reference_phrase(mode):
    with pytest.raises(RuntimeError, match="requires CUDA tensors"):
        render_field(*_field(), 10, 10, mode=mode)


def test_gsplat_requires_scales_and_rotations():
    with pytest.raises(ValueError, match="requires scales and rotations"):
        render_field(*_field(), 10, 10, mode="gsplat")


def test_unknown_mode_is_a_value_error():
    with pytest.raises(ValueError, match="unknown renderer"):
        render_field(*_field(), 10, 10, mode="nope")


@pytest.mark.parametrize("value", [0.0, -1e-12, float("inf"), float("nan")])
def test_normalization_eps_validation(value):
    with pytest.raises(ValueError, match="normalization_eps must be finite and > 0"):
        _validate_normalization_eps(value)
    assert _validate_normalization_eps(1e-12) == 1e-12


def test_render_hip_exact_rejects_cpu_and_non_float32():
    m, q, c, r = _field()
    with pytest.raises(RuntimeError, match="requires CUDA tensors"):
        render_hip_exact(m, q, c, r, 10, 10)


def test_fade_alpha_resolution_matches_reference():
    # render.py:94-98
    assert _resolve_support_fade_alpha(False) == 0.0
    assert _resolve_support_fade_alpha(True) == 1.0
    assert _resolve_support_fade_alpha(True, 0.25) == 0.25
    assert _resolve_support_fade_alpha(False, 7.0) == 1.0
    assert _resolve_support_fade_alpha(False, -1.0) == 0.0


def test_field_producer_dilation_validation():
    g = GaussianField.from_numpy(np.zeros((2, 2)), np.ones((2, 2)), np.zeros(2), np.zeros((2, 3)))
    with pytest.raises(ValueError, match="dilation must be >= 0"):
        g.conics(-1.0)
    with pytest.raises(ValueError, match="dilation must be >= 0"):
        g.radii(3.0, -0.5)
    assert g.radii(3.0).dtype == torch.int64


def test_affine_and_coverage_entry_points_reject_cpu_tensors():
    """Affine colors and the raw weight map run only on the GPU, like every mode."""
    from structsplat_amd import raw_weight_map, render_hip_affine

    m, q, c, r = _field()
    g = torch.zeros(1, 2, 3)
    with pytest.raises(RuntimeError, match="requires CUDA tensors"):
        render_field(m, q, c, r, 10, 10, color_grads=g, scales=torch.ones(1, 2),
                     rotations=torch.zeros(1))
    with pytest.raises(RuntimeError, match="requires CUDA tensors"):
        render_hip_affine(m, q, c, r, 10, 10, g, torch.ones(1, 2), torch.zeros(1))
    with pytest.raises(RuntimeError, match="requires CUDA tensors"):
        raw_weight_map(m, q, r, 10, 10)


def test_raw_weight_map_field_keeps_the_reference_signature():
    """fit.py:338-340: (field, cfg, H, W, support_fade_alpha=None, detach_opacities=False)."""
    import inspect

    from structsplat_amd.fit import raw_weight_map_field

    params = list(inspect.signature(raw_weight_map_field).parameters.items())
    assert [k for k, _ in params] == ["field", "cfg", "H", "W", "support_fade_alpha",
                                      "detach_opacities"]
    assert params[4][1].default is None and params[5][1].default is False


def test_gaussian_field_keeps_the_reference_constructor_and_helpers(tmp_path):
    """gaussians.py:15-29, :193-268, :324-358: argument order, affine coefficients as a
    parameter group, effective scales with the covariance filter, and save/load round trips."""
    import math

    g = GaussianField.from_numpy(np.array([[1.0, 2.0]]), np.array([[2.0, 3.0]]), np.zeros(1),
                                 np.ones((1, 3)), opacities=np.zeros(1),
                                 scale_max=np.full((1, 2), 9.0),
                                 background_mask=np.array([True]),
                                 filter_variance=np.array([0.25]))
    assert g.scale_max.shape == (1, 2) and g.background_mask.dtype == torch.bool
    a = g.with_affine_colors()
    assert a.color_grads.shape == (1, 2, 3) and not a.color_grads.any()
    groups = a.parameter_groups(1, 2, 3, 4, 5)
    assert len(groups) == 6 and groups[-1]["lr"] == 4  # color_grads at the color rate
    eff = g.effective_scales()
    assert torch.allclose(eff, torch.tensor([[math.sqrt(4.25), math.sqrt(9.25)]]))
    m = g.materialize_covariance_filter()
    assert m.filter_variance is None and torch.allclose(m.scales(), eff)
    a.save(str(tmp_path / "f.npz"))
    b = GaussianField.load(str(tmp_path / "f.npz"))
    for name in ("means", "log_scales", "rotations", "colors", "opacities", "scale_max",
                 "color_grads", "background_mask", "filter_variance"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
    positional = GaussianField(g.means, g.log_scales, g.rotations, g.colors, g.opacities,
                               g.scale_max)
    assert positional.scale_max is g.scale_max and positional.filter_variance is None


def test_fit_color_basis_and_l2_gradient():
    """fit() validates color_basis before any device work, and the in-place regularizer gradient
    equals autograd's gradient of color_grad_l2 * mean(color_grads^2) (fit.py:3256-3259)."""
    from structsplat_amd import fit as F
    from structsplat_amd.batch import fit_batch

    field = GaussianField.from_numpy(np.array([[3.0, 3.0]]), np.full((1, 2), 1.5), np.zeros(1),
                                     np.array([[1.0, 0, 0]]))
    target = torch.zeros(8, 8, 3)
    with pytest.raises(ValueError, match="color_basis must be constant or affine"):
        F.fit(field, target, F.FitConfig(iters=1, color_basis="quadratic"))
    with pytest.raises(ValueError, match="affine-color fields with fit"):
        fit_batch([field], [target], F.FitConfig(iters=1, color_basis="affine"))

    torch.manual_seed(0)
    cg = torch.randn(7, 2, 3, dtype=torch.float64, requires_grad=True)
    aff = GaussianField(torch.zeros(7, 2), torch.zeros(7, 2), torch.zeros(7), torch.zeros(7, 3),
                        color_grads=cg)
    cfg = F.FitConfig(color_grad_l2=0.3)
    (cfg.color_grad_l2 * (cg * cg).mean()).backward()
    expected = cg.grad.clone()
    cg.grad = torch.full_like(cg, 0.25)
    F.add_color_grad_l2(aff, cfg)
    assert torch.allclose(cg.grad - 0.25, expected, rtol=1e-12, atol=0)
    assert F.color_grad_l2_term(aff, cfg) == pytest.approx(0.3 * float((cg * cg).mean()),
                                                          rel=1e-12)


def test_exact_tile_size_any_edge_up_to_1024_pixels():
    """render_cuda_exact accepts any tile edge with tile_size^2 <= 1024 (render_ext.cpp:162-163);
    the kernels run 16 and 32 as asked and every other edge at 16 (the image does not depend on
    the tiling)."""
    from structsplat_amd.render import _check_exact_args, render_tile_size
    for ts in range(1, 33):
        expect = ts if ts in (16, 32) else 16
        assert render_tile_size(ts) == expect
        assert _check_exact_args(8, 8, True, ts, "baseline", "cuda") == expect
        # untiled (the reference default) ignores tile_size
        assert _check_exact_args(8, 8, False, ts, "baseline", "cuda") == 16
    for ts in (0, -4, 33, 64):
        with pytest.raises(ValueError, match=r"tile_size must be positive with tile_size\^2"):
            _check_exact_args(8, 8, True, ts, "baseline", "cuda")


def test_exact_block_reduce_is_untiled_only():
    """cuda_render.py:283-284: a tiled render with the block_reduce backward is a ValueError;
    untiled it is accepted, and unknown variants or backends are ValueErrors (:278-288)."""
    from structsplat_amd.render import _check_exact_args
    with pytest.raises(ValueError, match="block_reduce backward is experimental and untiled-only"):
        _check_exact_args(8, 8, True, 16, "block_reduce", "cuda")
    assert _check_exact_args(8, 8, False, 16, "block_reduce", "cuda") == 16
    with pytest.raises(ValueError, match="backward_variant must be baseline or block_reduce"):
        _check_exact_args(8, 8, False, 16, "atomic", "cuda")
    with pytest.raises(ValueError, match="tile_index_backend must be"):
        _check_exact_args(8, 8, True, 16, "baseline", "nope")


def test_concat_fields_carries_filter_variance_and_caps():
    """concat_fields keeps each field's covariance filter variance and scale caps with its rows;
    fields without them get the neutral values (variance 0, cap +inf)."""
    import math as _m

    import torch as _t

    from structsplat_amd.batch import concat_fields

    def f(n, fv=None, cap=None):
        g = GaussianField.from_numpy(np.zeros((n, 2)), np.ones((n, 2)), np.zeros(n),
                                     np.zeros((n, 3)))
        g.filter_variance = fv
        g.scale_max = cap
        return g

    a = f(2, fv=_t.tensor([0.5, 0.25]))
    b = f(3, cap=_t.full((3, 2), 2.0))
    c = f(1)
    big = concat_fields([a, b, c])
    assert big.filter_variance.tolist() == [0.5, 0.25, 0.0, 0.0, 0.0, 0.0]
    assert big.scale_max.shape == (6, 2)
    assert big.scale_max[2:5].eq(2.0).all() and _t.isinf(big.scale_max[[0, 1, 5]]).all()
    assert concat_fields([c, f(2)]).filter_variance is None
    assert concat_fields([c, f(2)]).scale_max is None
    # A neutral cap leaves the clamp's upper bound ln max(H, W).
    from structsplat_amd.fit import log_scale_upper
    hi = log_scale_upper(big, 40, 30)
    assert _t.is_tensor(hi) and hi.shape == (12,)
    assert hi[0].item() == _t.tensor(_m.log(40), dtype=_t.float32).item()
    assert hi[4].item() == _t.log(_t.tensor(2.0)).item()
    assert log_scale_upper(c, 40, 30) == _m.log(40)
