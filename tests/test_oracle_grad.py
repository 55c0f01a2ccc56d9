"""Pins oracle.exact64.backward: float64 autograd of the torch restatement of the reference and
float64 central finite differences with the support rectangles frozen."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import exact64
from oracle import reference_render as R
from oracle.tile_index import support_rects
from tests.helpers import field_tensors, random_field, to_np


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True, 0.6])
@pytest.mark.parametrize("opacity", [False, True])
def test_analytic_backward_matches_float64_autograd(normalize, fade, opacity):
    H, W = 23, 29
    f = random_field(31, 60, H, W, scale_lo=0.7, scale_hi=4.0, opacity=opacity)
    means, conics, colors, radii, op = field_tensors(f, "cpu")
    G = np.random.default_rng(1).random((H, W, 3))
    leaves = [means.double().requires_grad_(), conics.double().requires_grad_(),
              colors.double().requires_grad_()]
    opd = op.double().requires_grad_() if op is not None else None
    alpha = None if isinstance(fade, bool) else fade
    img = R.accumulate(*leaves, radii, H, W, 1, opd, normalize, bool(fade), 3.0, alpha)
    (img * torch.from_numpy(G)).sum().backward()
    got = exact64.backward(G, to_np(means), to_np(conics), to_np(colors), to_np(radii), to_np(op),
                           H, W, normalize, 1e-8, fade, 3.0)
    for t, g in zip(leaves, got[:3]):
        assert np.allclose(g, t.grad.numpy(), rtol=1e-9, atol=1e-12)
    if opacity:
        assert np.allclose(got[3], opd.grad.numpy(), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True])
def test_analytic_backward_matches_central_differences(normalize, fade):
    H, W = 19, 17
    f = random_field(8, 12, H, W, scale_lo=0.8, scale_hi=3.0, opacity=True, margin=0.0)
    means, conics, colors, radii, op = (to_np(t).astype(np.float64) if t is not None else None
                                        for t in field_tensors(f, "cpu"))
    radii = radii.astype(np.int64)
    rects = support_rects(means.astype(np.float32), radii, H, W)
    G = np.random.default_rng(2).random((H, W, 3))

    def loss(m, q, c, o):
        out, _ = exact64.forward(m, q, c, radii, o, H, W, normalize, 1e-8, fade, 3.0,
                                 rects=rects)
        return float((out * G).sum())

    # exact64 casts inputs to float32 first; differentiate a float64 copy of the same path.
    def loss64(m, q, c, o):
        return _loss_f64(m, q, c, o, rects, G, H, W, normalize, fade)

    analytic = exact64.backward(G, means, conics, colors, radii, op, H, W, normalize, 1e-8, fade,
                                3.0, rects=rects)
    params = [means, conics, colors, op]
    h = 1e-6
    for p_idx, p in enumerate(params):
        for flat in range(0, p.size, max(1, p.size // 12)):
            plus = [x.copy() for x in params]
            minus = [x.copy() for x in params]
            plus[p_idx].flat[flat] += h
            minus[p_idx].flat[flat] -= h
            fd = (loss64(*plus) - loss64(*minus)) / (2 * h)
            assert abs(fd - analytic[p_idx].flat[flat]) <= 1e-6 * max(1.0, abs(fd)), (p_idx, flat)
    assert np.isfinite(loss(means, conics, colors, op))


def _loss_f64(m, q, c, o, rects, G, H, W, normalize, fade):
    """Float64 forward with no float32 rounding of the parameters (for finite differences)."""
    x0, x1, y0, y1, valid = rects
    tail = np.exp(-4.5)
    num = np.zeros((H, W, 3))
    den = np.zeros((H, W))
    for i in np.nonzero(valid)[0]:
        ys, xs = np.mgrid[y0[i]:y1[i], x0[i]:x1[i]]
        dx = xs - m[i, 0]
        dy = ys - m[i, 1]
        quad = q[i, 0] * dx * dx + 2 * q[i, 1] * dx * dy + q[i, 2] * dy * dy
        w = np.exp(-0.5 * quad)
        if fade:
            w = np.maximum(w - tail, 0.0)
        w = w * o[i]
        num[y0[i]:y1[i], x0[i]:x1[i]] += w[..., None] * c[i]
        den[y0[i]:y1[i], x0[i]:x1[i]] += w
    out = num / (den[..., None] + 1e-8) if normalize else num
    return float((out * G).sum())


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True])
def test_float64_params_forward_matches_per_gaussian_loop(normalize, fade):
    """exact64.forward(float64_params=True) keeps parameters that float32 cannot hold: it equals
    the per-Gaussian float64 loop above on such parameters, and differs from the default
    (float32-rounded) forward."""
    H, W = 21, 18
    f = random_field(14, 15, H, W, scale_lo=0.8, scale_hi=3.0, opacity=True, margin=0.0)
    means, conics, colors, radii, op = (to_np(t).astype(np.float64) if t is not None else None
                                        for t in field_tensors(f, "cpu"))
    radii = radii.astype(np.int64)
    rects = support_rects(means.astype(np.float32), radii, H, W)
    jitter = np.random.default_rng(5)
    params = [p + jitter.uniform(-1e-9, 1e-9, p.shape) for p in (means, conics, colors, op)]
    G = jitter.random((H, W, 3))
    out, _ = exact64.forward(*params[:3], radii, params[3], H, W, normalize, 1e-8, fade, 3.0,
                             rects=rects, float64_params=True)
    ref = _loss_f64(*params, rects, G, H, W, normalize, fade)
    assert abs(float((out * G).sum()) - ref) <= 1e-12 * abs(ref)
    rounded, _ = exact64.forward(*params[:3], radii, params[3], H, W, normalize, 1e-8, fade, 3.0,
                                 rects=rects)
    assert not np.array_equal(out, rounded)
    with pytest.raises(ValueError):
        exact64.forward(*params[:3], radii, params[3], H, W, float64_params=True)


U32 = 2.0 ** -24  # float32 unit roundoff


def _f32_render_and_grads(f, G, H, W, normalize, fade):
    """The reference's own float32 evaluation (the torch restatement in float32, autograd)."""
    means, conics, colors, radii, op = field_tensors(f, "cpu")
    leaves = [t.clone().requires_grad_() for t in (means, conics, colors)]
    opl = op.clone().requires_grad_() if op is not None else None
    img = R.accumulate(*leaves, radii, H, W, 1, opl, normalize, bool(fade), 3.0)
    (img * torch.from_numpy(G.astype(np.float32))).sum().backward()
    grads = [t.grad.double().numpy() for t in leaves]
    grads.append(opl.grad.double().numpy() if opl is not None else None)
    return img.detach().double().numpy(), grads, (means, conics, colors, radii, op)


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True])
@pytest.mark.parametrize("uniform_colors", [False, True])
def test_conditioning_bounds_the_float32_error(normalize, fade, uniform_colors):
    """forward_conditioning and backward_conditioning bound the error of the reference's own
    float32 evaluation against float64: |f32 - f64| <= 64 u sens for every pixel and gradient.
    Needle-like Gaussians (aspect up to 40) make the quadratic's terms cancel; equal colors make
    the normalized per-pixel chain dw = sum G (c - out) / D cancel to zero in exact arithmetic,
    so a bound without those cancellation terms fails here."""
    H, W = 21, 27
    f = random_field(5, 40, H, W, opacity=True, aniso=(0.4, 1.2, 4.0, 16.0))
    if uniform_colors:
        f["colors"][:] = np.float32([0.3, 0.6, 0.9])
    G = np.random.default_rng(9).standard_normal((H, W, 3))
    img32, g32, t32 = _f32_render_and_grads(f, G, H, W, normalize, fade)
    args = tuple(to_np(t) for t in t32)
    img64, _ = exact64.forward(*args, H, W, normalize, 1e-8, fade, 3.0)
    sens_f = exact64.forward_conditioning(*args, H, W, normalize, 1e-8, fade, 3.0)
    assert np.all(np.abs(img32 - img64) <= 64 * U32 * sens_f + 1e-30)
    g64 = exact64.backward(G, *args, H, W, normalize, 1e-8, fade, 3.0)
    sens_b = exact64.backward_conditioning(G, *args, H, W, normalize, 1e-8, fade, 3.0)
    for name, a, b, s in zip(("means", "conics", "colors", "opacities"), g32, g64, sens_b):
        err = np.abs(a - b)
        assert np.all(err <= 64 * U32 * s + 1e-30), (name, float(np.max(err / (U32 * s + 1e-300))))


@pytest.mark.parametrize("normalize", [True, False])
def test_conditioning_bounds_the_float32_error_of_needles(normalize):
    """Extreme needles (minor scale 0.08, major up to 24, any angle): the quadratic's terms are
    thousands of times q itself, so raw's float32 rounding is relative to qabs, not to q; a bound
    without the qabs term fails here."""
    H, W = 25, 25
    f = random_field(6, 12, H, W, opacity=True, aniso=(0.08, 0.08, 8.0, 24.0))
    G = np.random.default_rng(10).standard_normal((H, W, 3))
    img32, g32, t32 = _f32_render_and_grads(f, G, H, W, normalize, False)
    args = tuple(to_np(t) for t in t32)
    img64, _ = exact64.forward(*args, H, W, normalize, 1e-8, False, 3.0)
    sens_f = exact64.forward_conditioning(*args, H, W, normalize, 1e-8, False, 3.0)
    assert np.all(np.abs(img32 - img64) <= 64 * U32 * sens_f + 1e-30)
    g64 = exact64.backward(G, *args, H, W, normalize, 1e-8, False, 3.0)
    sens_b = exact64.backward_conditioning(G, *args, H, W, normalize, 1e-8, False, 3.0)
    for name, a, b, s in zip(("means", "conics", "colors", "opacities"), g32, g64, sens_b):
        err = np.abs(a - b)
        assert np.all(err <= 64 * U32 * s + 1e-30), (name, float(np.max(err / (U32 * s + 1e-300))))


def test_backward_conditioning_is_close_to_the_gradient_where_well_conditioned():
    """Round Gaussians, additive, all-positive G: no cancellation anywhere, so the sensitivity is
    a small multiple of the gradient's own magnitude (the bound is not vacuous)."""
    H, W = 17, 19
    f = random_field(3, 20, H, W, scale_lo=1.0, scale_hi=2.0, opacity=True)
    f["angles"][:] = 0.0
    f["scales"][:, 1] = f["scales"][:, 0]
    G = np.random.default_rng(4).random((H, W, 3))
    means, conics, colors, radii, op = (to_np(t) for t in field_tensors(f, "cpu"))
    g64 = exact64.backward(G, means, conics, colors, radii, op, H, W, False, 1e-8, False, 3.0)
    s = exact64.backward_conditioning(G, means, conics, colors, radii, op, H, W, False, 1e-8,
                                      False, 3.0)
    # colors: sum G w, sensitivity sum G w (1 + qabs/2) with qabs <= 9 inside the 3-sigma box.
    assert np.all(s[2] >= np.abs(g64[2]) * (1 - 1e-12))
    assert np.all(s[2] <= 5.5 * np.abs(g64[2]) + 1e-30)
    assert np.all(s[3] >= np.abs(g64[3]) * (1 - 1e-12))
