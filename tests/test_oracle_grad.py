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
