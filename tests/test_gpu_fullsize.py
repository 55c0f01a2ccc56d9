"""GPU parity at the full sizes of the benchmark workloads.

* 1024 x 1024 / 50,000 Gaussians (normalized and additive): the whole image against the float64
  oracle, the tile lists bit-exact, and the gradients of 64 random Gaussians against float64
  central finite differences and the oracle's analytic backward.
* 2048 x 2048 / 200,000 and 4096 x 4096 / 1,000,000 Gaussians (heavy-tailed scales, aspect up to
  20; the backward recomputes the normalizer): bit-exact tile lists at both tile sizes, image
  windows and per-Gaussian gradients against the float64 oracle on the pixels they cover,
  bitwise-repeatable forwards, and the stored-normalizer backward agreeing with the recomputing
  one.

A pixel sums only the Gaussians whose support rectangle holds it, so the oracle evaluated on a
window (``tests.helpers.window_field``) gives the image there, and on a Gaussian's own rectangle
gives that Gaussian's whole gradient; the float64 oracle never has to render a full image of the
two larger workloads.
"""
from __future__ import annotations

import functools

import numpy as np
import pytest
import torch

from oracle import exact64
from oracle import tile_index as TI
from structsplat_amd import render as HR
from structsplat_amd import synth
from tests.helpers import (field_tensors, local_gradient, pixel_bar_err, pixel_rel_err,
                           sample_gaussians, scaled_err, window_field)

pytestmark = pytest.mark.gpu

WORKLOADS = {
    "single": (synth.single_field, 1024, 1024),
    "metric": (synth.metric_field, 2048, 2048),
    "stress": (synth.stress_field, 4096, 4096),
}
GRAD_NAMES = ("means", "conics", "colors", "opacities")


def _case(**kw):
    print("case:", ", ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@functools.lru_cache(maxsize=1)
def _host(name):
    """(means, conics, colors, radii, opacities) as CPU tensors, and H, W."""
    fn, H, W = WORKLOADS[name]
    return tuple(field_tensors(fn(0), "cpu")), H, W


def _device(name, dev, grad=False):
    fields, H, W = _host(name)
    out = [t.to(dev) for t in fields]
    if grad:
        for k in (0, 1, 2, 4):
            out[k].requires_grad_(True)
    return out, H, W


def _np(t):
    return t.detach().cpu().numpy()


def _grad_image(H, W, dev, seed=3):
    return torch.rand(H, W, 3, device=dev, generator=torch.Generator(device=dev).manual_seed(seed))


@pytest.mark.parametrize("ts", [16, 32])
@pytest.mark.parametrize("name", ["single", "metric", "stress"])
def test_full_size_tile_lists_bit_exact(cuda, name, ts):
    _case(test="full-size tile lists", workload=name, tile=ts)
    (m, q, c, r, o), H, W = _device(name, cuda)
    gids, offsets = HR.build_tile_index(m, q, r, H, W, ts, True, 3.0)
    torch.cuda.synchronize()
    ref_g, ref_o = TI.build_tile_index(_np(m), _np(q), _np(r), H, W, ts, True, 3.0)
    assert np.array_equal(_np(offsets).astype(np.int64), ref_o)
    assert np.array_equal(_np(gids), ref_g)


@pytest.mark.parametrize("normalize", [True, False])
def test_single_workload_whole_image_matches_fp64(cuda, normalize):
    _case(test="1024x1024/50k whole image", normalize=normalize)
    (m, q, c, r, o), H, W = _device("single", cuda)
    out = HR.render_hip_exact(m, q, c, r, H, W, opacities=o, normalize=normalize)
    torch.cuda.synchronize()
    args = (_np(m), _np(q), _np(c), _np(r), _np(o), H, W, normalize, 1e-8, False, 3.0)
    ref, _ = exact64.forward(*args)
    sens = exact64.forward_conditioning(*args)
    got = _np(out)
    assert np.isfinite(got).all()
    err = scaled_err(got, ref)
    bar = pixel_bar_err(got, ref, sens)
    print("scaled err", err, "per-pixel rel err (floor 1e-6)", pixel_rel_err(got, ref),
          "per-pixel bar", bar)
    assert err <= 1e-5
    assert pixel_rel_err(got, ref) <= 1e-5  # per pixel, floor 1e-6 of the image scale
    assert bar <= 1.0


@pytest.mark.parametrize("normalize", [True, False])
def test_single_workload_gradients_match_central_differences(cuda, normalize):
    """64 random Gaussians of the 1024x1024/50k field: the HIP gradients against float64 central
    differences of the oracle (support rectangles held fixed), which in turn agree with the
    oracle's analytic backward."""
    _case(test="1024x1024/50k gradients vs central differences", normalize=normalize)
    leaves, H, W = _device("single", cuda, grad=True)
    m, q, c, r, o = leaves
    out = HR.render_hip_exact(m, q, c, r, H, W, opacities=o, normalize=normalize)
    G = _grad_image(H, W, cuda)
    (out * G).sum().backward()
    torch.cuda.synchronize()
    host = _host("single")[0]
    Gn = _np(G).astype(np.float64)
    gpu = [_np(t.grad) for t in (m, q, c, o)]
    picks = sample_gaussians(host, H, W, 64, seed=11)
    fd = {k: [] for k in GRAD_NAMES}
    an = {k: [] for k in GRAD_NAMES}
    got = {k: [] for k in GRAD_NAMES}
    steps = (1e-5, 1e-6, 1e-5, 1e-5)  # means, conics (values ~1e-2), colors, opacities
    for i in picks:
        analytic, (k, Gw, h, w, mw, qw, cw, rw, ow, rects) = local_gradient(host, H, W, Gn, i,
                                                                              normalize)
        params = [mw, qw, cw, ow]

        def loss(ps):
            img, _ = exact64.forward(ps[0], ps[1], ps[2], rw, ps[3], h, w, normalize, 1e-8,
                                     False, 3.0, rects=rects, float64_params=True)
            return float((img * Gw).sum())

        for slot, name in enumerate(GRAD_NAMES):
            row = params[slot][k]
            vals = []
            for j in range(np.size(row)):
                plus = [p.copy() for p in params]
                minus = [p.copy() for p in params]
                if slot == 3:
                    plus[3][k] += steps[3]
                    minus[3][k] -= steps[3]
                else:
                    plus[slot][k, j] += steps[slot]
                    minus[slot][k, j] -= steps[slot]
                vals.append((loss(plus) - loss(minus)) / (2 * steps[slot]))
            fd[name].append(vals)
            an[name].append(np.atleast_1d(analytic[slot]))
            got[name].append(np.atleast_1d(gpu[slot][i]))
    for name in GRAD_NAMES:
        f, a, g = (np.array(d[name], dtype=np.float64) for d in (fd, an, got))
        e_fd = scaled_err(a, f)
        e_gpu = scaled_err(g, f)
        print(name, "oracle vs differences", e_fd, "HIP vs differences", e_gpu)
        assert e_fd <= 1e-6, name
        assert e_gpu <= 1e-4, name


@pytest.mark.parametrize("name", ["metric", "stress"])
def test_full_size_windows_and_gradients_match_fp64(cuda, name):
    """Image windows and 48 random Gaussians' gradients (normalizer recomputed in the backward)
    against the float64 oracle on the pixels they cover."""
    _case(test="full-size windows and gradients", workload=name)
    leaves, H, W = _device(name, cuda, grad=True)
    m, q, c, r, o = leaves
    out = HR.render_hip_exact(m, q, c, r, H, W, opacities=o, recompute_normalizer=True)
    G = _grad_image(H, W, cuda, seed=5)
    (out * G).sum().backward()
    torch.cuda.synchronize()
    host = _host(name)[0]
    hm, hq, hc, hr, ho = (_np(t) for t in host)
    img = _np(out)
    rng = np.random.default_rng(7)
    for _ in range(12):
        x0 = int(rng.integers(0, W - 48))
        y0 = int(rng.integers(0, H - 48))
        ids, mw, qw, cw, rw, ow, rects = window_field(hm, hq, hc, hr, ho, H, W, x0, x0 + 48,
                                                      y0, y0 + 48)
        wargs = (mw, qw, cw, rw, ow, 48, 48, True, 1e-8, False, 3.0)
        ref, _ = exact64.forward(*wargs, rects=rects, float64_params=True)
        sens = exact64.forward_conditioning(*wargs, rects=rects, float64_params=True)
        win = img[y0:y0 + 48, x0:x0 + 48]
        err, bar = scaled_err(win, ref), pixel_bar_err(win, ref, sens)
        print("window", x0, y0, "scaled", err, "per-pixel rel", pixel_rel_err(win, ref), "bar", bar)
        assert err <= 1e-5 and bar <= 1.0, (x0, y0, err, bar)
        assert pixel_rel_err(win, ref) <= 1e-5, (x0, y0)  # floor 1e-6 of the window's scale
    Gn = _np(G).astype(np.float64)
    gpu = [_np(t.grad) for t in (m, q, c, o)]
    an = {k: [] for k in GRAD_NAMES}
    got = {k: [] for k in GRAD_NAMES}
    for i in sample_gaussians(host, H, W, 48, seed=13):
        analytic, _ = local_gradient(host, H, W, Gn, i, True)
        for slot, key in enumerate(GRAD_NAMES):
            an[key].append(np.atleast_1d(analytic[slot]))
            got[key].append(np.atleast_1d(gpu[slot][i]))
    for key in GRAD_NAMES:
        err = scaled_err(np.array(got[key]), np.array(an[key]))
        print(key, "scaled err", err)
        assert err <= 1e-4, key


@pytest.mark.parametrize("name", ["metric", "stress"])
def test_full_size_repeatable_and_normalizer_modes_agree(cuda, name):
    """Two forwards are bitwise equal; the backward that reads a stored normalizer and the one
    that recomputes it give the same gradients up to atomic summation order."""
    _case(test="full-size determinism and normalizer modes", workload=name)
    (m, q, c, r, o), H, W = _device(name, cuda)
    gids, offsets = HR.build_tile_index(m, q, r, H, W, 16, True, 3.0)
    den = torch.empty(H, W, device=cuda)
    out, _ = HR.forward_tiled(m, q, c, r, o, gids, offsets, H, W, True, 1e-8, 16, False, 3.0,
                              den=den)
    again, _ = HR.forward_tiled(m, q, c, r, o, gids, offsets, H, W, True, 1e-8, 16, False, 3.0)
    torch.cuda.synchronize()
    assert torch.equal(out, again)
    G = _grad_image(H, W, cuda, seed=9)
    saved = HR.backward_tiled(G, m, q, c, r, o, den, out, gids, offsets, H, W, True, 1e-8, 16,
                              False, 3.0)
    recomputed = HR.backward_tiled(G, m, q, c, r, o, None, out, gids, offsets, H, W, True, 1e-8,
                                   16, False, 3.0)
    torch.cuda.synchronize()
    for key, a, b in zip(GRAD_NAMES, saved, recomputed):
        assert torch.isfinite(a).all() and torch.isfinite(b).all(), key
        err = scaled_err(_np(b), _np(a))
        print(key, "saved vs recomputed scaled err", err)
        assert err <= 1e-5, key
