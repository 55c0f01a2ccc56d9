"""GPU parity of affine (gradient) colors: the HIP forward and backward against the float64
torch restatement of the reference's renderer with ``color_grads`` (render.py:110-123,
:180-263), differentiated by autograd."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import reference_render as R
from structsplat_amd import render as HR
from tests.helpers import field_tensors, random_field, scaled_err, to_np

pytestmark = pytest.mark.gpu


def _case(**kw):
    print("case:", ", ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _affine_field(seed, N, H, W, cuda, opacity=True, scale_hi=5.0, grad_scale=0.3):
    f = random_field(seed, N, H, W, scale_lo=0.7, scale_hi=scale_hi, opacity=opacity)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    rng = np.random.default_rng(seed + 1000)
    cg = torch.tensor(rng.normal(0.0, grad_scale, (N, 2, 3)), dtype=torch.float32, device=cuda)
    scales = torch.tensor(f["scales"], device=cuda)
    rot = torch.tensor(rng.uniform(-4.0, 7.0, N), dtype=torch.float32, device=cuda)
    return means, conics, colors, radii, op, cg, scales, rot


def _oracle(means, conics, colors, radii, op, cg, scales, rot, H, W, normalize, fade,
            grad_out=None):
    """float64 image (and with grad_out the gradients of sum(out * grad_out)) on the CPU."""
    ts = [None if t is None else t.detach().cpu().double().requires_grad_(True)
          for t in (means, conics, colors, op, cg, scales, rot)]
    m, q, c, o, g, s, r = ts
    kw = dict(opacities=o, color_grads=g, scales=s, rotations=r)
    if isinstance(fade, bool):
        kw["support_fade"] = fade
    else:
        kw.update(support_fade=True, support_fade_alpha=float(fade))
    if normalize:
        out = R.render(m, q, c, radii.cpu(), H, W, **kw)
    else:
        out = R.render_additive(m, q, c, radii.cpu(), H, W, **kw)
    if grad_out is None:
        return out.detach().numpy(), None
    (out * grad_out.cpu().double()).sum().backward()
    return out.detach().numpy(), [None if t is None else t.grad.numpy() for t in ts]


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True, 0.5])
@pytest.mark.parametrize("opacity", [False, True])
def test_affine_forward_matches_fp64(cuda, normalize, fade, opacity):
    _case(test="affine forward", normalize=normalize, fade=fade, opacity=opacity)
    H, W = 58, 83
    args = _affine_field(5, 300, H, W, cuda, opacity)
    out = HR.render_hip_affine(*args[:4], H, W, *args[5:], opacities=args[4],
                               normalize=normalize, support_fade=fade)
    torch.cuda.synchronize()
    ref, _ = _oracle(*args, H, W, normalize, fade)
    err = scaled_err(to_np(out), ref)
    print("scaled err", err)
    assert np.isfinite(to_np(out)).all() and err <= 1e-5


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True, 0.5])
@pytest.mark.parametrize("big", [False, True])
def test_affine_backward_matches_fp64(cuda, normalize, fade, big):
    """All seven gradients.  ``big`` adds Gaussians spanning more than 64 tiles, whose slots
    are summed by a whole wave instead of one lane."""
    _case(test="affine backward", normalize=normalize, fade=fade, big=big)
    H, W = (150, 170) if big else (61, 77)
    args = _affine_field(12, 160 if big else 250, H, W, cuda, True,
                         scale_hi=30.0 if big else 5.0)
    leaves = [args[i] for i in (0, 1, 2, 4, 5, 6, 7)]
    for t in leaves:
        t.requires_grad_(True)
    out = HR.render_hip_affine(*args[:4], H, W, *args[5:], opacities=args[4],
                               normalize=normalize, support_fade=fade)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(3))
    (out * g).sum().backward()
    torch.cuda.synchronize()
    if big:
        assert int(to_np(args[3]).max()) > 4 * 16
    _, ref = _oracle(*args, H, W, normalize, fade, g)
    names = ("means", "conics", "colors", "opacities", "color_grads", "scales", "rotations")
    worst = {}
    for name, leaf, exp in zip(names, leaves, ref):
        worst[name] = scaled_err(to_np(leaf.grad), exp)
    print("scaled errors", worst)
    for name, err in worst.items():
        assert err <= 1e-4, name


def test_affine_zero_color_grads_equal_plain_render(cuda):
    """G = 0 leaves every color c_i: the image equals the plain render's, and the gradients of
    means, conics, colors and opacities equal the plain backward's."""
    _case(test="affine zero gradients")
    H, W = 64, 80
    means, conics, colors, radii, op, cg, scales, rot = _affine_field(8, 280, H, W, cuda)
    cg = torch.zeros_like(cg)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(1))
    res = []
    for affine in (False, True):
        leaves = [t.detach().clone().requires_grad_(True) for t in (means, conics, colors, op)]
        if affine:
            out = HR.render_hip_affine(leaves[0], leaves[1], leaves[2], radii, H, W, cg, scales,
                                       rot, opacities=leaves[3])
        else:
            out = HR.render_hip_exact(leaves[0], leaves[1], leaves[2], radii, H, W,
                                      opacities=leaves[3])
        (out * g).sum().backward()
        res.append([out.detach()] + [t.grad for t in leaves])
    torch.cuda.synchronize()
    for a, b in zip(*res):
        assert scaled_err(to_np(b), to_np(a)) <= 2e-6
    print("image bit-identical:", torch.equal(res[0][0], res[1][0]))


def test_affine_backward_is_deterministic(cuda):
    _case(test="affine determinism")
    H, W = 96, 112
    args = _affine_field(31, 900, H, W, cuda, True, scale_hi=12.0)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(9))
    runs = []
    for _ in range(2):
        leaves = [args[i].detach().clone().requires_grad_(True) for i in (0, 1, 2, 4, 5, 6, 7)]
        out = HR.render_hip_affine(leaves[0], leaves[1], leaves[2], args[3], H, W, *leaves[4:],
                                   opacities=leaves[3])
        (out * g).sum().backward()
        runs.append([t.grad for t in leaves])
    torch.cuda.synchronize()
    for a, b in zip(*runs):
        assert torch.equal(a, b)


def test_clamped_scales_get_no_gradient(cuda):
    """A scale below the 1e-6 clamp (render.py:119-120) gets gradient 0, as through torch's
    clamp_min; the other scale of the same Gaussian still gets its gradient."""
    _case(test="affine scale clamp")
    H, W = 40, 48
    means, conics, colors, radii, op, cg, scales, rot = _affine_field(3, 60, H, W, cuda)
    scales = scales.clone()
    scales[::3, 0] = 1e-7
    scales = scales.requires_grad_(True)
    out = HR.render_hip_affine(means, conics, colors, radii, H, W, cg * 1e-6, scales, rot,
                               opacities=op)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(2))
    (out * g).sum().backward()
    torch.cuda.synchronize()
    assert np.isfinite(to_np(out)).all()
    _, ref = _oracle(means, conics, colors, radii, op, cg * 1e-6, scales, rot, H, W, True, False,
                     g)
    got = to_np(scales.grad)
    assert (got[::3, 0] == 0).all() and (ref[5][::3, 0] == 0).all()
    assert scaled_err(got, ref[5]) <= 1e-4


def test_render_field_dispatches_affine_colors(cuda):
    """render.py:359-374: normalized modes and ``additive`` render affine colors, the fixed-fade
    additive kernel modes reject them, and scales / rotations are required."""
    _case(test="render_field affine dispatch")
    H, W = 40, 52
    means, conics, colors, radii, op, cg, scales, rot = _affine_field(4, 120, H, W, cuda)
    kw = dict(opacities=op, scales=scales, rotations=rot, color_grads=cg)
    a = HR.render_field(means, conics, colors, radii, H, W, mode="cuda", **kw)
    b = HR.render_hip_affine(means, conics, colors, radii, H, W, cg, scales, rot, opacities=op)
    assert torch.equal(a, b)
    c = HR.render_field(means, conics, colors, radii, H, W, mode="additive", **kw)
    d = HR.render_hip_affine(means, conics, colors, radii, H, W, cg, scales, rot, opacities=op,
                             normalize=False)
    assert torch.equal(c, d)
    e = HR.render_field(means, conics, colors, radii, H, W, mode="cuda_additive",
                        support_fade_alpha=0.5, **kw)
    ref, _ = _oracle(means, conics, colors, radii, op, cg, scales, rot, H, W, False, 0.5)
    assert scaled_err(to_np(e), ref) <= 1e-5
    with pytest.raises(ValueError, match="normalized/additive renderers only"):
        HR.render_field(means, conics, colors, radii, H, W, mode="cuda_additive", **kw)
    with pytest.raises(ValueError, match="requires scales and rotations"):
        HR.render_field(means, conics, colors, radii, H, W, opacities=op, color_grads=cg)


def test_affine_empty_field_renders_zero(cuda):
    _case(test="affine empty field")
    H, W = 8, 9
    leaves = [torch.zeros(0, k, device=cuda, requires_grad=True) for k in (2, 3, 3)]
    cg = torch.zeros(0, 2, 3, device=cuda, requires_grad=True)
    scales = torch.zeros(0, 2, device=cuda, requires_grad=True)
    rot = torch.zeros(0, device=cuda, requires_grad=True)
    r = torch.zeros(0, 2, dtype=torch.long, device=cuda)
    out = HR.render_hip_affine(leaves[0], leaves[1], leaves[2], r, H, W, cg, scales, rot)
    out.sum().backward()
    torch.cuda.synchronize()
    assert out.shape == (H, W, 3) and torch.count_nonzero(out) == 0
    assert cg.grad is not None and cg.grad.shape == (0, 2, 3)


def test_affine_entry_points_reject_other_tile_sizes(cuda):
    """The affine kernels exist for 16-pixel tiles; the C ABI says so for 32."""
    from structsplat_amd import _native

    _case(test="affine tile size 32")
    H, W = 32, 32
    means, conics, colors, radii, op, cg, scales, rot = _affine_field(2, 12, H, W, cuda)
    radii = radii.to(torch.int64).contiguous()
    gids, offsets = HR.build_tile_index(means, conics, radii, H, W, 32)
    ws = torch.empty(_native.affine_workspace_bytes(12, H, W, 32), dtype=torch.uint8,
                     device=cuda)
    out = torch.empty(H, W, 3, device=cuda)
    lib = _native.load()
    rc = lib.ssplat_forward_tiled_affine(
        HR._ptr(means), HR._ptr(conics), HR._ptr(colors), HR._ptr(radii), HR._ptr(op),
        HR._ptr(cg), HR._ptr(scales), HR._ptr(rot), 12, HR._ptr(gids), HR._ptr(offsets),
        gids.numel(), H, W, 32, 1, 1e-8, 0.0, 3.0, HR._ptr(out), None, HR._ptr(ws), ws.numel(),
        HR._stream(means.device))
    with pytest.raises(ValueError, match="tile_size 16"):
        _native.check(rc, "ssplat_forward_tiled_affine")
