"""GPU parity of the HIP path against the CPU oracle (tile lists bit-exact, forward vs float64,
backward vs the float64 analytic gradients)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import exact64, tile_index
from structsplat_amd import render as HR
from tests.helpers import (field_tensors, pixel_bar_err, pixel_rel_err, random_field,
                           scaled_err, to_np)

pytestmark = pytest.mark.gpu


def _print_case(**kw):
    print("case:", ", ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("ts", [4, 8, 16, 32])
@pytest.mark.parametrize("cull", [False, True])
def test_tile_index_bit_exact(cuda, ts, cull):
    _print_case(test="tile_index", ts=ts, cull=cull)
    N, H, W = 300, 67, 45
    f = random_field(21, N, H, W, scale_lo=0.6, scale_hi=9.0, margin=6.0,
                     aniso=(1.0, 10.0, 0.4, 2.0))
    means, conics, colors, radii, _ = field_tensors(f, cuda)
    gids, offsets = HR.build_tile_index(means, conics, radii, H, W, ts, cull, 3.0)
    torch.cuda.synchronize()
    ref_g, ref_o = tile_index.build_tile_index(to_np(means), to_np(conics), to_np(radii), H, W,
                                               ts, cull, 3.0)
    assert offsets.dtype == torch.int32 and gids.dtype == torch.int32
    assert to_np(offsets).tolist() == ref_o.tolist()
    assert to_np(gids).tolist() == ref_g.tolist()


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("H,W,ts", [(1100, 1100, 4), (300, 200, 4), (1100, 1100, 2)])
def test_tile_index_scan_paths_bit_exact(cuda, H, W, ts, fused):
    """Both binning schemes: 75,625 and 3,750 tiles bin by strips of 16 tiles (4,727 and 235
    strips); 302,500 tiles, more than 7,680 strips of 32, bin by the per-pair kernels, whose
    counters take the three-launch scan.  Through the two-call and the one-call binning."""
    _print_case(test="tile_index_scan_paths", H=H, W=W, ts=ts, fused=fused)
    f = random_field(33, 3000, H, W, scale_lo=0.6, scale_hi=12.0, margin=8.0)
    means, conics, colors, radii, _ = field_tensors(f, cuda)
    ref_g, ref_o = tile_index.build_tile_index(to_np(means), to_np(conics), to_np(radii), H, W,
                                               ts, False, 3.0)
    if fused:
        gids, offsets, status = HR.build_tile_index_capacity(means, conics, radii, H, W,
                                                             len(ref_g) + 7, ts)
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        gids = gids[:len(ref_g)]
    else:
        gids, offsets = HR.build_tile_index(means, conics, radii, H, W, ts)
        torch.cuda.synchronize()
    assert to_np(offsets).tolist() == ref_o.tolist()
    assert to_np(gids).tolist() == ref_g.tolist()


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("cull", [False, True])
def test_tile_index_bucket_overflow_bit_exact(cuda, cull, fused):
    """Gaussians whose rectangles cover the whole image meet every strip: 100 needles 300-600 px
    long at 1024x1024 with 4-pixel tiles (4,096 strips of 16 tiles) make ~400,000 strip entries,
    beyond the bucket's 8 N + 65,536 = 68,736 (N = 400), so the strips past it take every
    Gaussian as a candidate.  The ellipse cull drops most of the needles' pairs.  The lists equal
    the oracle's."""
    _print_case(test="bucket_overflow", cull=cull, fused=fused)
    rng = np.random.default_rng(5)
    H = W = 1024
    nb, ns = 100, 300
    means_np = np.concatenate([
        np.stack([rng.uniform(0, W, nb), rng.uniform(0, H, nb)], 1),
        np.stack([rng.uniform(0, W, ns), rng.uniform(0, H, ns)], 1)]).astype(np.float32)
    scales = np.concatenate([
        np.stack([rng.uniform(300, 600, nb), rng.uniform(20, 60, nb)], 1),
        np.exp(rng.uniform(np.log(0.6), np.log(8.0), (ns, 2)))]).astype(np.float32)
    order = rng.permutation(nb + ns)
    f = {"means": means_np[order], "scales": scales[order],
         "angles": rng.uniform(0, np.pi, nb + ns).astype(np.float32),
         "colors": rng.random((nb + ns, 3)).astype(np.float32), "opacity": None}
    means, conics, colors, radii, _ = field_tensors(f, cuda)
    ref_g, ref_o = tile_index.build_tile_index(to_np(means), to_np(conics), to_np(radii), H, W,
                                               4, cull, 3.0)
    if fused:
        gids, offsets, status = HR.build_tile_index_capacity(means, conics, radii, H, W,
                                                             len(ref_g) + 3, 4, cull, 3.0)
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        gids = gids[:len(ref_g)]
    else:
        gids, offsets = HR.build_tile_index(means, conics, radii, H, W, 4, cull, 3.0)
        torch.cuda.synchronize()
    assert to_np(offsets).tolist() == ref_o.tolist()
    assert np.array_equal(to_np(gids), ref_g)


def test_tile_index_reference_golden(cuda):
    # test_render.py:167-172
    means = torch.tensor([[2.0, 2.0], [20.0, 20.0]], device=cuda)
    radii = torch.tensor([[1, 1], [1, 1]], dtype=torch.long, device=cuda)
    conics = torch.ones(2, 3, device=cuda)
    gids, offsets = HR.build_tile_index(means, conics, radii, 8, 8, 4)
    assert offsets.tolist() == [0, 1, 1, 1, 1]
    assert gids.tolist() == [0]


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True, 0.5])
@pytest.mark.parametrize("opacity", [False, True])
def test_forward_matches_fp64(cuda, normalize, fade, opacity):
    _print_case(test="forward", normalize=normalize, fade=fade, opacity=opacity)
    N, H, W = 400, 70, 93
    f = random_field(7, N, H, W, scale_lo=0.7, scale_hi=6.0, opacity=opacity)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    out = HR.render_hip_exact(means, conics, colors, radii, H, W, opacities=op,
                              normalize=normalize, support_fade=fade)
    torch.cuda.synchronize()
    args = (to_np(means), to_np(conics), to_np(colors), to_np(radii), to_np(op), H, W,
            normalize, 1e-8, fade, 3.0)
    ref, _ = exact64.forward(*args)
    sens = exact64.forward_conditioning(*args)
    got = to_np(out)
    assert np.isfinite(got).all()
    err = scaled_err(got, ref)
    bar = pixel_bar_err(got, ref, sens)
    print("scaled err", err, "per-pixel rel err (floor 1e-6)", pixel_rel_err(got, ref),
          "per-pixel bar", bar)
    assert err <= 1e-5
    assert pixel_rel_err(got, ref) <= 1e-5  # per pixel, floor 1e-6 of the image scale
    assert bar <= 1.0


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("fade", [False, True, 0.5])
def test_backward_matches_fp64(cuda, normalize, fade):
    _print_case(test="backward", normalize=normalize, fade=fade)
    N, H, W = 250, 61, 77
    f = random_field(12, N, H, W, scale_lo=0.8, scale_hi=5.0, opacity=True)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    for t in (means, conics, colors, op):
        t.requires_grad_(True)
    out = HR.render_hip_exact(means, conics, colors, radii, H, W, opacities=op,
                              normalize=normalize, support_fade=fade)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(3))
    (out * g).sum().backward()
    torch.cuda.synchronize()
    ref = exact64.backward(to_np(g), to_np(means), to_np(conics), to_np(colors), to_np(radii),
                           to_np(op), H, W, normalize, 1e-8, fade, 3.0)
    for name, got, exp in zip(("means", "conics", "colors", "opacities"),
                              (means.grad, conics.grad, colors.grad, op.grad), ref):
        err = scaled_err(to_np(got), exp)
        print(name, "scaled err", err)
        assert err <= 1e-4, name


@pytest.mark.parametrize("fade", [False, True, 0.5])
def test_backward_normalizer_modes_agree(cuda, fade):
    """Saved den and out, recomputed den, and recomputed numerator and den: same gradients
    (the recomputed values equal the forward's bit for bit; only atomics reorder sums)."""
    _print_case(test="backward_modes", fade=fade)
    N, H, W = 300, 66, 83
    f = random_field(19, N, H, W, scale_lo=0.8, scale_hi=6.0, opacity=True)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    alpha = HR._fade_alpha(fade)
    gids, offsets = HR.build_tile_index(means, conics, radii, H, W, 16, alpha == 1.0, 3.0)
    den = torch.empty(H, W, device=cuda)
    out, _ = HR.forward_tiled(means, conics, colors, radii, op, gids, offsets, H, W, True, 1e-8,
                              16, alpha, 3.0, den=den)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(4))
    ref = exact64.backward(to_np(g), to_np(means), to_np(conics), to_np(colors), to_np(radii),
                           to_np(op), H, W, True, 1e-8, fade, 3.0)
    results = []
    for d, o in ((den, out), (None, out), (None, None)):
        grads = HR.backward_tiled(g, means, conics, colors, radii, op, d, o, gids, offsets, H, W,
                                  True, 1e-8, 16, alpha, 3.0)
        torch.cuda.synchronize()
        results.append([to_np(x) for x in grads])
        for got, exp in zip(results[-1], ref):
            assert scaled_err(got, exp) <= 1e-4
    for other in results[1:]:
        for a, b in zip(results[0], other):
            assert scaled_err(b, a) <= 1e-6


@pytest.mark.parametrize("fade", [False, True, 0.5])
@pytest.mark.parametrize("normalize", [True, False])
def test_tile32_forward_equals_tile16_and_gradients_match_fp64(cuda, fade, normalize):
    """Every pixel accumulates the Gaussians whose rectangle holds it in ascending id at any tile
    size, so the 32-pixel-tile forward equals the 16-pixel-tile forward bit for bit; the
    32-pixel-tile backward (all normalizer modes) matches the float64 oracle."""
    _print_case(test="tile32", fade=fade, normalize=normalize)
    N, H, W = 700, 90, 111
    f = random_field(23, N, H, W, scale_lo=0.8, scale_hi=9.0, opacity=True)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    alpha = HR._fade_alpha(fade)
    outs = {}
    for ts in (16, 32):
        gids, offsets = HR.build_tile_index(means, conics, radii, H, W, ts, alpha == 1.0, 3.0)
        den = torch.empty(H, W, device=cuda)
        out, _ = HR.forward_tiled(means, conics, colors, radii, op, gids, offsets, H, W,
                                  normalize, 1e-8, ts, alpha, 3.0, den=den)
        outs[ts] = (out, den, gids, offsets)
    torch.cuda.synchronize()
    assert torch.equal(outs[16][0], outs[32][0])
    assert torch.equal(outs[16][1], outs[32][1])
    out, den, gids, offsets = outs[32]
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(8))
    ref = exact64.backward(to_np(g), to_np(means), to_np(conics), to_np(colors), to_np(radii),
                           to_np(op), H, W, normalize, 1e-8, fade, 3.0)
    modes = ((den, out), (None, out), (None, None)) if normalize else ((None, None),)
    for d, o in modes:
        grads = HR.backward_tiled(g, means, conics, colors, radii, op, d, o, gids, offsets, H, W,
                                  normalize, 1e-8, 32, alpha, 3.0)
        torch.cuda.synchronize()
        for got, exp in zip(grads, ref):
            assert scaled_err(to_np(got), exp) <= 1e-4


def test_tile32_long_lists(cuda):
    """32-pixel tiles whose lists span several 256-record batches."""
    _print_case(test="tile32 long lists")
    rng = np.random.default_rng(99)
    N, H, W = 1500, 40, 40
    f = {"means": rng.uniform(0, 40, (N, 2)).astype(np.float32),
         "scales": np.exp(rng.uniform(np.log(1.0), np.log(8.0), (N, 2))).astype(np.float32),
         "angles": rng.uniform(0, np.pi, N).astype(np.float32),
         "colors": rng.random((N, 3)).astype(np.float32),
         "opacity": rng.uniform(0.1, 1.0, N).astype(np.float32)}
    means, conics, colors, radii, op = field_tensors(f, cuda)
    out16 = HR.render_hip_exact(means, conics, colors, radii, H, W, opacities=op, tiled=True,
                                tile_size=16)
    x = means.clone().requires_grad_(True)
    out32 = HR.render_hip_exact(x, conics, colors, radii, H, W, opacities=op, tiled=True,
                                tile_size=32)
    g = torch.rand(H, W, 3, device=cuda, generator=torch.Generator(device=cuda).manual_seed(9))
    (out32 * g).sum().backward()
    torch.cuda.synchronize()
    assert torch.equal(out16, out32)
    ref = exact64.backward(to_np(g), to_np(means), to_np(conics), to_np(colors), to_np(radii),
                           to_np(op), H, W, True, 1e-8, False, 3.0)
    assert scaled_err(to_np(x.grad), ref[0]) <= 1e-4


def test_empty_field_renders_zero(cuda):
    H, W = 8, 9
    z2 = torch.zeros(0, 2, device=cuda)
    z3 = torch.zeros(0, 3, device=cuda)
    r = torch.zeros(0, 2, dtype=torch.long, device=cuda)
    out = HR.render_field(z2, z3, z3, r, H, W, mode="hip")
    torch.cuda.synchronize()
    assert out.shape == (H, W, 3)
    assert torch.count_nonzero(out) == 0


def test_nonfinite_means_are_dropped(cuda):
    H, W = 12, 12
    f = {"means": np.array([[5.0, 5.0], [np.nan, 3.0], [1e12, 1e12], [np.inf, 2.0]], np.float32),
         "scales": np.full((4, 2), 1.5, np.float32), "angles": np.zeros(4, np.float32),
         "colors": np.ones((4, 3), np.float32), "opacity": None}
    means, conics, colors, radii, _ = field_tensors(f, cuda)
    out = HR.render_field(means, conics, colors, radii, H, W, mode="hip")
    ref, _ = exact64.forward(to_np(means), to_np(conics), to_np(colors), to_np(radii), None, H, W)
    torch.cuda.synchronize()
    assert torch.isfinite(out).all()
    assert scaled_err(to_np(out), ref) <= 1e-5


def test_forward_is_deterministic(cuda):
    N, H, W = 2000, 128, 160
    f = random_field(5, N, H, W, scale_lo=0.7, scale_hi=8.0, opacity=True)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    a = HR.render_hip_exact(means, conics, colors, radii, H, W, opacities=op)
    b = HR.render_hip_exact(means, conics, colors, radii, H, W, opacities=op)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_capacity_status_is_sticky_and_exact_capacity_is_enough(cuda):
    N, H, W = 500, 80, 96
    f = random_field(17, N, H, W, scale_lo=0.7, scale_hi=7.0)
    means, conics, colors, radii, _ = field_tensors(f, cuda)
    ref_g, ref_o = HR.build_tile_index(means, conics, radii, H, W)
    P = int(ref_o[-1])
    gids, offsets, status = HR.build_tile_index_capacity(means, conics, radii, H, W, P)
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    assert torch.equal(offsets, ref_o) and torch.equal(gids[:P], ref_g)
    small = HR.build_tile_index_capacity(means, conics, radii, H, W, P // 2)
    torch.cuda.synchronize()
    assert int(small[2].item()) == 2  # capacity exceeded
    # The word stays set after a later binning that fits (sticky until the caller clears it).
    HR.build_tile_index_capacity(means, conics, radii, H, W, P, out=(gids, offsets, small[2],
                                 HR.workspace(N, H, W, 16, cuda)))
    torch.cuda.synchronize()
    assert int(small[2].item()) == 2


def test_render_kernels_reject_other_tile_sizes(cuda):
    f = random_field(1, 10, 16, 16)
    means, conics, colors, radii, _ = field_tensors(f, cuda)
    gids, offsets = HR.build_tile_index(means, conics, radii, 16, 16, 8)
    with pytest.raises(ValueError, match="tile_size"):
        HR.forward_tiled(means, conics, colors, radii, None, gids, offsets, 16, 16, True, 1e-8,
                         8, False, 3.0)


@pytest.mark.parametrize("n_dense", [70, 200, 900, 2000, 6000])
def test_tile_index_long_lists_bit_exact(cuda, n_dense):
    """Dense tiles: lists of ~n_dense entries exercise every sort path (registers with one, two
    and four entries per lane, wave LDS, block LDS, global memory)."""
    _print_case(test="long lists", n_dense=n_dense)
    rng = np.random.default_rng(n_dense)
    H, W = 40, 52
    dense = np.stack([rng.uniform(17, 30, n_dense), rng.uniform(17, 30, n_dense)], 1)
    sparse = np.stack([rng.uniform(0, W, 200), rng.uniform(0, H, 200)], 1)
    means_np = np.concatenate([dense, sparse]).astype(np.float32)
    order = rng.permutation(len(means_np))  # interleave dense and sparse ids
    f = {"means": means_np[order],
         "scales": np.exp(rng.uniform(np.log(0.5), np.log(3.0), (len(order), 2))).astype(np.float32),
         "angles": rng.uniform(0, np.pi, len(order)).astype(np.float32),
         "colors": rng.random((len(order), 3)).astype(np.float32), "opacity": None}
    means, conics, colors, radii, _ = field_tensors(f, cuda)
    gids, offsets = HR.build_tile_index(means, conics, radii, H, W, 16)
    torch.cuda.synchronize()
    ref_g, ref_o = tile_index.build_tile_index(to_np(means), to_np(conics), to_np(radii), H, W,
                                               16)
    assert np.diff(ref_o).max() >= min(n_dense, 64)
    assert to_np(offsets).tolist() == ref_o.tolist()
    assert to_np(gids).tolist() == ref_g.tolist()


@pytest.mark.parametrize("cull", [False, True])
def test_tile_index_strip_cache_rounds_bit_exact(cuda, cull):
    """One strip met by about 2000 Gaussians: at 2048x2048 with 16-pixel tiles a strip is 16
    tiles of one tile row, so the strip of tile row 10, tiles 0-15, gets more candidates than
    level 2's LDS cache holds (1024) and takes them in two rounds, while every tile's list stays
    within its LDS segment (256); the lists must equal the oracle's."""
    _print_case(test="strip cache rounds", cull=cull)
    rng = np.random.default_rng(5)
    H = W = 2048
    n_dense = 2000
    dense = np.stack([rng.uniform(4, 252, n_dense), rng.uniform(163, 173, n_dense)], 1)
    sparse = np.stack([rng.uniform(0, W, 3000), rng.uniform(0, H, 3000)], 1)
    means_np = np.concatenate([dense, sparse]).astype(np.float32)
    order = rng.permutation(len(means_np))
    n = len(order)
    f = {"means": means_np[order],
         "scales": np.exp(rng.uniform(np.log(0.5), np.log(1.0), (n, 2))).astype(np.float32),
         "angles": rng.uniform(0, np.pi, n).astype(np.float32),
         "colors": rng.random((n, 3)).astype(np.float32), "opacity": None}
    means, conics, colors, radii, _ = field_tensors(f, cuda)
    gids, offsets = HR.build_tile_index(means, conics, radii, H, W, 16, ellipse_cull=cull)
    torch.cuda.synchronize()
    ref_g, ref_o = tile_index.build_tile_index(to_np(means), to_np(conics), to_np(radii), H, W,
                                               16, ellipse_cull=cull)
    strip = range(10 * 128, 10 * 128 + 16)
    members = {int(g) for t in strip for g in ref_g[ref_o[t]:ref_o[t + 1]]}
    assert len(members) > 1024
    assert max(int(ref_o[t + 1] - ref_o[t]) for t in strip) <= 256
    assert to_np(offsets).tolist() == ref_o.tolist()
    assert to_np(gids).tolist() == ref_g.tolist()


@pytest.mark.parametrize("fade", [False, True, 0.5])
def test_normalized_render_is_invariant_to_opacity_scale(cuda, fade):
    """Scaling every opacity by k = 4 (a power of two, so every weight and partial sum scales
    exactly) multiplies the coverage by exactly 4 and leaves the normalized image unchanged up to
    the eps term: out = 4N / (4D + eps).  A dropped or doubled opacity factor in either sum
    breaks one of the two."""
    _print_case(test="opacity scale invariance", fade=fade)
    N, H, W = 500, 72, 88
    f = random_field(27, N, H, W, scale_lo=0.7, scale_hi=6.0, opacity=True)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    out1, den1 = HR.render_hip_exact_with_coverage(means, conics, colors, radii, H, W,
                                                    opacities=op, support_fade=fade)
    out4, den4 = HR.render_hip_exact_with_coverage(means, conics, colors, radii, H, W,
                                                    opacities=op * 4.0, support_fade=fade)
    torch.cuda.synchronize()
    # Exact wherever the sums stayed clear of the subnormal range (a subnormal partial weight
    # rounds on a fixed grid that scaling does not commute with).
    normal = den1 > 1e-20
    assert torch.equal(den4[normal], den1[normal] * 4.0)
    live = to_np(den1) > 1e-3
    assert np.abs(to_np(out4) - to_np(out1))[live].max() <= 1e-6 * max(
        1.0, float(np.abs(to_np(out1)).max()))


def test_reference_renderer_names_route_to_render_field(cuda):
    """render / render_additive (render.py:238-263) and the reference's CUDA entry names give
    render_field's images bit for bit."""
    _print_case(test="reference renderer names")
    N, H, W = 200, 40, 56
    f = random_field(31, N, H, W, scale_lo=0.7, scale_hi=5.0, opacity=True)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    a = HR.render(means, conics, colors, radii, H, W, opacities=op, support_fade=True,
                  support_fade_alpha=0.5)
    b = HR.render_field(means, conics, colors, radii, H, W, mode="normalized", opacities=op,
                        support_fade=True, support_fade_alpha=0.5)
    c = HR.render_additive(means, conics, colors, radii, H, W, opacities=op)
    d = HR.render_field(means, conics, colors, radii, H, W, mode="additive", opacities=op)
    e = HR.render_cuda_exact(means, conics, colors, radii, H, W, opacities=op)
    g = HR.render_field(means, conics, colors, radii, H, W, mode="cuda", opacities=op)
    torch.cuda.synchronize()
    assert torch.equal(a, b) and torch.equal(c, d) and torch.equal(e, g)
    assert HR.render_cuda_exact_with_coverage is HR.render_hip_exact_with_coverage


def test_tile_index_backend_names(cuda):
    """render_cuda_exact's tile_index_backend (cuda_render.py:285-288): "cuda" and "torch" give
    the same image, and an unknown name is a ValueError (test_render.py:411-414)."""
    _print_case(test="tile_index_backend names")
    N, H, W = 150, 36, 44
    f = random_field(41, N, H, W, scale_lo=0.7, scale_hi=5.0, opacity=True)
    means, conics, colors, radii, op = field_tensors(f, cuda)
    a = HR.render_cuda_exact(means, conics, colors, radii, H, W, opacities=op, tiled=True,
                             tile_index_backend="cuda")
    b = HR.render_cuda_exact(means, conics, colors, radii, H, W, opacities=op, tiled=True,
                             tile_index_backend="torch")
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    with pytest.raises(ValueError, match="tile_index_backend must be"):
        HR.render_cuda_exact(means, conics, colors, radii, H, W, tiled=True,
                             tile_index_backend="nope")
