"""The allocation-free forward (ssplat_forward_tiled_noalloc, static device scratch) against
ssplat_forward_tiled with a caller workspace, through the C ABI: bit for bit on two calls in a
row, the second with a smaller field (the scratch keeps the first call's contents), and sizes
beyond the scratch rejected before any device work."""
from __future__ import annotations

import ctypes

import pytest
import torch

from structsplat_amd import _native
from structsplat_amd import render as HR
from tests.helpers import field_tensors, random_field


def _ptr(t):
    return None if t is None or t.numel() == 0 else ctypes.c_void_p(t.data_ptr())


def test_noalloc_rejects_sizes_beyond_the_scratch():
    """Host-side check, no device memory touched (pointers are never dereferenced)."""
    lib = _native.load()
    fake = ctypes.c_void_p(256)
    rc = lib.ssplat_forward_tiled_noalloc(fake, fake, fake, fake, None, (1 << 18) + 1, fake,
                                          fake, 16, 64, 64, 16, 1, 1e-8, 0.0, 3.0, fake, None,
                                          None)
    assert rc == _native.ERR_INVALID_ARGUMENT
    assert "static scratch holds" in _native.last_error()
    rc = lib.ssplat_forward_tiled_noalloc(fake, fake, fake, fake, None, 100, fake, fake, 16,
                                          8192, 8192, 16, 1, 1e-8, 0.0, 3.0, fake, None, None)
    assert rc == _native.ERR_INVALID_ARGUMENT
    assert "tiles" in _native.last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("fade", [False, True])
def test_noalloc_forward_equals_workspace_forward(cuda, fade):
    print(f"case: noalloc forward, fade={fade}", flush=True)
    lib = _native.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
    # Two calls in a row, the second smaller: 1,200 Gaussians on 150x170, then 300 on 40x56.
    for seed, (N, H, W) in ((3, (1200, 150, 170)), (4, (300, 40, 56))):
        f = random_field(seed, N, H, W, scale_lo=0.6, scale_hi=9.0, opacity=True)
        means, conics, colors, radii, op = field_tensors(f, cuda)
        gids, offsets = HR.build_tile_index(means, conics, radii, H, W, 16, fade, 3.0)
        ref = torch.full((H, W, 3), float("nan"), device=cuda)
        ref_den = torch.full((H, W), float("nan"), device=cuda)
        ws = HR.workspace(N, H, W, 16, cuda)
        _native.check(lib.ssplat_forward_tiled(
            _ptr(means), _ptr(conics), _ptr(colors), _ptr(radii), _ptr(op), N, _ptr(gids),
            _ptr(offsets), gids.numel(), H, W, 16, 1, 1e-8, 1.0 if fade else 0.0, 3.0, _ptr(ref),
            _ptr(ref_den), _ptr(ws), ws.numel(), stream), "ssplat_forward_tiled")
        out = torch.full((H, W, 3), float("nan"), device=cuda)
        den = torch.full((H, W), float("nan"), device=cuda)
        _native.check(lib.ssplat_forward_tiled_noalloc(
            _ptr(means), _ptr(conics), _ptr(colors), _ptr(radii), _ptr(op), N, _ptr(gids),
            _ptr(offsets), gids.numel(), H, W, 16, 1, 1e-8, 1.0 if fade else 0.0, 3.0,
            _ptr(out), _ptr(den), stream), "ssplat_forward_tiled_noalloc")
        torch.cuda.synchronize()
        print("N", N, "H", H, "W", W, "max |diff|", float((out - ref).abs().max()), flush=True)
        assert torch.isfinite(out).all()
        assert torch.equal(out, ref)
        assert torch.equal(den, ref_den)
