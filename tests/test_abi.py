"""The C-ABI library loads and exports every symbol include/ssplat.h declares (no GPU calls)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

from structsplat_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ssplat.h")


def _declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ssplat_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_symbols():
    assert set(_declared_symbols()) == set(_native.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    for name in _declared_symbols():
        assert hasattr(lib, name), name


def test_abi_version_and_workspace_query():
    lib = _native.load()
    assert lib.ssplat_abi_version() == 6
    nbytes = _native.workspace_bytes(1000, 256, 256, 16)
    # 1000 records of 64 B + (256 tiles + 1) counters + scan sums + status, 256-byte aligned.
    assert nbytes >= 1000 * 64 + 257 * 4
    assert nbytes % 256 == 0


@pytest.mark.parametrize("args,phrase", [
    ((10, 0, 5, 16), "height and width must be positive"),
    ((10, 5, 5, 0), "tile_size must be positive"),
    ((10, 5, 5, 33), "tile_size must be positive"),
    ((-1, 5, 5, 16), "n must be"),
])
def test_invalid_arguments_report_errors(args, phrase):
    with pytest.raises(ValueError, match=phrase):
        _native.workspace_bytes(*args)


def test_null_output_pointer_is_rejected():
    lib = _native.load()
    rc = lib.ssplat_workspace_bytes(10, 5, 5, 16, None)
    assert rc == _native.ERR_INVALID_ARGUMENT
    assert "NULL" in _native.last_error()
    assert isinstance(ctypes.c_size_t(0).value, int)


def test_normalized_backward_given_den_needs_out():
    """Argument checks run on the host before any device work: den without out is rejected (den
    alone cannot give the adjoints; out alone, or neither, selects the recomputing backward)."""
    lib = _native.load()
    dummy = ctypes.create_string_buffer(64)
    p = ctypes.cast(dummy, ctypes.c_void_p)
    rc = lib.ssplat_backward_tiled(p, p, p, p, p, None, p, None, 1, p, p, 1, 16, 16, 16, 1,
                                   1e-8, 0.0, 3.0, p, p, p, None, p, 1, None, p, 64, None)
    assert rc == _native.ERR_INVALID_ARGUMENT
    assert "also needs the forward's out" in _native.last_error()


def test_backward_needs_a_gradient_slot_buffer():
    lib = _native.load()
    dummy = ctypes.create_string_buffer(64)
    p = ctypes.cast(dummy, ctypes.c_void_p)
    rc = lib.ssplat_backward_tiled(p, p, p, p, p, None, p, p, 1, p, p, 1, 16, 16, 16, 1,
                                   1e-8, 0.0, 3.0, p, p, p, None, None, 1, None, p, 64, None)
    assert rc == _native.ERR_INVALID_ARGUMENT
    assert "grad_slots is NULL" in _native.last_error()
    rc = lib.ssplat_backward_tiled(p, p, p, p, p, None, p, p, 1, p, p, 1, 16, 16, 16, 1,
                                   1e-8, 0.0, 3.0, p, p, p, None, p, -1, None, p, 64, None)
    assert rc == _native.ERR_INVALID_ARGUMENT


def test_grad_slot_buffer_size():
    """Header (the call counter) plus 10 words per slot, rounded to 256 bytes."""
    assert _native.grad_slots_bytes(0) == 256
    assert _native.grad_slots_bytes(1) == 512
    assert _native.grad_slots_bytes(1000) == 256 + ((1000 * 40 + 255) // 256) * 256
    with pytest.raises(ValueError):
        _native.grad_slots_bytes(-1)


def test_kernel_timing_rejects_unknown_kernels():
    lib = _native.load()
    assert lib.ssplat_kernel_timing_read(7, None, None) == _native.ERR_INVALID_ARGUMENT
    n, ms = _native.kernel_timing_read(_native.KERNEL_FORWARD)  # nothing recorded: no GPU call
    assert (n, ms) == (0, 0.0)


def test_adam_step_rejects_bad_groups():
    lib = _native.load()
    assert lib.ssplat_adam_step(None, 0, None, 0.9, 0.999, 1e-8, None) == _native.ERR_INVALID_ARGUMENT
    groups = (_native.AdamGroup * 9)()
    assert lib.ssplat_adam_step(groups, 9, None, 0.9, 0.999, 1e-8, None) == \
        _native.ERR_INVALID_ARGUMENT
