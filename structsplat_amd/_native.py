"""ctypes binding of ``libssplat_hip.so``, the C ABI declared in ``include/ssplat.h``.

torch is imported before the library is opened: torch ships its own HIP runtime with the same
soname, and loading torch first makes the library bind to that one runtime instead of a second
copy.  Every wrapper here takes raw device addresses (ints) and a stream handle; tensors never
cross the ABI.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (one HIP runtime per process, see module docstring)

from ._build import LIB_PATH

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_CAPACITY = 2
ERR_HIP = 3
ERR_UNSUPPORTED = 4

_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double
_c_ptr = ctypes.c_void_p
_c_size = ctypes.c_size_t

# Exported symbol -> (restype, argtypes).  Mirrors include/ssplat.h.
SIGNATURES = {
    "ssplat_abi_version": (_c_int, []),
    "ssplat_last_error": (ctypes.c_char_p, []),
    "ssplat_workspace_bytes": (_c_int, [_c_i64, _c_int, _c_int, _c_int,
                                        ctypes.POINTER(_c_size)]),
    "ssplat_tile_index_count": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_i64, _c_int, _c_int, _c_int,
                                         _c_int, _c_dbl, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "ssplat_tile_index_fill": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_i64, _c_int, _c_int, _c_int,
                                        _c_int, _c_dbl, _c_ptr, _c_ptr, _c_i64, _c_ptr, _c_ptr,
                                        _c_size, _c_ptr]),
    "ssplat_build_tile_index": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_i64, _c_int, _c_int, _c_int,
                                         _c_int, _c_dbl, _c_ptr, _c_ptr, _c_i64, _c_ptr, _c_ptr,
                                         _c_size, _c_ptr]),
    "ssplat_forward_tiled": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_ptr,
                                      _c_ptr, _c_i64, _c_int, _c_int, _c_int, _c_int, _c_dbl,
                                      _c_dbl, _c_dbl, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "ssplat_forward_tiled_noalloc": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_i64,
                                              _c_ptr, _c_ptr, _c_i64, _c_int, _c_int, _c_int,
                                              _c_int, _c_dbl, _c_dbl, _c_dbl, _c_ptr, _c_ptr,
                                              _c_ptr]),
    "ssplat_backward_tiled": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr,
                                       _c_ptr, _c_i64, _c_ptr, _c_ptr, _c_i64, _c_int, _c_int,
                                       _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_ptr, _c_ptr,
                                       _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_ptr, _c_ptr, _c_size,
                                       _c_ptr]),
    "ssplat_grad_slots_bytes": (_c_int, [_c_i64, ctypes.POINTER(_c_size)]),
    "ssplat_forward_tiled_affine": (_c_int, [_c_ptr] * 8 + [_c_i64, _c_ptr, _c_ptr, _c_i64,
                                             _c_int, _c_int, _c_int, _c_int, _c_dbl, _c_dbl,
                                             _c_dbl, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "ssplat_backward_tiled_affine": (_c_int, [_c_ptr] * 11 + [_c_i64, _c_ptr, _c_ptr, _c_i64,
                                              _c_int, _c_int, _c_int, _c_int, _c_dbl, _c_dbl,
                                              _c_dbl] + [_c_ptr] * 8 + [_c_i64, _c_ptr, _c_ptr,
                                                                        _c_size, _c_ptr]),
    "ssplat_affine_workspace_bytes": (_c_int, [_c_i64, _c_int, _c_int, _c_int,
                                               ctypes.POINTER(_c_size)]),
    "ssplat_affine_grad_slots_bytes": (_c_int, [_c_i64, ctypes.POINTER(_c_size)]),
    "ssplat_sspl1_workspace_bytes": (_c_int, [_c_i64, ctypes.POINTER(_c_size)]),
    "ssplat_sspl1_decode": (_c_int, [_c_ptr] * 11 + [_c_ptr, _c_size, _c_ptr]),
    "ssplat_gaussian_activity": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_i64, _c_int, _c_int,
                                          _c_dbl, _c_dbl, _c_ptr, _c_ptr]),
    "ssplat_rect_pair_count": (_c_int, [_c_ptr, _c_ptr, _c_i64, _c_int, _c_int, _c_int, _c_ptr,
                                        _c_ptr, _c_size, _c_ptr]),
    "ssplat_kernel_timing_enable": (_c_int, [_c_int]),
    "ssplat_kernel_timing_read": (_c_int, [_c_int, ctypes.POINTER(_c_i64),
                                           ctypes.POINTER(_c_dbl)]),
    "ssplat_loss_workspace_bytes": (_c_int, [_c_i64, _c_i64, ctypes.POINTER(_c_size)]),
    "ssplat_loss_plan": (_c_int, [_c_ptr, _c_int, ctypes.POINTER(_c_i64),
                                  ctypes.POINTER(_c_i64), ctypes.POINTER(_c_size)]),
    "ssplat_l1_ssim_loss": (_c_int, [_c_ptr, _c_ptr, _c_int, _c_int, _c_dbl, _c_ptr, _c_ptr,
                                     _c_ptr, _c_size, _c_ptr]),
    "ssplat_batch_l1_ssim_loss": (_c_int, [_c_ptr, _c_int, _c_i64, _c_i64, _c_ptr, _c_ptr,
                                           _c_dbl, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "ssplat_adam_step": (_c_int, [_c_ptr, _c_int, _c_ptr, _c_dbl, _c_dbl, _c_dbl, _c_ptr]),
    "ssplat_produce_field": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_dbl, _c_dbl,
                                      _c_ptr, _c_ptr, _c_ptr, _c_ptr]),
    "ssplat_produce_field_backward": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_dbl,
                                               _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr]),
    "ssplat_loss_target_moments": (_c_int, [_c_ptr, _c_int, _c_int, _c_ptr, _c_ptr]),
    "ssplat_batch_loss_target_moments": (_c_int, [_c_ptr, _c_int, _c_i64, _c_i64, _c_ptr,
                                                  _c_ptr, _c_ptr]),
    "ssplat_l1_ssim_loss_forward": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_int, _c_int, _c_dbl,
                                             _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "ssplat_l1_ssim_loss_backward": (_c_int, [_c_ptr, _c_ptr, _c_int, _c_int, _c_dbl, _c_ptr,
                                              _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "ssplat_batch_l1_ssim_loss_forward": (_c_int, [_c_ptr, _c_int, _c_i64, _c_i64, _c_ptr,
                                                   _c_ptr, _c_ptr, _c_dbl, _c_ptr, _c_ptr,
                                                   _c_size, _c_ptr]),
    "ssplat_batch_l1_ssim_loss_backward": (_c_int, [_c_ptr, _c_int, _c_i64, _c_i64, _c_ptr,
                                                    _c_ptr, _c_dbl, _c_ptr, _c_ptr, _c_ptr,
                                                    _c_size, _c_ptr]),
    "ssplat_batch_plan": (_c_int, [_c_ptr, _c_int, _c_int, ctypes.POINTER(_c_i64),
                                   ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64),
                                   ctypes.POINTER(_c_size)]),
    "ssplat_batch_build_tile_index": (_c_int, [_c_ptr, _c_int, _c_i64, _c_i64, _c_int, _c_ptr,
                                               _c_ptr, _c_ptr, _c_int, _c_dbl, _c_ptr, _c_ptr,
                                               _c_i64, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "ssplat_batch_forward_tiled": (_c_int, [_c_ptr, _c_int, _c_i64, _c_i64, _c_int, _c_ptr,
                                            _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr,
                                            _c_i64, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_ptr,
                                            _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "ssplat_batch_backward_tiled": (_c_int, [_c_ptr, _c_int, _c_i64, _c_i64, _c_int, _c_ptr,
                                             _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr,
                                             _c_ptr, _c_ptr, _c_ptr, _c_i64, _c_int, _c_dbl,
                                             _c_dbl, _c_dbl, _c_ptr, _c_ptr, _c_ptr, _c_ptr,
                                             _c_ptr, _c_i64, _c_ptr, _c_ptr, _c_size, _c_ptr]),
}
# The after-forward backwards take the plain backwards' arguments.
SIGNATURES["ssplat_backward_tiled_after_forward"] = SIGNATURES["ssplat_backward_tiled"]
SIGNATURES["ssplat_batch_backward_tiled_after_forward"] = SIGNATURES["ssplat_batch_backward_tiled"]


class Field(ctypes.Structure):
    """ssplat_field of include/ssplat.h (one image of a batch)."""

    _fields_ = [
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("first_gaussian", ctypes.c_int64),
        ("num_gaussians", ctypes.c_int64),
        ("first_pixel", ctypes.c_int64),
        ("first_tile", ctypes.c_int64),
        ("tiles_x", ctypes.c_int32),
        ("num_tiles", ctypes.c_int32),
    ]

_LIB = None
_LOCK = threading.Lock()


class AdamGroup(ctypes.Structure):
    """ssplat_adam_group of include/ssplat.h (one parameter group of an Adam step)."""

    _fields_ = [
        ("param", ctypes.c_void_p),
        ("grad", ctypes.c_void_p),
        ("exp_avg", ctypes.c_void_p),
        ("exp_avg_sq", ctypes.c_void_p),
        ("count", ctypes.c_int64),
        ("lr", ctypes.c_double),
        ("clamp", ctypes.c_int32),
        ("lo", ctypes.c_double),
        ("hi", ctypes.c_double),
        ("hi_rows", ctypes.c_void_p),
        ("row_width", ctypes.c_int64),
    ]


class SSPL1Layout(ctypes.Structure):
    """ssplat_sspl1 of include/ssplat.h (an SSPL1 header's decoding parameters)."""

    _fields_ = [
        ("n", ctypes.c_int64),
        ("bits_means", ctypes.c_int32),
        ("bits_scales", ctypes.c_int32),
        ("bits_rot", ctypes.c_int32),
        ("bits_colors", ctypes.c_int32),
        ("bits_opacity", ctypes.c_int32),
        ("morton", ctypes.c_int32),
        ("color_delta", ctypes.c_int32),
        ("means_byte_planar", ctypes.c_int32),
        ("rot_circular", ctypes.c_int32),
        ("has_opacity", ctypes.c_int32),
        ("means_lo", ctypes.c_double * 2),
        ("means_hi", ctypes.c_double * 2),
        ("scale_lo", ctypes.c_double * 2),
        ("scale_hi", ctypes.c_double * 2),
        ("color_lo", ctypes.c_double * 3),
        ("color_hi", ctypes.c_double * 3),
    ]


class NativeLibraryMissing(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Open the in-tree HIP library (fails loudly when it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    with _LOCK:
        if _LIB is None:
            if not os.path.exists(LIB_PATH):
                raise NativeLibraryMissing(
                    f"StructSplat HIP library not found at {LIB_PATH}; build it with "
                    "`python -m structsplat_amd._build` (hipcc, gfx950)."
                )
            lib = ctypes.CDLL(LIB_PATH)
            for name, (restype, argtypes) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _LIB = lib
    return _LIB


def last_error() -> str:
    msg = load().ssplat_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    """Raise the Python exception matching a non-zero status code."""
    if rc == OK:
        return
    msg = last_error()
    if rc == ERR_INVALID_ARGUMENT:
        raise ValueError(f"{what}: {msg}")
    if rc == ERR_UNSUPPORTED:
        raise ValueError(f"{what}: unsupported: {msg}")
    if rc == ERR_CAPACITY:
        raise RuntimeError(f"{what}: capacity exceeded: {msg}")
    raise RuntimeError(f"{what}: HIP error: {msg}")


def workspace_bytes(n: int, height: int, width: int, tile_size: int) -> int:
    out = _c_size(0)
    check(load().ssplat_workspace_bytes(int(n), int(height), int(width), int(tile_size),
                                        ctypes.byref(out)), "ssplat_workspace_bytes")
    return int(out.value)


def grad_slots_bytes(slot_capacity: int) -> int:
    out = _c_size(0)
    check(load().ssplat_grad_slots_bytes(int(slot_capacity), ctypes.byref(out)),
          "ssplat_grad_slots_bytes")
    return int(out.value)


def affine_workspace_bytes(n: int, height: int, width: int, tile_size: int) -> int:
    out = _c_size(0)
    check(load().ssplat_affine_workspace_bytes(int(n), int(height), int(width), int(tile_size),
                                               ctypes.byref(out)), "ssplat_affine_workspace_bytes")
    return int(out.value)


def affine_grad_slots_bytes(slot_capacity: int) -> int:
    out = _c_size(0)
    check(load().ssplat_affine_grad_slots_bytes(int(slot_capacity), ctypes.byref(out)),
          "ssplat_affine_grad_slots_bytes")
    return int(out.value)


KERNEL_FORWARD = 0
KERNEL_BACKWARD = 1


def kernel_timing(enable: bool) -> None:
    """Switch the library's render-kernel event timing on or off (outside graph capture)."""
    check(load().ssplat_kernel_timing_enable(int(bool(enable))), "ssplat_kernel_timing_enable")


def kernel_timing_read(kernel: int) -> tuple[int, float]:
    """(launches, summed milliseconds) of the timed launches of one render kernel since the last
    read; waits for them to finish."""
    n = _c_i64(0)
    ms = _c_dbl(0.0)
    check(load().ssplat_kernel_timing_read(int(kernel), ctypes.byref(n), ctypes.byref(ms)),
          "ssplat_kernel_timing_read")
    return int(n.value), float(ms.value)
