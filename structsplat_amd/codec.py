"""SSPL1 decoding and decode-and-render on the GPU (the reference's codec.py:333-496).

The container is parsed on the host (magic ``SSPL1``, a uint32 header length, the JSON header,
one uint32-framed stream per attribute) and its streams are inflated there (zlib, or the PNG
plane codec); everything after that runs in ``libssplat_hip.so``: un-packing the quantized
integers, undoing the Morton / color delta coding (prefix sums mod 2^bits), dequantizing to the
field arrays (``ssplat_sspl1_decode``), the conic / radius / opacity producer, binning and the
forward render.  The decoded arrays are bit-identical to the reference's NumPy decode.

Entry points keep the reference's names: ``decode``, ``decode_and_render``, ``decode_alpha``,
``blob_header``, ``blob_components``; ``decode_and_render_with_coverage`` also returns the raw
coverage ``den`` of a normalized render.
"""
from __future__ import annotations

import ctypes
import json
import struct
import zlib
from io import BytesIO

import numpy as np
import torch

from . import _native
from .gaussians import GaussianField
from .render import (_EPS, _NORMALIZED_MODES, _ptr, _stream, render_field,
                     render_hip_exact_with_coverage)

_MAGIC = b"SSPL1"
_SCALE_LO = float(np.log(0.35))  # codec.py:37, the fitter's log-scale clamp


def blob_header(blob: bytes) -> dict:
    """The JSON header of a blob (codec.py:422-426)."""
    if blob[:5] != _MAGIC:
        raise ValueError("not a structsplat codec blob")
    (hlen,) = struct.unpack_from("<I", blob, 5)
    return json.loads(blob[9:9 + hlen])


def _stream_names(header: dict) -> list[str]:
    names = ["means", "scales", "rotation", "colors"]
    if bool(header.get("has_opacity", False)):
        names.append("opacity")
    if bool(header.get("has_alpha", False)):
        names.append("alpha")
    return names


def blob_components(blob: bytes) -> dict:
    """Exact byte accounting of one container (codec.py:429-479): header bytes plus each
    stream's framed bytes sum to the blob's length."""
    if len(blob) < 9 or blob[:5] != _MAGIC:
        raise ValueError("not a structsplat codec blob")
    (hlen,) = struct.unpack_from("<I", blob, 5)
    header_end = 9 + int(hlen)
    if header_end > len(blob):
        raise ValueError("truncated structsplat codec header")
    try:
        header = json.loads(blob[9:header_end])
    except (UnicodeDecodeError, json.JSONDecodeError) as exc:
        raise ValueError("invalid structsplat codec header") from exc
    raw_lengths = header.get("stream_raw_lengths", [])
    streams = {}
    off = header_end
    for index, name in enumerate(_stream_names(header)):
        if off + 4 > len(blob):
            raise ValueError(f"truncated {name} stream framing")
        (payload,) = struct.unpack_from("<I", blob, off)
        off += 4
        if off + int(payload) > len(blob):
            raise ValueError(f"truncated {name} stream payload")
        streams[name] = {
            "payload_bytes": int(payload),
            "framing_bytes": 4,
            "framed_bytes": 4 + int(payload),
            "raw_bytes": int(raw_lengths[index]) if index < len(raw_lengths) else None,
        }
        off += int(payload)
    if off != len(blob):
        raise ValueError(f"unexpected {len(blob) - off} trailing codec bytes")
    return {"total_bytes": len(blob), "magic_bytes": len(_MAGIC), "header_length_prefix_bytes": 4,
            "header_json_bytes": int(hlen), "header_bytes": header_end, "streams": streams}


def _inflate(payload: bytes, codec: str, raw_len: int | None) -> bytes:
    """codec.py:147-172: zlib, or a PNG image of the bytes (row-major, zero padded)."""
    if codec == "zlib":
        return zlib.decompress(payload)
    if codec == "png":
        if raw_len is None:
            raise ValueError("PNG stream requires stream_raw_lengths in the codec header")
        from PIL import Image

        arr = np.asarray(Image.open(BytesIO(payload)).convert("L"), dtype=np.uint8).reshape(-1)
        if arr.size < raw_len:
            raise ValueError(f"PNG stream too short: got {arr.size} bytes, expected {raw_len}")
        return arr[:raw_len].tobytes()
    raise ValueError(f"unknown stream_codec {codec!r}; expected 'zlib' or 'png'")


def _raw_streams(blob: bytes, count: int) -> tuple[dict, list[bytes]]:
    header = blob_header(blob)
    off = 9 + struct.unpack_from("<I", blob, 5)[0]
    codec = header.get("stream_codec", "zlib")
    raw_lengths = header.get("stream_raw_lengths", [])
    raws = []
    for i in range(count):
        (zlen,) = struct.unpack_from("<I", blob, off)
        off += 4
        raw_len = int(raw_lengths[i]) if i < len(raw_lengths) else None
        raws.append(_inflate(blob[off:off + zlen], codec, raw_len))
        off += zlen
    return header, raws


def _width(bits: int) -> int:
    return 1 if bits <= 8 else (2 if bits <= 16 else 4)


def decode(blob: bytes, device: str = "cuda") -> GaussianField:
    """The field of a blob (codec.py:333-391), decoded by the HIP kernels on ``device``."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError("SSPL1 decoding runs on the GPU; pass a CUDA device")
    header = blob_header(blob)
    has_opacity = bool(header.get("has_opacity", False))
    header, raws = _raw_streams(blob, 5 if has_opacity else 4)
    n, H, W = int(header["n"]), int(header["H"]), int(header["W"])
    b_means, b_scales, b_rot, b_colors = (int(b) for b in header["bits"])
    b_op = int(header.get("bits_opacity", 8))
    need = [2 * n * _width(b_means), 2 * n * _width(b_scales), n * _width(b_rot),
            3 * n * _width(b_colors)] + ([n * _width(b_op)] if has_opacity else [])
    for raw, size, name in zip(raws, need, _stream_names(header)):
        if len(raw) < size:
            raise ValueError(f"{name} stream holds {len(raw)} bytes, needs {size}")
    big = float(np.log(max(H, W)))
    lay = _native.SSPL1Layout()
    lay.n = n
    lay.bits_means, lay.bits_scales, lay.bits_rot, lay.bits_colors = b_means, b_scales, b_rot, \
        b_colors
    lay.bits_opacity = b_op
    lay.morton = int(bool(header["morton"]))
    lay.color_delta = int(bool(header.get("color_delta", False)))
    lay.means_byte_planar = int(bool(header.get("means_byte_planar", False)))
    lay.rot_circular = int(bool(header.get("rot_circular", False)))
    lay.has_opacity = int(has_opacity)
    for dst, src in ((lay.means_lo, header.get("means_lo", [0.0, 0.0])),
                     (lay.means_hi, header.get("means_hi", [W - 1.0, H - 1.0])),
                     (lay.scale_lo, header.get("scale_lo", [_SCALE_LO, _SCALE_LO])),
                     (lay.scale_hi, header.get("scale_hi", [big, big])),
                     (lay.color_lo, header["color_lo"]), (lay.color_hi, header["color_hi"])):
        for k, v in enumerate(src):
            dst[k] = float(v)
    # One host-to-device copy of all raw streams, each at a 256-byte aligned offset.
    offsets, total = [], 0
    for size in need:
        offsets.append(total)
        total += (size + 255) // 256 * 256
    host = np.zeros(max(total, 1), dtype=np.uint8)
    for raw, off, size in zip(raws, offsets, need):
        host[off:off + size] = np.frombuffer(raw, dtype=np.uint8, count=size)
    dev_raw = torch.from_numpy(host).to(dev)
    base = dev_raw.data_ptr()
    ptrs = [ctypes.c_void_p(base + off) for off in offsets] + [None] * (5 - len(offsets))
    means = torch.empty(n, 2, dtype=torch.float32, device=dev)
    log_scales = torch.empty(n, 2, dtype=torch.float32, device=dev)
    rotations = torch.empty(n, dtype=torch.float32, device=dev)
    colors = torch.empty(n, 3, dtype=torch.float32, device=dev)
    logits = torch.empty(n, dtype=torch.float32, device=dev) if has_opacity else None
    nbytes = ctypes.c_size_t(0)
    lib = _native.load()
    _native.check(lib.ssplat_sspl1_workspace_bytes(n, ctypes.byref(nbytes)),
                  "ssplat_sspl1_workspace_bytes")
    ws = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8, device=dev)
    _native.check(lib.ssplat_sspl1_decode(
        ctypes.byref(lay), *ptrs, _ptr(means), _ptr(log_scales), _ptr(rotations), _ptr(colors),
        _ptr(logits), _ptr(ws), ws.numel(), _stream(dev)), "ssplat_sspl1_decode")
    return GaussianField(means, log_scales, rotations, colors, logits)


def decode_alpha(blob: bytes) -> np.ndarray | None:
    """The (H, W) float32 alpha map in [0, 1], or None for alpha-free blobs (codec.py:394-419)."""
    header = blob_header(blob)
    if not header.get("has_alpha", False):
        return None
    H, W = int(header["H"]), int(header["W"])
    alpha_index = 5 if bool(header.get("has_opacity", False)) else 4
    header, raws = _raw_streams(blob, alpha_index + 1)
    raw = raws[alpha_index]
    if len(raw) < H * W:
        raise ValueError("alpha stream missing despite has_alpha header flag")
    return np.frombuffer(raw, dtype=np.uint8, count=H * W).reshape(H, W).astype(np.float32) / 255


def _render_args(blob: bytes, device: str):
    h = blob_header(blob)
    field = decode(blob, device=device)
    aa = float(h.get("aa_dilation", 0.0))
    sigma = float(h.get("sigma_cutoff", 3.0))
    conics, radii, op = field.produce(sigma, aa)
    return h, field, conics, radii, op, sigma


@torch.no_grad()
def decode_and_render(blob: bytes, device: str = "cuda") -> torch.Tensor:
    """Decode a blob and render it with the render semantics its header stores
    (codec.py:482-500): renderer mode, anti-alias dilation, sigma cutoff, support fade,
    normalization eps.  Returns the (H, W, 3) image."""
    h, field, conics, radii, op, sigma = _render_args(blob, device)
    return render_field(field.means, conics, field.colors, radii, int(h["H"]), int(h["W"]),
                        int(h.get("render_chunk", 512)), h.get("renderer", "normalized"), op,
                        support_fade=bool(h.get("support_fade", False)), sigma_cutoff=sigma,
                        normalization_eps=float(h.get("normalization_eps", _EPS)))


@torch.no_grad()
def decode_and_render_with_coverage(blob: bytes, device: str = "cuda"):
    """Decode and render a normalized blob; returns ``(out (H,W,3), den (H,W))`` where ``den``
    is the raw coverage sum_i o_i G_i of the same forward (cuda_render.py:229-253)."""
    h, field, conics, radii, op, sigma = _render_args(blob, device)
    if h.get("renderer", "normalized") not in _NORMALIZED_MODES:
        raise ValueError("coverage is defined for normalized renders")
    return render_hip_exact_with_coverage(field.means, conics, field.colors, radii,
                                          int(h["H"]), int(h["W"]), opacities=op,
                                          eps=float(h.get("normalization_eps", _EPS)),
                                          support_fade=bool(h.get("support_fade", False)),
                                          sigma_cutoff=sigma)
