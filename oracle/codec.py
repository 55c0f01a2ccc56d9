"""NumPy restatement of the SSPL1 container (codec.py:36-496).  Test infrastructure only.

An SSPL1 blob is ``b"SSPL1"``, a little-endian uint32 header length, a JSON header, then one
framed stream per attribute (uint32 payload length + payload): means, log-scales, rotations,
colors and, with ``has_opacity``, opacities (an optional alpha plane follows).  Each stream is
zlib- (or PNG-) compressed raw bytes of uniformly quantized values:

* means: 2 x n values (all x, then all y) over ``[means_lo, means_hi]``, ``bits[0]`` bits,
  Morton-ordered and delta-coded mod 2^bits when ``morton``, byte-planar when
  ``means_byte_planar`` (all low bytes, then the next byte plane, ...);
* log-scales: 2 x n values over ``[scale_lo, scale_hi]`` per axis;
* rotations: n values, ``2^bits`` bins over ``[0, pi)`` when ``rot_circular``, else
  ``2^bits - 1`` levels over ``[0, pi]``;
* colors: 3 x n values over ``[color_lo, color_hi]`` per channel, delta-coded when
  ``color_delta``;
* opacities: n probabilities over ``[0, 1]`` (``bits_opacity``), decoded to logits after a clip
  to ``[1e-4, 1 - 1e-4]``.

Values are stored as uint8 / uint16 / uint32 by bit width.  ``decode_arrays`` follows
``codec.decode`` (codec.py:333-391) step for step and returns the field arrays; ``encode_arrays``
follows ``codec.encode`` (codec.py:235-330) from field arrays.

Pinned by: re-encoding the decoded real SSPL1 blobs of the reference repository
(``tests/golden/sspl1``) reproduces their raw streams byte for byte, and the reference's codec
tests restated in ``tests/test_oracle_codec.py`` (test_codec.py:29-136, :190-228).
"""
from __future__ import annotations

import json
import struct
import zlib

import numpy as np

MAGIC = b"SSPL1"
SCALE_LO = float(np.log(0.35))  # codec.py:37, the fitter's log-scale clamp


def pack_dtype(bits: int) -> np.dtype:
    """codec.py:102-103."""
    return np.dtype(np.uint8 if bits <= 8 else (np.uint16 if bits <= 16 else np.uint32))


def quant(x, lo, hi, bits: int) -> np.ndarray:
    """codec.py:80-83."""
    levels = (1 << bits) - 1
    t = np.clip((x - lo) / np.maximum(np.asarray(hi) - lo, 1e-12), 0.0, 1.0)
    return np.round(t * levels).astype(np.uint32)


def dequant(q, lo, hi, bits: int) -> np.ndarray:
    """codec.py:86-88: float64 arithmetic, float32 result."""
    levels = (1 << bits) - 1
    return (lo + (q.astype(np.float64) / levels) * (np.asarray(hi) - lo)).astype(np.float32)


def quant_circular(x, period: float, bits: int) -> np.ndarray:
    """codec.py:91-94."""
    levels = 1 << bits
    t = np.mod(x, period) / period
    return np.mod(np.round(t * levels).astype(np.int64), levels).astype(np.uint32)


def dequant_circular(q, period: float, bits: int) -> np.ndarray:
    """codec.py:97-99."""
    levels = 1 << bits
    return ((q.astype(np.float64) / levels) * period).astype(np.float32)


def pack(q, bits: int, byte_planar: bool = False) -> bytes:
    """codec.py:106-121."""
    dtype = pack_dtype(bits)
    raw = q.astype(dtype).tobytes()
    if not byte_planar or dtype.itemsize == 1:
        return raw
    return np.frombuffer(raw, dtype=np.uint8).reshape(-1, dtype.itemsize).T.ravel().tobytes()


def unpack(raw: bytes, bits: int, count: int, byte_planar: bool = False) -> np.ndarray:
    """codec.py:110-131."""
    dtype = pack_dtype(bits)
    if not byte_planar or dtype.itemsize == 1:
        return np.frombuffer(raw, dtype=dtype, count=count).astype(np.uint32)
    planes = np.frombuffer(raw, dtype=np.uint8, count=count * dtype.itemsize)
    interleaved = np.ascontiguousarray(planes.reshape(dtype.itemsize, count).T)
    return np.frombuffer(interleaved.tobytes(), dtype=dtype, count=count).astype(np.uint32)


def morton_order(x, y, H: int, W: int) -> np.ndarray:
    """codec.py:69-77: stable sort by a 10-bit-per-axis Z-order code."""
    xi = np.clip((x / max(W - 1, 1) * 1023).round().astype(np.uint32), 0, 1023).astype(np.uint64)
    yi = np.clip((y / max(H - 1, 1) * 1023).round().astype(np.uint32), 0, 1023).astype(np.uint64)
    code = np.zeros_like(xi)
    for b in range(10):
        code |= ((xi >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
        code |= ((yi >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b + 1)
    return np.argsort(code, kind="stable")


def means_extent(means, H: int, W: int):
    """codec.py:212-222: the image box unioned with the means' range."""
    lo = [float(min(0.0, np.floor(means[:, 0].min()))),
          float(min(0.0, np.floor(means[:, 1].min())))]
    hi = [float(max(W - 1.0, np.ceil(means[:, 0].max()))),
          float(max(H - 1.0, np.ceil(means[:, 1].max())))]
    return lo, hi


def parse(blob: bytes):
    """(header dict, [raw stream bytes]) of a blob (codec.py:333-352, 429-479 framing checks)."""
    if len(blob) < 9 or blob[:5] != MAGIC:
        raise ValueError("not a structsplat codec blob")
    (hlen,) = struct.unpack_from("<I", blob, 5)
    end = 9 + hlen
    if end > len(blob):
        raise ValueError("truncated structsplat codec header")
    header = json.loads(blob[9:end])
    count = 5 if header.get("has_opacity", False) else 4
    raw_lengths = header.get("stream_raw_lengths", [])
    codec = header.get("stream_codec", "zlib")
    off = end
    raws = []
    for i in range(count):
        if off + 4 > len(blob):
            raise ValueError("truncated stream framing")
        (plen,) = struct.unpack_from("<I", blob, off)
        off += 4
        if off + plen > len(blob):
            raise ValueError("truncated stream payload")
        payload = blob[off:off + plen]
        off += plen
        if codec != "zlib":
            raise ValueError(f"stream codec {codec!r} is not restated here")
        raw = zlib.decompress(payload)
        if i < len(raw_lengths) and len(raw) != int(raw_lengths[i]):
            raise ValueError("stream length differs from the header's")
        raws.append(raw)
    return header, raws


def quantized(header, raws):
    """The quantized integer columns after un-packing and un-delta-coding (codec.py:354-371)."""
    n = int(header["n"])
    b_means, b_scales, b_rot, b_colors = header["bits"]
    q_means = unpack(raws[0], b_means, 2 * n,
                     bool(header.get("means_byte_planar", False))).reshape(2, n).T
    if header["morton"]:
        q_means = np.cumsum(q_means, axis=0, dtype=np.uint64).astype(np.uint32)
        q_means &= (1 << b_means) - 1
    q_scales = unpack(raws[1], b_scales, 2 * n).reshape(2, n).T
    q_rot = unpack(raws[2], b_rot, n)
    q_colors = unpack(raws[3], b_colors, 3 * n).reshape(3, n).T
    if header.get("color_delta", False):
        q_colors = np.cumsum(q_colors, axis=0, dtype=np.uint64).astype(np.uint32)
        q_colors &= (1 << b_colors) - 1
    q_op = None
    if header.get("has_opacity", False):
        q_op = unpack(raws[4], int(header["bits_opacity"]), n)
    return q_means, q_scales, q_rot, q_colors, q_op


def decode_arrays(blob: bytes) -> dict:
    """Field arrays of a blob: means (n,2), log_scales (n,2), rotations (n,), colors (n,3),
    opacity logits (n,) or None, all float32 (codec.py:333-391)."""
    header, raws = parse(blob)
    H, W = int(header["H"]), int(header["W"])
    b_means, b_scales, b_rot, b_colors = header["bits"]
    q_means, q_scales, q_rot, q_colors, q_op = quantized(header, raws)
    means = dequant(q_means, header.get("means_lo", [0.0, 0.0]),
                    header.get("means_hi", [W - 1.0, H - 1.0]), b_means)
    big = float(np.log(max(H, W)))
    log_scales = dequant(q_scales, header.get("scale_lo", [SCALE_LO, SCALE_LO]),
                         header.get("scale_hi", [big, big]), b_scales)
    if header.get("rot_circular", False):
        rotations = dequant_circular(q_rot, np.pi, b_rot)
    else:
        rotations = dequant(q_rot, 0.0, np.pi, b_rot)
    colors = dequant(q_colors, np.asarray(header["color_lo"]), np.asarray(header["color_hi"]),
                     b_colors)
    logits = None
    if q_op is not None:
        b_op = int(header["bits_opacity"])
        p = dequant(q_op, 0.0, 1.0, b_op).clip(1e-4, 1.0 - 1e-4)
        logits = np.log(p / (1.0 - p)).astype(np.float32)
    return {"header": header, "means": means, "log_scales": log_scales, "rotations": rotations,
            "colors": colors, "opacity_logits": logits}


def encode_arrays(means, log_scales, rotations, colors, H: int, W: int, opacity_p=None,
                  bits=(16, 8, 8, 8), bits_opacity: int = 8, morton: bool = True,
                  color_delta: bool = True, means_byte_planar: bool = True,
                  rot_circular: bool = True, color_range=None, scale_range=None,
                  render=None, zlib_level: int = 9, keep_order: bool = False,
                  means_range=None) -> tuple[bytes, list]:
    """Blob of float64 field arrays (codec.py:235-330); returns (blob, raw streams).

    ``keep_order`` skips the Morton sort (the arrays are already in the encoded order, e.g. a
    decoded blob's) and ``means_range`` replaces the data-derived means extent, so a decoded
    blob re-encodes to its own streams."""
    means = np.asarray(means, np.float64)
    log_scales = np.asarray(log_scales, np.float64)
    theta = np.mod(np.asarray(rotations, np.float64), np.pi)
    colors = np.asarray(colors, np.float64)
    n = means.shape[0]
    clo, chi = ((colors.min(axis=0), colors.max(axis=0)) if color_range is None
                else map(np.asarray, color_range))
    slo, shi = ((log_scales.min(axis=0), log_scales.max(axis=0)) if scale_range is None
                else map(np.asarray, scale_range))
    clo, chi, slo, shi = (np.asarray(v, np.float64) for v in (clo, chi, slo, shi))
    means_lo, means_hi = (means_extent(means, H, W) if means_range is None
                          else (list(means_range[0]), list(means_range[1])))
    if morton and not keep_order:
        order = morton_order(means[:, 0], means[:, 1], H, W)
        means, log_scales, theta, colors = (a[order] for a in (means, log_scales, theta, colors))
        if opacity_p is not None:
            opacity_p = np.asarray(opacity_p, np.float64)[order]
    b_means, b_scales, b_rot, b_colors = bits
    q_means = quant(means, means_lo, means_hi, b_means)
    if morton:
        q_means = np.diff(q_means, axis=0, prepend=np.zeros((1, 2), np.uint32)).astype(np.uint32)
        q_means &= (1 << b_means) - 1
    q_scales = quant(log_scales, slo, shi, b_scales)
    q_rot = (quant_circular(theta, np.pi, b_rot) if rot_circular
             else quant(theta, 0.0, np.pi, b_rot))
    q_colors = quant(colors, clo, chi, b_colors)
    color_delta = bool(color_delta and morton)
    if color_delta:
        q_colors = np.diff(q_colors, axis=0, prepend=np.zeros((1, 3), np.uint32)).astype(np.uint32)
        q_colors &= (1 << b_colors) - 1
    streams = [
        pack(q_means.T.ravel(), b_means, means_byte_planar),
        pack(q_scales.T.ravel(), b_scales),
        pack(q_rot, b_rot),
        pack(q_colors.T.ravel(), b_colors),
    ]
    if opacity_p is not None:
        streams.append(pack(quant(np.asarray(opacity_p, np.float64), 0.0, 1.0, bits_opacity),
                            bits_opacity))
    render = render or {}
    header = {
        "n": int(n), "H": int(H), "W": int(W), "bits": list(bits), "morton": bool(morton),
        "color_delta": color_delta, "means_byte_planar": bool(means_byte_planar),
        "rot_circular": bool(rot_circular), "stream_codec": "zlib",
        "stream_raw_lengths": [len(s) for s in streams],
        "color_lo": clo.tolist(), "color_hi": chi.tolist(),
        "scale_lo": slo.tolist(), "scale_hi": shi.tolist(),
        "means_lo": means_lo, "means_hi": means_hi,
        "has_opacity": opacity_p is not None, "bits_opacity": int(bits_opacity),
        "renderer": render.get("renderer", "cuda"),
        "aa_dilation": float(render.get("aa_dilation", 0.0)),
        "sigma_cutoff": float(render.get("sigma_cutoff", 3.0)),
        "support_fade": bool(render.get("support_fade", False)),
        "render_chunk": int(render.get("render_chunk", 512)),
    }
    if "normalization_eps" in render:
        header["normalization_eps"] = float(render["normalization_eps"])
    hdr = json.dumps(header).encode()
    blob = MAGIC + struct.pack("<I", len(hdr)) + hdr
    for raw in streams:
        payload = zlib.compress(raw, zlib_level)
        blob += struct.pack("<I", len(payload)) + payload
    return blob, streams
