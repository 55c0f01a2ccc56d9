"""Pins oracle.codec (the SSPL1 restatement) and checks the product's host-side container
parsing (structsplat_amd.codec: blob_header, blob_components, decode_alpha).

The real blobs under tests/golden/sspl1 were written by the reference's encoder (its COMP-009
evidence inputs: 8192 fitted Gaussians on 768x512 images, bit tuples (16, 8, 8, 8) and
(12, 6, 6, 8), Morton + color-delta + byte-planar + circular rotation).  Decoding them and
re-encoding the decoded arrays in the decoded order reproduces every raw stream byte for byte.
"""
from __future__ import annotations

import glob
import os
import struct

import numpy as np
import pytest

from oracle import codec as C
from structsplat_amd import codec as P

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "sspl1", "*.sspl1")))


def _field(n=300, H=48, W=64, seed=0, opacity=True):
    rng = np.random.default_rng(seed)
    means = np.stack([rng.uniform(-3, W + 2, n), rng.uniform(-2, H + 3, n)], 1)
    log_scales = rng.uniform(np.log(0.4), np.log(8.0), (n, 2))
    rot = rng.uniform(-4.0, 7.0, n)
    colors = rng.normal(0.4, 1.5, (n, 3))
    op = rng.uniform(0.0, 1.0, n) if opacity else None
    return means, log_scales, rot, colors, op


def test_golden_blobs_exist():
    assert len(GOLDEN) >= 3


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_reference_blobs_reencode_to_their_own_streams(path):
    blob = open(path, "rb").read()
    h, raws = C.parse(blob)
    d = C.decode_arrays(blob)
    _, streams = C.encode_arrays(
        d["means"], d["log_scales"], d["rotations"], d["colors"], h["H"], h["W"],
        bits=tuple(h["bits"]), morton=h["morton"], color_delta=h["color_delta"],
        means_byte_planar=h["means_byte_planar"], rot_circular=h["rot_circular"],
        color_range=(h["color_lo"], h["color_hi"]), scale_range=(h["scale_lo"], h["scale_hi"]),
        keep_order=True, means_range=(h["means_lo"], h["means_hi"]))
    assert [len(r) for r in raws] == h["stream_raw_lengths"]
    for k, (a, b) in enumerate(zip(raws, streams)):
        assert a == b, k
    # Decoded values lie in the header's ranges; the rotations in [0, pi).
    assert (d["means"] >= np.float32(h["means_lo"]) - 1e-3).all()
    assert (d["means"] <= np.float32(h["means_hi"]) + 1e-3).all()
    assert (d["colors"] >= np.float32(h["color_lo"]) - 1e-5).all()
    assert (d["colors"] <= np.float32(h["color_hi"]) + 1e-5).all()
    assert (d["rotations"] >= 0).all() and (d["rotations"] < np.pi).all()


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_blob_components_partition_reference_blobs(path):
    """test_codec.py:116-127 on the reference's own blobs."""
    blob = open(path, "rb").read()
    parts = P.blob_components(blob)
    assert parts["total_bytes"] == len(blob)
    assert parts["header_bytes"] + sum(s["framed_bytes"] for s in parts["streams"].values()) \
        == len(blob)
    assert list(parts["streams"]) == ["means", "scales", "rotation", "colors"]
    assert P.blob_header(blob) == C.parse(blob)[0]
    assert P.decode_alpha(blob) is None


def test_blob_components_rejects_truncated_or_trailing_bytes():
    """test_codec.py:129-135."""
    blob, _ = C.encode_arrays(*_field(16)[:4], 48, 64)
    with pytest.raises(ValueError, match="truncated"):
        P.blob_components(blob[:-1])
    with pytest.raises(ValueError, match="trailing"):
        P.blob_components(blob + b"x")
    with pytest.raises(ValueError, match="not a structsplat codec blob"):
        P.blob_components(b"XXXXX" + blob[5:])


def test_means_error_bounded_by_one_lattice_step():
    """test_codec.py:29-45 and :190-208 (off-image means): without reordering, every decoded
    mean is within one lattice step of the original."""
    means, ls, rot, col, _ = _field(400)
    blob, _ = C.encode_arrays(means, ls, rot, col, 48, 64, morton=False)
    h, _ = C.parse(blob)
    d = C.decode_arrays(blob)
    step = (np.asarray(h["means_hi"]) - np.asarray(h["means_lo"])) / (2**16 - 1)
    assert (np.abs(d["means"] - means) <= step + 1e-4).all()
    assert h["means_lo"][0] <= means[:, 0].min() and h["means_hi"][1] >= means[:, 1].max()


def test_morton_reorder_changes_order_not_content():
    """test_codec.py:57-66: the same multiset of Gaussians either way."""
    means, ls, rot, col, _ = _field(300)
    a = C.decode_arrays(C.encode_arrays(means, ls, rot, col, 48, 64, morton=True)[0])
    b = C.decode_arrays(C.encode_arrays(means, ls, rot, col, 48, 64, morton=False)[0])
    key = lambda d: np.lexsort((d["colors"][:, 0], d["means"][:, 1], d["means"][:, 0]))
    for name in ("means", "log_scales", "rotations", "colors"):
        assert np.array_equal(a[name][key(a)], b[name][key(b)]), name


def test_circular_rotation_quantization_wraps_top_bin_to_zero():
    """test_codec.py:80-89."""
    means, ls, _, col, _ = _field(16)
    rot = np.full(16, np.pi - 1e-4)
    d = C.decode_arrays(C.encode_arrays(means, ls, rot, col, 48, 64, bits=(16, 8, 4, 8),
                                        morton=False)[0])
    assert float(d["rotations"].max()) < np.pi / 16


def test_rotation_period_invariance():
    """test_codec.py:69-77: theta and theta + pi encode identically."""
    means, ls, rot, col, _ = _field(64)
    a = C.encode_arrays(means, ls, rot, col, 48, 64)[1]
    b = C.encode_arrays(means, ls, rot + np.pi, col, 48, 64)[1]
    assert a[2] == b[2]


def test_flags_are_self_describing_and_plain_streams_decode():
    """test_codec.py:92-113."""
    means, ls, rot, col, _ = _field(48)
    d_blob = C.encode_arrays(means, ls, rot, col, 48, 64)[0]
    p_blob = C.encode_arrays(means, ls, rot, col, 48, 64, morton=False, color_delta=False,
                             means_byte_planar=False, rot_circular=False)[0]
    hd, hp = P.blob_header(d_blob), P.blob_header(p_blob)
    assert hd["color_delta"] and hd["means_byte_planar"] and hd["rot_circular"]
    assert not (hp["color_delta"] or hp["means_byte_planar"] or hp["rot_circular"])
    assert np.isfinite(C.decode_arrays(p_blob)["colors"]).all()


def test_opacity_roundtrip_to_logits():
    """test_codec.py:155-176: probabilities over [0, 1], logits after the clip."""
    means, ls, rot, col, op = _field(200)
    op[:3] = [0.0, 1.0, 0.5]
    d = C.decode_arrays(C.encode_arrays(means, ls, rot, col, 48, 64, opacity_p=op,
                                        morton=False)[0])
    p = 1.0 / (1.0 + np.exp(-d["opacity_logits"].astype(np.float64)))
    assert np.abs(p - np.clip(op, 1e-4, 1 - 1e-4)).max() <= 0.5 / 255 + 1e-6
    assert np.isfinite(d["opacity_logits"]).all()


def test_byte_planar_packing_is_a_transpose_of_little_endian_bytes():
    q = np.array([0x0102, 0xA0B0, 0xFFFF, 7], dtype=np.uint32)
    raw = C.pack(q, 16, byte_planar=True)
    assert raw == bytes([0x02, 0xB0, 0xFF, 0x07, 0x01, 0xA0, 0xFF, 0x00])
    assert np.array_equal(C.unpack(raw, 16, 4, byte_planar=True), q)
    q32 = np.array([0x01020304, 0xFFFFFFFF], dtype=np.uint32)
    assert np.array_equal(C.unpack(C.pack(q32, 20, True), 20, 2, True), q32)


def test_delta_coding_wraps_modulo_the_bit_width():
    """Decreasing values along the Morton order delta-code to wrapped differences."""
    means = np.array([[60.0, 40.0], [0.5, 0.5], [30.0, 1.0]])
    ls = np.zeros((3, 2))
    col = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.5, 0.5, 0.5]])
    blob, streams = C.encode_arrays(means, ls, np.zeros(3), col, 48, 64, bits=(8, 8, 8, 8),
                                    means_byte_planar=False)
    deltas = np.frombuffer(streams[0], np.uint8)
    assert (deltas.astype(np.int64) < 256).all()
    d = C.decode_arrays(blob)
    order = C.morton_order(means[:, 0], means[:, 1], 48, 64)
    h, _ = C.parse(blob)
    step = (np.asarray(h["means_hi"]) - np.asarray(h["means_lo"])) / 255
    assert (np.abs(d["means"] - means[order]) <= step).all()


def test_product_header_parse_matches_framing():
    blob, streams = C.encode_arrays(*_field(40)[:4], 48, 64)
    parts = P.blob_components(blob)
    assert [s["raw_bytes"] for s in parts["streams"].values()] == [len(s) for s in streams]
    (hlen,) = struct.unpack_from("<I", blob, 5)
    assert parts["header_json_bytes"] == hlen
