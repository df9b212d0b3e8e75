"""The output writers (utils.py:253-260): PNG checked by an independent
decoder (PIL), OpenEXR by a reader written from the OpenEXR 2 file layout
(no EXR library is in the image: the EXR encoder's parity with cv2.imwrite is
unpinned beyond that layout)."""
import io
import os
import struct

import numpy as np
import pytest
from PIL import Image

from sphere_stereo.io import encode_exr, encode_png, save_rgbd_panorama


def test_png_roundtrip_through_pil():
    rng = np.random.RandomState(0)
    bgr = rng.randint(0, 256, (37, 53, 3)).astype(np.uint8)
    got = np.array(Image.open(io.BytesIO(encode_png(bgr))))
    assert got.shape == (37, 53, 3)
    np.testing.assert_array_equal(got, bgr[:, :, ::-1])  # cv2 stores BGR arrays as RGB
    gray = rng.randint(0, 256, (5, 7)).astype(np.uint8)
    np.testing.assert_array_equal(np.array(Image.open(io.BytesIO(encode_png(gray)))), gray)
    with pytest.raises(TypeError):
        encode_png(bgr.astype(np.float32))


def read_exr(data):
    """A reader of single-part scanline OpenEXR files without compression,
    from the file layout: magic 20000630, version 2, attributes (name, type,
    size, value) up to an empty name, a table of one uint64 offset per
    block, blocks of (int32 y, int32 size, channels in header order)."""
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630 and version == 2
    pos, attrs = 8, {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        tend = data.index(b"\0", end + 1)
        kind = data[end + 1:tend].decode()
        size = struct.unpack_from("<i", data, tend + 1)[0]
        attrs[name] = (kind, data[tend + 5:tend + 5 + size])
        pos = tend + 5 + size
    pos += 1
    assert attrs["compression"] == ("compression", b"\0")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chans, c = [], attrs["channels"][1]
    while c[0] != 0:
        end = c.index(b"\0")
        ptype, _, xs, ys = struct.unpack_from("<iB3xii", c, end + 1)
        assert ptype == 2 and xs == ys == 1  # FLOAT, full resolution
        chans.append(c[:end].decode())
        c = c[end + 17:]
    offsets = struct.unpack_from("<%dQ" % h, data, pos)
    out = {n: np.empty((h, w), np.float32) for n in chans}
    for off in offsets:
        y, size = struct.unpack_from("<ii", data, off)
        assert size == 4 * w * len(chans)
        vals = np.frombuffer(data, "<f4", w * len(chans), off + 8).reshape(len(chans), w)
        for i, n in enumerate(chans):
            out[n][y - y0] = vals[i]
    return out


def test_exr_layout_and_roundtrip():
    rng = np.random.RandomState(1)
    inv = (rng.rand(19, 31).astype(np.float32) * 2).astype(np.float32)
    inv[3, 4] = np.inf
    inv[5, 6] = np.nan
    inv[0, 0] = -0.0
    data = encode_exr(inv)
    # the fixed start of the file: magic, version, then the channel list of "Y"
    assert data[:8] == bytes([0x76, 0x2F, 0x31, 0x01, 2, 0, 0, 0])
    assert data[8:8 + 16] == b"channels\0chlist\0"
    assert struct.unpack_from("<i", data, 24)[0] == 2 + 16 + 1
    got = read_exr(data)
    assert list(got) == ["Y"]
    np.testing.assert_array_equal(got["Y"].view(np.uint32), inv.view(np.uint32))
    rgb = rng.rand(4, 6, 3).astype(np.float32)
    got3 = read_exr(encode_exr(rgb))
    assert sorted(got3) == ["B", "G", "R"]
    for i, n in enumerate("BGR"):
        np.testing.assert_array_equal(got3[n], rgb[:, :, i])


def test_save_rgbd_panorama(tmp_path):
    rng = np.random.RandomState(2)
    pano = {"rgb": rng.randint(0, 256, (8, 16, 3)).astype(np.uint8),
            "inv_distance": rng.rand(8, 16).astype(np.float32)}
    os.makedirs(tmp_path / "output")
    save_rgbd_panorama({"frame_0001.png": pano}, "frame_0001.png", str(tmp_path))
    rgb = np.array(Image.open(tmp_path / "output" / "rgb_frame_0001.png"))
    np.testing.assert_array_equal(rgb, pano["rgb"][:, :, ::-1])
    inv = read_exr((tmp_path / "output" / "inv_distance_frame_0001.exr").read_bytes())["Y"]
    np.testing.assert_array_equal(inv, pano["inv_distance"])
    save_rgbd_panorama({}, "missing.png", str(tmp_path))  # a missing frame is skipped
    assert sorted(os.listdir(tmp_path / "output")) == ["inv_distance_frame_0001.exr", "rgb_frame_0001.png"]


def test_decode_exr_reads_the_encoder_and_refuses_others():
    """io.decode_exr (the ground-truth reader of the dataset loop) reads what
    encode_exr writes, bit for bit, and refuses other layouts."""
    from sphere_stereo.io import decode_exr

    inv = np.random.RandomState(3).rand(7, 9).astype(np.float32)
    np.testing.assert_array_equal(decode_exr(encode_exr(inv)), inv)
    assert decode_exr(encode_exr(np.zeros((2, 3, 3), np.float32))) is None  # B, G, R channels
    assert decode_exr(b"not an exr") is None


def _exr_blocks(data):
    """(first scanline, stored size) of each block of an EXR file whose
    header ends with the screenWindowWidth attribute (as the writers here
    order it)."""
    key = b"screenWindowWidth\0float\0"
    table = data.index(key) + len(key) + 4 + 4 + 1  # size, value, end of header
    out, pos = [], table
    while True:
        off = struct.unpack_from("<Q", data, pos)[0]
        out.append(struct.unpack_from("<ii", data, off))
        pos += 8
        if off + 8 + out[-1][1] == len(data):
            return out


def test_decode_exr_zip_and_zips():
    """ZIP (16 scanlines per block, cv2.imwrite's default) and ZIPS (one)
    blocks, written by a loop-based writer of the OpenEXR layout
    (helpers.encode_exr_zip), decode bit for bit: a height that is not a
    multiple of 16, non-finite values and -0.0, an image of few distinct
    values (every block compresses, so the predictor and the byte split are
    exercised), a constant one and one of random bits (blocks that do not
    compress are stored raw).  No EXR library is in the image, so the match
    with cv2's own files is unpinned beyond that layout."""
    from helpers import encode_exr_zip
    from sphere_stereo.io import decode_exr

    rng = np.random.RandomState(5)
    steps = (np.floor(np.cumsum(rng.rand(37, 21), axis=1) * 4) / 4).astype(np.float32)
    steps[2, 3], steps[30, 1], steps[36, 20], steps[0, 0] = np.nan, np.inf, -np.inf, -0.0
    noise = rng.randint(0, 2 ** 32, (19, 11), dtype=np.uint64).astype(np.uint32).view(np.float32)
    const = np.full((17, 5), 0.25, np.float32)
    for img, packed in ((steps, True), (const, True), (noise, False)):
        for lines in (16, 1):
            data = encode_exr_zip(img, lines)
            blocks = _exr_blocks(data)
            assert [y for y, _ in blocks] == list(range(0, img.shape[0], lines))
            raw = [4 * img.shape[1] * min(lines, img.shape[0] - y) for y, _ in blocks]
            assert all((size < r) == packed for (_, size), r in zip(blocks, raw))
            got = decode_exr(data)
            assert got is not None and got.shape == img.shape
            np.testing.assert_array_equal(got.view(np.uint32), img.view(np.uint32))


def test_dataset_loop_reader_and_arguments(tmp_path):
    """sphere_stereo.main's host side: images decoded as cv2.IMREAD_UNCHANGED
    gives them (BGR, greyscale, BGRA kept 4-channel so the frame's shape check
    refuses it, unreadable files None) and main.py's arguments."""
    from sphere_stereo import main as ss_main

    rng = np.random.RandomState(4)
    bgr = rng.randint(0, 256, (6, 5, 3)).astype(np.uint8)
    (tmp_path / "c.png").write_bytes(encode_png(bgr))
    np.testing.assert_array_equal(ss_main.read_image(str(tmp_path / "c.png")), bgr)
    gray = rng.randint(0, 256, (4, 7)).astype(np.uint8)
    (tmp_path / "g.png").write_bytes(encode_png(gray))
    np.testing.assert_array_equal(ss_main.read_image(str(tmp_path / "g.png")), gray)
    rgba = rng.randint(0, 256, (3, 4, 4)).astype(np.uint8)
    Image.fromarray(rgba, "RGBA").save(tmp_path / "a.png")
    np.testing.assert_array_equal(ss_main.read_image(str(tmp_path / "a.png")), rgba[:, :, [2, 1, 0, 3]])
    # 16-bit colour: PIL would keep only the high bytes; refused with a warning
    # (16-bit greyscale decodes as uint16, as cv2.IMREAD_UNCHANGED gives it)
    from helpers import encode_png16
    rgb16 = rng.randint(0, 65536, (3, 5, 3)).astype(np.uint16)
    (tmp_path / "c16.png").write_bytes(encode_png16(rgb16))
    assert np.array_equal(np.array(Image.open(tmp_path / "c16.png")), (rgb16 >> 8).astype(np.uint8))
    with pytest.warns(UserWarning, match="16-bit colour PNG"):
        assert ss_main.read_image(str(tmp_path / "c16.png")) is None
    g16 = rng.randint(0, 65536, (4, 6)).astype(np.uint16)
    Image.fromarray(g16).save(tmp_path / "g16.png")
    got16 = ss_main.read_image(str(tmp_path / "g16.png"))
    assert got16.dtype == np.uint16
    np.testing.assert_array_equal(got16, g16)
    (tmp_path / "x.png").write_bytes(b"not an image")
    assert ss_main.read_image(str(tmp_path / "x.png")) is None
    assert ss_main.read_image(str(tmp_path / "missing.png")) is None
    a = ss_main.parse_args([])
    assert (a.references_indices, a.candidate_count, a.matching_resolution, a.panorama_resolution, a.sigma_s) == \
        ([0, 2], 32, [1024, 1024], [2048, 1024], 25)
    assert a.saving and not a.evaluate and a.bad_px_ratio_thresholds == [0.1, 0.4]
    assert not ss_main.parse_args(["--saving", "False"]).saving


def test_psnr_closed_forms():
    """evaluation.psnr (utils.py:232): one channel of one pixel off by 255
    among n values gives 10 log10(n) dB; equal images give infinity; random
    images agree with the float64 formula."""
    from sphere_stereo.evaluation import psnr

    a = np.zeros((4, 5, 3), np.uint8)
    b = a.copy()
    b[1, 2, 0] = 255
    assert abs(psnr(a, b) - 10 * np.log10(60)) < 1e-12
    assert psnr(a, a) == float("inf")
    rng = np.random.RandomState(6)
    x, y = (rng.randint(0, 256, (9, 7, 3)).astype(np.uint8) for _ in range(2))
    ref = 10 * np.log10(1 / np.mean((x / 255.0 - y / 255.0) ** 2))
    assert abs(psnr(x, y) - ref) < 1e-5
    with pytest.raises(ValueError):
        psnr(a, b[:2])
