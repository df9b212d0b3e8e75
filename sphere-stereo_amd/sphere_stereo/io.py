"""Output files of the frame loop: save_rgbd_panorama (utils.py:253-260) writes
a panorama's colour as an 8-bit PNG and its inverse distance as an OpenEXR
image, as the reference does through cv2.imwrite.  cv2 is not a dependency
here: the two encoders below write the formats directly (PNG: RFC 2083
chunks, zlib-compressed scanlines with filter type 0; OpenEXR 2: one part,
scanline blocks of one line, no compression, 32-bit float channels).  Host
code: the files are written from host copies of the outputs.
"""
import os
import struct
import zlib

import numpy as np


def _host(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(a)


def _png_chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def encode_png(img, bgr=True):
    """PNG bytes of an H x W x 3 (or H x W) uint8 image.  With bgr, the
    channels are in cv2's BGR order, which cv2.imwrite stores as RGB."""
    a = _host(img)
    if a.dtype != np.uint8:
        raise TypeError("PNG encoder expects uint8, got %s" % a.dtype)
    if a.ndim == 2:
        color, rows = 0, a
    elif a.ndim == 3 and a.shape[2] == 3:
        color, rows = 2, (a[:, :, ::-1] if bgr else a)
    else:
        raise ValueError("PNG encoder expects H x W or H x W x 3, got %s" % (a.shape,))
    h, w = a.shape[:2]
    raw = np.concatenate([np.zeros((h, 1), np.uint8), np.ascontiguousarray(rows).reshape(h, -1)], axis=1)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, color, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", ihdr) + _png_chunk(b"IDAT", zlib.compress(raw.tobytes(), 6)) +
            _png_chunk(b"IEND", b""))


def _exr_attr(name, kind, value):
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value


def encode_exr(img):
    """OpenEXR bytes of an H x W (channel "Y", as cv2 names a single channel)
    or H x W x 3 (channels B, G, R) float32 image."""
    a = _host(img).astype(np.float32, copy=False)
    if a.ndim == 2:
        chans = {"Y": a}
    elif a.ndim == 3 and a.shape[2] == 3:
        chans = {"B": a[:, :, 0], "G": a[:, :, 1], "R": a[:, :, 2]}
    else:
        raise ValueError("EXR encoder expects H x W or H x W x 3, got %s" % (a.shape,))
    h, w = a.shape[:2]
    names = sorted(chans)
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", 2, 0, 1, 1) for n in names) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = (_exr_attr("channels", "chlist", chlist) + _exr_attr("compression", "compression", b"\0") +
              _exr_attr("dataWindow", "box2i", box) + _exr_attr("displayWindow", "box2i", box) +
              _exr_attr("lineOrder", "lineOrder", b"\0") + _exr_attr("pixelAspectRatio", "float", struct.pack("<f", 1)) +
              _exr_attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0)) +
              _exr_attr("screenWindowWidth", "float", struct.pack("<f", 1)) + b"\0")
    head = struct.pack("<ii", 20000630, 2) + header
    line = 4 * w * len(names)
    first = len(head) + 8 * h  # after the offset table
    offsets = struct.pack("<%dQ" % h, *[first + y * (8 + line) for y in range(h)])
    planes = np.stack([np.ascontiguousarray(chans[n]) for n in names], axis=1).astype("<f4")  # H x C x W
    blocks = b"".join(struct.pack("<ii", y, line) + planes[y].tobytes() for y in range(h))
    return head + offsets + blocks


def png_bit_depth(data):
    """(bit depth, colour type) from a PNG file's IHDR chunk, or None when the
    bytes are not a PNG."""
    if len(data) < 33 or data[:8] != b"\x89PNG\r\n\x1a\n" or data[12:16] != b"IHDR":
        return None
    return data[24], data[25]


# OpenEXR compression codes: none, ZIPS (one scanline per block), ZIP (16)
_EXR_LINES_PER_BLOCK = {0: 1, 2: 1, 3: 16}


def _exr_unzip(block, size):
    """OpenEXR ZIP / ZIPS block data: zlib, then the byte-delta predictor
    (t[i] = t[i-1] + d[i] - 128 mod 256) and the split of even and odd bytes
    into two halves undone (OpenEXR file layout, "ZIP compression")."""
    d = np.frombuffer(zlib.decompress(block), np.uint8).astype(np.int64)
    if d.size != size:
        raise ValueError("EXR block decompresses to %d bytes, %d expected" % (d.size, size))
    if d.size:
        d[1:] -= 128
        d = np.cumsum(d) & 255
    t = d.astype(np.uint8)
    out = np.empty_like(t)
    half = (t.size + 1) // 2
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def decode_exr(data):
    """Inverse-distance ground truth: the Y channel of a single-part scanline
    OpenEXR file with 32-bit float pixels, uncompressed (as encode_exr writes)
    or ZIP / ZIPS compressed (cv2.imwrite's default for EXR); None for any
    other EXR."""
    if len(data) < 8 or struct.unpack_from("<ii", data, 0) != (20000630, 2):
        return None
    pos, attrs = 8, {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        tend = data.index(b"\0", end + 1)
        size = struct.unpack_from("<i", data, tend + 1)[0]
        attrs[data[pos:end].decode()] = data[tend + 5:tend + 5 + size]
        pos = tend + 5 + size
    pos += 1
    comp = attrs.get("compression", b"\xff")[0]
    if comp not in _EXR_LINES_PER_BLOCK or attrs.get("channels", b"")[:2] != b"Y\0":
        return None
    if struct.unpack_from("<i", attrs["channels"], 2)[0] != 2 or attrs["channels"][18:19] != b"\0":
        return None
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lines = _EXR_LINES_PER_BLOCK[comp]
    nblocks = (h + lines - 1) // lines
    out = np.empty((h, w), np.float32)
    for off in struct.unpack_from("<%dQ" % nblocks, data, pos):
        y, size = struct.unpack_from("<ii", data, off)
        n = min(lines, y0 + h - y)
        raw = data[off + 8:off + 8 + size]
        if comp != 0 and size < 4 * w * n:  # a block stored uncompressed when that is not larger
            raw = _exr_unzip(raw, 4 * w * n)
        out[y - y0:y - y0 + n] = np.frombuffer(raw, "<f4", w * n).reshape(n, w)
    return out


def save_rgbd_panorama(rgbd_panoramas, filename, dataset_path):
    """utils.py:253-260: output/rgb_<name>.png and output/inv_distance_<name>.exr
    under dataset_path for rgbd_panoramas[filename] = {"rgb": BGR uint8,
    "inv_distance": float32}; a missing entry is skipped, as there.  The
    output directory must exist (main.py:128 creates it before the loop)."""
    try:
        pano = rgbd_panoramas[filename]
    except KeyError:
        return
    name = os.path.splitext(filename)[0]
    out = os.path.join(dataset_path, "output")
    with open(os.path.join(out, "rgb_" + name + ".png"), "wb") as f:
        f.write(encode_png(pano["rgb"]))
    with open(os.path.join(out, "inv_distance_" + name + ".exr"), "wb") as f:
        f.write(encode_exr(pano["inv_distance"]))
