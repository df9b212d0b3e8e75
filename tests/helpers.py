"""Shared set-up for the tests: small seeded synthetic frames."""
import functools

import numpy as np
import torch

from sphere_stereo import synth
from sphere_stereo.utils import parse_json_calib


@functools.lru_cache(maxsize=8)
def rendered(seed, resolution, baseline=0.26):
    rig = synth.SyntheticRig(seed=seed, resolution=resolution, baseline=baseline)
    imgs, masks = rig.render("cpu")
    return rig, imgs, masks


def frame(seed=0, resolution=160, matching=(80, 72), stitch=(96, 96), refs=(0, 2), device="cpu", baseline=0.26):
    """Calibrations, viewpoint, masks and images of one synthetic frame."""
    rig, imgs, masks = rendered(seed, resolution, baseline)
    calibs = parse_json_calib(rig.basalt_json()["value0"], list(matching), device)
    viewpoint = torch.zeros([3], device=device)
    for r in refs:
        viewpoint += calibs[r].rt[:3, 3] / len(refs)
    to_match, to_stitch, masks_m = synth.make_frame(rig, matching, stitch, refs, device=device,
                                                    rendered=(imgs, masks))
    return dict(rig=rig, calibs=calibs, viewpoint=viewpoint, masks=masks_m, to_match=to_match, to_stitch=to_stitch,
                matching=tuple(matching), stitch=tuple(stitch), refs=list(refs))


def to_np(t):
    return t.detach().cpu().numpy()


def first_mismatch(a, b):
    a, b = np.asarray(a), np.asarray(b)
    bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b))
    if bad.size == 0:
        return None
    i = tuple(bad[0])
    return "%d mismatches, first at %s: %r vs %r" % (len(bad), i, a[i], b[i])


# ---------------------------------------------------------------------------
# the reference's README demo (README.md:27-30): resources/ frames, masks and
# calibration, and the panorama the reference produced from them
# ---------------------------------------------------------------------------
import json  # noqa: E402
import os  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEMO = os.path.join(GOLDEN, "demo")
DEMO_EXPECTED = os.path.join(GOLDEN, "demo_expected_rgb_0.png")
# python main.py --dataset_path resources --references_indices 2 3, the rest
# main.py's defaults (main.py:35-51)
DEMO_ARGS = dict(min_dist=0.55, max_dist=100, candidate_count=32, references_indices=[2, 3],
                 matching_resolution=(1024, 1024), rgb_to_stitch_resolution=(1216, 1216),
                 panorama_resolution=(2048, 1024), sigma_i=10, sigma_s=25)


def demo_frame(oracle):
    """The demo frame as main.py:62-81 and read_input_images (utils.py:157-204)
    hand it to the estimator, on the CPU: frames and masks decoded with PIL
    (BGR order), the frames resized to the matching resolution with the
    oracle's INTER_LINEAR (the GPU input stage's bit-exact target), the masks
    resized the same way, rounded to uint8 and divided by 255 the way torch
    divides a GPU tensor by a Python number (main.py:79): a multiply by the
    float32 reciprocal.  Returns
    (calibrations, estimator keyword arguments, images_to_match,
    images_to_stitch)."""
    from sphere_stereo.main import read_image
    from sphere_stereo.utils import parse_json_calib

    a = DEMO_ARGS
    mc, mr = a["matching_resolution"]
    with open(os.path.join(DEMO, "calibration.json")) as f:
        calibs = parse_json_calib(json.load(f)["value0"], list(a["matching_resolution"]), "cpu")
    viewpoint = torch.zeros([3])
    for r in a["references_indices"]:
        viewpoint += calibs[r].rt[:3, 3] / len(a["references_indices"])
    masks = []
    for k in range(len(calibs)):
        m = read_image(os.path.join(DEMO, "cam%d" % k, "mask.png")).astype(np.float32)
        m = np.round(oracle.resize_linear(m, mc, mr)).clip(0, 255).astype(np.float32)
        m = m * (np.float32(1) / np.float32(255))
        masks.append(torch.from_numpy(m).unsqueeze(0))
    imgs = [read_image(os.path.join(DEMO, "cam%d" % k, "0.jpg")).astype(np.float32) for k in range(len(calibs))]
    sc, sr = a["rgb_to_stitch_resolution"]
    to_match = [oracle.resize_linear(i, mc, mr) for i in imgs]
    to_stitch = [oracle.resize_linear(imgs[r], sc, sr) for r in a["references_indices"]]
    kw = dict(a, reprojection_viewpoint=viewpoint, masks=masks)
    return calibs, kw, to_match, to_stitch


def psnr(a, b):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return 10 * np.log10(255.0 ** 2 / np.mean(d * d))


def cpu_args(calibs, kw):
    """CPU copies of a frame's calibrations and estimator keyword arguments,
    for the oracle."""
    cc = [type(c)(c.original_resolution, c.principal.cpu(), c.fl.cpu(), c.xi, c.alpha, c.rt.cpu(),
                  c.matching_scale.cpu()) for c in calibs]
    kc = dict(kw)
    kc["reprojection_viewpoint"] = kw["reprojection_viewpoint"].cpu()
    kc["masks"] = [m.cpu() for m in kw["masks"]]
    return cc, kc


def oracle_frame(calibs, kw, to_match, to_stitch, **extra):
    """The oracle's run of one frame: (pipeline, kept intermediates, rgb,
    distance)."""
    from pipeline import OraclePipeline

    cc, kc = cpu_args(calibs, kw)
    ora = OraclePipeline(cc, **kc, **extra)
    keep = {}
    rgb, dist = ora.estimate_RGBD_panorama([t.cpu() for t in to_match], [t.cpu() for t in to_stitch], keep=keep)
    return ora, keep, rgb, dist


def encode_exr_zip(img, lines=16):
    """A single-channel ("Y", float32) OpenEXR file with ZIP (lines=16) or ZIPS
    (lines=1) compression, as cv2.imwrite writes EXR by default, written from
    the OpenEXR 2 file layout with plain loops: per block of scanlines the
    bytes are split into the even-indexed ones followed by the odd-indexed
    ones, each byte after the first replaced by (t[i] - t[i-1] + 128) mod 256,
    and the result zlib-compressed (stored raw when that is not smaller)."""
    import struct
    import zlib

    a = np.ascontiguousarray(img, dtype="<f4")
    h, w = a.shape

    def attr(name, kind, value):
        return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value

    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    code = {16: 3, 1: 2}[lines]
    header = (attr("channels", "chlist", b"Y\0" + struct.pack("<iB3xii", 2, 0, 1, 1) + b"\0") +
              attr("compression", "compression", bytes([code])) + attr("dataWindow", "box2i", box) +
              attr("displayWindow", "box2i", box) + attr("lineOrder", "lineOrder", b"\0") +
              attr("pixelAspectRatio", "float", struct.pack("<f", 1)) +
              attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0)) +
              attr("screenWindowWidth", "float", struct.pack("<f", 1)) + b"\0")
    head = struct.pack("<ii", 20000630, 2) + header
    blocks = []
    for y in range(0, h, lines):
        raw = a[y:y + lines].tobytes()
        split = raw[0::2] + raw[1::2]
        pred = bytearray(split)
        for i in range(1, len(split)):
            pred[i] = (split[i] - split[i - 1] + 128) % 256
        packed = zlib.compress(bytes(pred), 6)
        if len(packed) >= len(raw):
            packed = raw
        blocks.append(struct.pack("<ii", y, len(packed)) + packed)
    offsets, pos = [], len(head) + 8 * len(blocks)
    for b in blocks:
        offsets.append(pos)
        pos += len(b)
    return head + struct.pack("<%dQ" % len(blocks), *offsets) + b"".join(blocks)


def encode_png16(rgb):
    """A 16-bit RGB PNG (colour type 2, filter type 0, big-endian samples)
    from an H x W x 3 uint16 array."""
    import struct
    import zlib

    h, w, _ = rgb.shape
    rows = np.ascontiguousarray(rgb.astype(">u2")).view(np.uint8).reshape(h, -1)
    raw = np.concatenate([np.zeros((h, 1), np.uint8), rows], axis=1).tobytes()

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 16, 2, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))
