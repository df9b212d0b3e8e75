"""The dataset loop of main.py (main.py:44-140) on this package:
``python -m sphere_stereo.main --dataset_path DIR [options]`` with the
reference's arguments.  DIR holds calibration.json (basalt, {"value0": ...}),
cam<k>/ folders of frames (and an optional cam<k>/mask.png); the panoramas
go to DIR/output/rgb_<frame>.png and inv_distance_<frame>.exr.

Differences, because cv2, skimage and a display are not available: images
are decoded with PIL (8-bit and 16-bit greyscale, 8-bit RGB/RGBA as
cv2.IMREAD_UNCHANGED gives them, RGB turned to cv2's BGR order; other files
count as unreadable, as cv2.imread's None does there; 16-bit colour PNGs,
whose low bytes PIL drops, are refused with a warning); a mask is resized in
float32 and rounded to uint8, where cv2 resizes uint8 in fixed point (parity
unpinned; on the README demo frame some mask pixels differ by one level and
the panorama's PSNR against the reference's output moves by 0.001 dB);
--visualize is ignored; --evaluate reports the inverse-distance MAE / RMSE /
bad-pixel ratios (utils.py:233-242, on the GPU) against
DIR/gt/inv_distance_<frame>.exr (single-channel float32 OpenEXR, uncompressed
or ZIP / ZIPS; resized to the panorama as utils.py:226-227 does), raises when
no frame has a readable one, adds the colour PSNR (utils.py:232) when
DIR/gt/rgb_<frame>.png is readable (resized in float32 and rounded where cv2
resizes uint8: parity unpinned when the sizes differ) and reports no SSIM
(skimage's windowed statistics, not restated here).
"""
import argparse
import json
import os
import warnings

import numpy as np
import torch
from PIL import Image

from .depth_estimation import RGBD_Estimator
from .evaluation import error_stats, psnr
from .io import decode_exr, png_bit_depth, save_rgbd_panorama
from .utils import parse_json_calib, prepare_inputs, resize_linear


def read_image(path):
    """cv2.imread(path, cv2.IMREAD_UNCHANGED) for the files PIL decodes:
    HW (greyscale) or HWC in BGR / BGRA order; None when unreadable.  A PNG
    of 16-bit colour, which PIL would reduce to its high bytes (cv2 keeps
    uint16, which read_input_images divides by 255), is refused with a
    warning instead (decode it with the caller's own reader and hand the
    uint16 array to prepare_inputs)."""
    try:
        with open(path, "rb") as f:
            head = f.read(64)
    except OSError:
        return None
    depth = png_bit_depth(head)
    if depth is not None and depth[0] == 16 and depth[1] in (2, 6):
        warnings.warn("16-bit colour PNG not supported by this reader: " + path)
        return None
    try:
        with Image.open(path) as im:
            if im.mode not in ("L", "I;16", "RGB", "RGBA"):
                return None
            a = np.array(im)
    except (OSError, ValueError):
        return None
    if a.ndim == 3:
        a = a[:, :, [2, 1, 0, 3][:a.shape[2]]]
    return np.ascontiguousarray(a)


def read_frame(filename, dataset_path, calibrations):
    """read_input_images (utils.py:157-204) up to the resize, which runs on
    the GPU: the C decoded images (uint8 / uint16 as read, float32 clipped
    to [0, 1] and scaled to [0, 255]) and whether the frame is valid."""
    images = []
    for k, calib in enumerate(calibrations):
        path = os.path.join(dataset_path, "cam%d" % k, filename)
        img = read_image(path)
        if img is None:
            warnings.warn("Cannot read image for file " + path)
            return [], False
        if img.shape != (calib.original_resolution[1], calib.original_resolution[0], 3):
            warnings.warn("Invalid image size / channels for file " + path)
            return [], False
        if img.dtype == np.float32:
            if img.max() > 1:
                img = np.clip(img, 0, 1)
                warnings.warn("Image has out-of-range float values for file " + path + ". Clipped for processing.")
            img = img * 255
        elif img.dtype not in (np.uint8, np.uint16):
            warnings.warn("Invalide image type for file " + path)
            return [], False
        images.append(img)
    return images, True


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--dataset_path", type=str, default="evaluation_dataset")
    p.add_argument("--references_indices", nargs="*", type=int, default=[0, 2])
    p.add_argument("--min_dist", type=float, default=0.55)
    p.add_argument("--max_dist", type=float, default=100)
    p.add_argument("--candidate_count", type=int, default=32)
    p.add_argument("--sigma_i", type=float, default=10)
    p.add_argument("--sigma_s", type=float, default=25)
    p.add_argument("--matching_resolution", nargs=2, type=int, default=[1024, 1024])
    p.add_argument("--rgb_to_stitch_resolution", nargs=2, type=int, default=[1216, 1216])
    p.add_argument("--panorama_resolution", nargs=2, type=int, default=[2048, 1024])
    p.add_argument("--device", type=str, default="cuda:0")
    p.add_argument("--saving", type=lambda s: s.lower() not in ("0", "false", "no"), default=True)
    p.add_argument("--visualize", type=lambda s: s.lower() not in ("0", "false", "no"), default=False)
    p.add_argument("--evaluate", type=lambda s: s.lower() not in ("0", "false", "no"), default=False)
    p.add_argument("--bad_px_ratio_thresholds", nargs="*", type=float, default=[0.1, 0.4])
    return p.parse_args(argv)


def run(args):
    """The loop of main.py:62-140; returns {filename: {"rgb", "inv_distance"}}
    (and, with --evaluate, {filename: metrics} as a second value)."""
    with open(os.path.join(args.dataset_path, "calibration.json")) as f:
        raw = json.load(f)["value0"]
    calibrations = parse_json_calib(raw, args.matching_resolution, args.device)
    viewpoint = torch.zeros([3], device=args.device)
    for r in args.references_indices:
        viewpoint += calibrations[r].rt[:3, 3] / len(args.references_indices)
    mc, mr = args.matching_resolution
    masks = []
    for k in range(len(calibrations)):
        mask = read_image(os.path.join(args.dataset_path, "cam%d" % k, "mask.png"))
        if mask is not None and mask.ndim == 2:
            m = resize_linear(torch.from_numpy(mask).to(args.device), mc, mr)
            masks.append(torch.round(m).clamp(0, 255).unsqueeze(0) / 255)
        else:
            masks.append(torch.ones([mr, mc], device=args.device).unsqueeze(0))
    est = RGBD_Estimator(calibrations, args.min_dist, args.max_dist, args.candidate_count, args.references_indices,
                         viewpoint, masks, args.matching_resolution, args.rgb_to_stitch_resolution,
                         args.panorama_resolution, args.sigma_i, args.sigma_s, args.device)
    filenames = sorted(f for f in os.listdir(os.path.join(args.dataset_path, "cam0")) if f != "mask.png")
    panoramas = {}
    for filename in filenames:
        images, valid = read_frame(filename, args.dataset_path, calibrations)
        if not valid:
            continue
        dev = [torch.from_numpy(i).to(args.device) for i in images]
        to_match, to_stitch = prepare_inputs(dev, args.matching_resolution, args.rgb_to_stitch_resolution,
                                             args.references_indices)
        rgb, distance = est.estimate_RGBD_panorama(to_match, to_stitch)
        panoramas[filename] = {"rgb": rgb.cpu().numpy(), "inv_distance": 1 / distance.cpu().numpy()}
    if args.saving:
        os.makedirs(os.path.join(args.dataset_path, "output"), exist_ok=True)
        for filename in filenames:
            save_rgbd_panorama(panoramas, filename, args.dataset_path)
    if not args.evaluate:
        return panoramas
    metrics = {}
    pc, pr = args.panorama_resolution
    for filename, pano in panoramas.items():
        gt_path = os.path.join(args.dataset_path, "gt", "inv_distance_" + os.path.splitext(filename)[0] + ".exr")
        gt = decode_exr(open(gt_path, "rb").read()) if os.path.isfile(gt_path) else None
        if gt is None:
            warnings.warn("Invalid ground truth for file " + filename + ". Will be ignored for evaluation")
            continue
        gt = torch.from_numpy(gt).to(args.device)
        if tuple(gt.shape) != (pr, pc):
            gt = resize_linear(gt, pc, pr)  # utils.py:226-227 resizes it to the panorama
        s = error_stats(torch.from_numpy(pano["inv_distance"]).to(args.device), gt, args.bad_px_ratio_thresholds,
                        finite_only=False)
        gt_rgb = read_image(os.path.join(args.dataset_path, "gt", "rgb_" + os.path.splitext(filename)[0] + ".png"))
        if gt_rgb is not None and gt_rgb.dtype == np.uint8 and gt_rgb.ndim == 3 and gt_rgb.shape[2] == 3:
            if gt_rgb.shape[:2] != (pr, pc):
                m = resize_linear(torch.from_numpy(gt_rgb).to(args.device), pc, pr)
                gt_rgb = torch.round(m).clamp(0, 255).to(torch.uint8).cpu().numpy()
            s["psnr"] = psnr(gt_rgb, pano["rgb"])
        metrics[filename] = s
    if panoramas and not metrics:
        raise RuntimeError("--evaluate: no frame has a readable ground truth under %s/gt (inv_distance_<frame>.exr, "
                           "single-channel float32, uncompressed or ZIP)" % args.dataset_path)
    return panoramas, metrics


def main(argv=None):
    args = parse_args(argv)
    out = run(args)
    if args.evaluate:
        for name, m in sorted(out[1].items()):
            print(name, m)


if __name__ == "__main__":
    main()
