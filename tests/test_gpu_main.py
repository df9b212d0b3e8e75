"""main.py's dataset loop (sphere_stereo.main) end to end on a synthetic
dataset written to disk: calibration.json, cam<k>/ PNG frames and masks,
ground-truth inverse-distance EXRs (one uncompressed at the panorama size,
one ZIP-compressed at twice that size, which the loop resizes, as
utils.py:226-227 does) and a ground-truth colour PNG; one frame of the wrong
size is skipped."""
import json
import os

import numpy as np
import pytest
import torch
from PIL import Image

from sphere_stereo import main as ss_main
from sphere_stereo import synth
from sphere_stereo.evaluation import psnr
from sphere_stereo.io import decode_exr, encode_exr, encode_png

from helpers import encode_exr_zip

pytestmark = pytest.mark.gpu


def test_main_dataset_loop(gpu, tmp_path):
    rig = synth.SyntheticRig(seed=7, resolution=256, baseline=0.26)
    imgs, masks = rig.render(gpu)
    root = tmp_path / "dataset"
    for k in range(imgs.shape[0]):
        d = root / ("cam%d" % k)
        os.makedirs(d)
        (d / "f0.png").write_bytes(encode_png(imgs[k].cpu().numpy()))
        (d / "f1.png").write_bytes(encode_png(imgs[k].cpu().numpy()[:128]))  # wrong size: skipped
        (d / "f2.png").write_bytes(encode_png(imgs[k].cpu().numpy()))  # f0 again, other ground truth
        (d / "mask.png").write_bytes(encode_png(masks[k].cpu().numpy()))
    (root / "calibration.json").write_text(json.dumps(rig.basalt_json()))
    refs = [0, 2]
    calib = __import__("sphere_stereo").parse_json_calib(rig.basalt_json()["value0"], [128, 128], gpu)
    vp = sum(calib[r].rt[:3, 3] for r in refs) / len(refs)
    gt = rig.panorama_inverse_distance(vp, 128, 64, device=gpu).cpu().numpy().astype(np.float32)
    os.makedirs(root / "gt")
    (root / "gt" / "inv_distance_f0.exr").write_bytes(encode_exr(gt))
    # f2: the ground truth at twice the panorama size (each value repeated 2 x 2,
    # which the bilinear half-size resize returns exactly), ZIP-compressed as
    # cv2.imwrite writes EXR, and a colour ground truth likewise
    gt2 = np.repeat(np.repeat(gt, 2, axis=0), 2, axis=1)
    (root / "gt" / "inv_distance_f2.exr").write_bytes(encode_exr_zip(gt2, 16))
    gt_rgb = np.random.RandomState(9).randint(0, 256, (64, 128, 3)).astype(np.uint8)
    (root / "gt" / "rgb_f2.png").write_bytes(encode_png(np.repeat(np.repeat(gt_rgb, 2, axis=0), 2, axis=1)))
    args = ss_main.parse_args(["--dataset_path", str(root), "--matching_resolution", "128", "128",
                               "--rgb_to_stitch_resolution", "160", "160", "--panorama_resolution", "128", "64",
                               "--candidate_count", "16", "--sigma_s", "12.5", "--device", str(gpu), "--evaluate", "1"])
    with pytest.warns(UserWarning, match="Invalid image size"):
        panoramas, metrics = ss_main.run(args)
    assert sorted(panoramas) == ["f0.png", "f2.png"]
    out = root / "output"
    assert sorted(os.listdir(out)) == ["inv_distance_f0.exr", "inv_distance_f2.exr", "rgb_f0.png", "rgb_f2.png"]
    p = panoramas["f0.png"]
    assert p["rgb"].shape == (64, 128, 3) and p["inv_distance"].shape == (64, 128)
    np.testing.assert_array_equal(np.array(Image.open(out / "rgb_f0.png")), p["rgb"][:, :, ::-1])
    got = decode_exr((out / "inv_distance_f0.exr").read_bytes())
    np.testing.assert_array_equal(got.view(np.uint32), p["inv_distance"].view(np.uint32))
    # the GPU metrics against the same numbers on the host; as in the
    # reference (utils.py:233-242 sums every pixel), a NaN of the ground truth
    # propagates into the MAE
    err = np.abs(p["inv_distance"].astype(np.float64) - gt.astype(np.float64))
    m = metrics["f0.png"]
    assert m["count"] == err.size
    if np.isnan(err).any():
        assert np.isnan(m["mae"]) and np.isnan(m["rmse"])
    else:
        assert abs(m["mae"] - err.mean()) <= 1e-9 * max(1.0, err.mean())
    assert abs(m["bad_px_ratios"][0] - 100.0 * (err > 0.1).mean()) < 1e-9
    # f2: the same frame against the resized ZIP ground truth gives the same
    # figures, plus the colour PSNR of utils.py:232
    p2, m2 = panoramas["f2.png"], metrics["f2.png"]
    np.testing.assert_array_equal(p2["inv_distance"].view(np.uint32), p["inv_distance"].view(np.uint32))
    for key in ("mae", "rmse", "bad_px_ratios", "count"):
        np.testing.assert_array_equal(np.asarray(m2[key]), np.asarray(m[key]))
    assert "psnr" not in m and m2["psnr"] == psnr(gt_rgb, p2["rgb"])
    # the inverse distances: most finite pixels in the candidate range (pixels
    # no reference blends into stay NaN at this tiny panorama size, and holes
    # beyond the inpainting's reach blend in 1e8, as in the reference)
    inv = p["inv_distance"]
    fin = inv[np.isfinite(inv)]
    assert fin.size > 1000 and 1 / 100.5 <= np.median(fin) <= 1 / 0.5 and fin.max() <= 1 / 0.5
    # without readable ground truths, --evaluate fails loudly instead of
    # returning no metrics
    for name in os.listdir(root / "gt"):
        os.remove(root / "gt" / name)
    args.saving = False
    with pytest.warns(UserWarning, match="Invalid ground truth"):
        with pytest.raises(RuntimeError, match="no frame has a readable ground truth"):
            ss_main.run(args)
