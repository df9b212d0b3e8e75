"""Inverse distance and panorama error statistics on the GPU.

``inverse_distance`` is main.py:114 (``1 / distance``); ``error_stats`` is the
metric part of utils.evaluate_rgbd_panorama (utils.py:233-241): mean absolute
error, root mean square error and bad pixel ratios (in percent) of an
evaluated inverse-distance map against a reference one.  The per-pixel error
is the fp32 ``|a - b|`` of the reference; the sums are fp64 in a fixed order.
``psnr`` is the panorama colour's PSNR of utils.py:229-232 (on the host).
"""
import math

import numpy as np
import torch

from . import _lib


def inverse_distance(distance, out=None):
    """1 / distance, elementwise, as a new (or the given) float32 tensor."""
    if distance.dtype != torch.float32 or not distance.is_contiguous():
        raise ValueError("distance must be a contiguous float32 tensor")
    out = torch.empty_like(distance) if out is None else out
    _lib.check(_lib.lib.ss_inverse_distance(distance.data_ptr(), out.data_ptr(), distance.numel(),
                                            _lib.stream_of(distance)))
    return out


def error_stats(evaluated, reference, thresholds=(0.1, 0.4), finite_only=True):
    """MAE, RMSE and bad pixel ratios (utils.py:233-241, default thresholds of
    main.py:60) of ``evaluated`` against ``reference`` (same shape, float32,
    on the GPU).  With ``finite_only`` pixels where either map is NaN or
    infinite are left out and counted in ``skipped``."""
    if evaluated.shape != reference.shape:
        raise ValueError("maps differ in shape")
    a = evaluated.contiguous().float()
    b = reference.to(a.device, torch.float32).contiguous()
    n = a.numel()
    ws = torch.empty(_lib.lib.ss_error_stats_workspace_bytes(n), dtype=torch.uint8, device=a.device)
    out = torch.zeros(7, dtype=torch.float64, device=a.device)
    thr = _lib.float_array(thresholds)
    _lib.check(_lib.lib.ss_error_stats(a.data_ptr(), b.data_ptr(), n, thr, len(thresholds), int(bool(finite_only)),
                                       out.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_of(a)))
    s = out.cpu().tolist()
    count = int(s[2])
    if count == 0:
        return {"mae": float("nan"), "rmse": float("nan"), "bad_px_ratios": [float("nan")] * len(thresholds),
                "count": 0, "skipped": n}
    return {"mae": s[0] / count, "rmse": math.sqrt(s[1] / count),
            "bad_px_ratios": [100.0 * s[3 + k] / count for k in range(len(thresholds))],
            "count": count, "skipped": n - count}


def psnr(gt_rgb, rgb):
    """utils.py:229-232's peak_signal_noise_ratio(gt / 255, rgb / 255) of two
    uint8 images of one shape, as skimage computes it for float32 images in
    [0, 1]: data range 1, the squared differences in float32, their mean in
    float64, 10 * log10(1 / mean)."""
    a = np.asarray(gt_rgb)
    b = np.asarray(rgb)
    if a.shape != b.shape or a.dtype != np.uint8 or b.dtype != np.uint8:
        raise ValueError("psnr expects two uint8 images of one shape")
    d = a.astype(np.float32) / np.float32(255) - b.astype(np.float32) / np.float32(255)
    mse = np.mean(d * d, dtype=np.float64)
    return float(10 * np.log10(1.0 / mse)) if mse > 0 else float("inf")
