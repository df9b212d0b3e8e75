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
