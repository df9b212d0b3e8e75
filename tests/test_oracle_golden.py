"""The oracle against the reference's own outputs (tests/golden/reference_vectors.npz).

The vectors were written by tests/golden/make_golden.py, which runs the
reference's python/utils.py and python/depth_estimation.py on torch-CPU.  Every
comparison here is bit-exact.
"""
import json

import numpy as np
import pytest
import torch

from helpers import first_mismatch
from oracle import Oracle

Z = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "reference_vectors.npz"))
O = Oracle("f32")
H, W = Z["images"].shape[1:3]


def cams():
    out = []
    for k in range(4):
        fs = (torch.tensor(Z["fl"][k]) * torch.tensor(Z["matching_scale"][k])).tolist()
        cs = (torch.tensor(Z["principal"][k]) * torch.tensor(Z["matching_scale"][k])).tolist()
        out.append([*fs, *cs, *Z["xi_alpha"][k]])
    return np.array(out)


CAMS = cams()


def rts(r):
    rt = torch.tensor(Z["rt"])
    return np.concatenate([torch.matmul(torch.inverse(rt[k]), rt[r])[:3].reshape(-1).numpy()
                           for k in range(4)]).astype(np.float64)


def same(a, b, what):
    msg = first_mismatch(a, b)
    assert msg is None, "%s: %s" % (what, msg)


@pytest.mark.parametrize("k", [0, 1])
def test_unproject_grid(k):
    u, v = np.meshgrid(np.arange(W), np.arange(H))
    p, val = O.unproject_torch(CAMS[k], np.stack([u, v], -1).reshape(-1, 2))
    same(p.reshape(H, W, 3), Z["unproject_grid_%d" % k], "points")
    same(val.reshape(H, W), Z["unproject_grid_valid_%d" % k], "valid")


@pytest.mark.parametrize("k", [0, 1])
def test_unproject_random(k):
    p, val = O.unproject_torch(CAMS[k], Z["uv_rand"])
    same(p, Z["unproject_rand_%d" % k], "points")
    same(val, Z["unproject_rand_valid_%d" % k], "valid")


@pytest.mark.parametrize("k", [0, 1])
def test_project_random(k):
    uv, val = O.project_torch(CAMS[k], Z["pts_rand"])
    same(uv, Z["project_rand_%d" % k], "uv")
    same(val, Z["project_rand_valid_%d" % k], "valid")
    assert not val.all() and val.any()


def test_guide():
    for k in range(4):
        same(O.guide(Z["images"][k]), Z["guide_ref"][k], "guide %d" % k)


@pytest.mark.parametrize("slot,r", [(0, 0), (1, 2)])
def test_select_camera(slot, r):
    sel = O.select_camera(CAMS, rts(r), Z["masks"], r, 0.55, 100)
    same(sel, Z["selected"][slot], "selection")
    assert not np.any(sel == r)


@pytest.mark.parametrize("slot,r", [(0, 0), (1, 2)])
def test_sweep_cost(slot, r):
    cost = O.sweep(Z["images"], Z["selected"][slot], Z["candidates"].astype(np.float64), CAMS, rts(r), r)
    same(cost, Z["cost_raw"][slot], "cost volume")


def _A():
    c = Z["candidates"]
    return c, float(np.float32(c[0]) / np.float32(c[-1]) - np.float32(1))


@pytest.mark.parametrize("slot", [0, 1])
def test_argmin_on_sweep(slot):
    c, A = _A()
    d, _ = O.argmin_refine(Z["cost_raw"][slot], c[0], c[-1], A)
    same(d, Z["dist_unfiltered"][slot], "distance")


def test_argmin_designed_volumes():
    c, A = _A()
    d, _ = O.argmin_refine(Z["argmin_volume"], c[0], c[-1], A)
    same(d, Z["argmin_distance"], "distance")


def test_parse_json_calib_matches_reference():
    from sphere_stereo.utils import parse_json_calib

    for pre in ("", "demo_"):
        raw = json.loads(bytes(Z[pre + "calib_json"]).decode())["value0"]
        res = [W, H] if pre == "" else [1024, 1024]
        cal = parse_json_calib(raw, res, "cpu")
        same(np.stack([c.rt.numpy() for c in cal]), Z[pre + "rt"], pre + "rt")
        same(np.stack([c.fl.numpy() for c in cal]), Z[pre + "fl"], pre + "fl")
        same(np.stack([c.principal.numpy() for c in cal]), Z[pre + "principal"], pre + "principal")
        same(np.stack([c.matching_scale.numpy() for c in cal]), Z[pre + "matching_scale"], pre + "scale")
        same(np.array([[c.xi, c.alpha] for c in cal]), Z[pre + "xi_alpha"], pre + "xi/alpha")


def test_rgb2ycbcr_host_helper_matches_reference():
    from sphere_stereo.utils import rgb2yCbCr

    for k in range(4):
        same(rgb2yCbCr(torch.tensor(Z["images"][k])).type(torch.uint8).numpy(), Z["guide_ref"][k], "guide")
