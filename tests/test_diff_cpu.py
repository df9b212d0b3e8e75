"""Differentiation form, CPU side: the oracle's soft-argmin readout against
closed forms, and the float64 torch reference (tests/diff_reference.py) against
the fp64 oracle forward and against its own central differences."""
import numpy as np
import pytest
import torch

import diff_reference as dr
from helpers import frame
from oracle import Oracle
from sphere_stereo.differentiable import camera_parameters, camera_poses, pose_exp, relative_poses
from sphere_stereo.utils import distance_candidates

O32, O64 = Oracle("f32"), Oracle("f64")


@pytest.mark.parametrize("o", [O32, O64], ids=["f32", "f64"])
def test_soft_argmin_closed_forms(o):
    N, beta, inv0, inv1 = 12, 0.04, 0.01, 1 / 0.55
    # constant cost: uniform weights, iota = (N - 1) / 2
    got = o.soft_argmin(np.full((N, 3, 4), 7.0), beta, inv0, inv1)
    np.testing.assert_allclose(got, inv0 + (inv1 - inv0) * 0.5, rtol=1e-6)
    # one dominant minimum: every other weight is exp(-16) relative
    c = np.full((N, 1, 1), 400.0)
    c[5] = 0.0
    r = np.exp(-beta * 400.0)
    iota = (5 + r * (sum(range(N)) - 5)) / (1 + r * (N - 1))
    np.testing.assert_allclose(o.soft_argmin(c, beta, inv0, inv1), inv0 + (inv1 - inv0) * iota / (N - 1), rtol=1e-6)
    # linear cost a*n: geometric weights r^n
    a = 3.0
    c = (a * np.arange(N, dtype=np.float64))[:, None, None] * np.ones((1, 2, 2))
    r = np.exp(-beta * a)
    w = r ** np.arange(N)
    iota = (np.arange(N) * w).sum() / w.sum()
    np.testing.assert_allclose(o.soft_argmin(c, beta, inv0, inv1), inv0 + (inv1 - inv0) * iota / (N - 1), rtol=1e-5)
    # shifting every cost by a constant changes nothing (the minimum is taken out)
    c2 = c + 123.0
    np.testing.assert_allclose(o.soft_argmin(c2, beta, inv0, inv1), o.soft_argmin(c, beta, inv0, inv1), rtol=1e-6)


def test_pose_exp_is_a_rigid_transform():
    d = torch.tensor([0.01, -0.02, 0.005, 0.003, -0.001, 0.002], dtype=torch.float64)
    T = pose_exp(d)
    R = T[:3, :3]
    assert torch.allclose(R @ R.T, torch.eye(3, dtype=torch.float64), atol=1e-14)
    assert abs(float(torch.det(R)) - 1) < 1e-14
    # exp of a rotation about z by theta
    th = 0.3
    Tz = pose_exp(torch.tensor([0, 0, th, 0, 0, 0], dtype=torch.float64))
    ref = torch.tensor([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]], dtype=torch.float64)
    assert torch.allclose(Tz[:3, :3], ref, atol=1e-14)
    assert torch.equal(pose_exp(torch.zeros(6, dtype=torch.float64)), torch.eye(4, dtype=torch.float64))


def problem(seed=0, matching=(40, 36), N=8, ref=0):
    f = frame(seed=seed, resolution=160, matching=matching, refs=(0, 2))
    calibs = f["calibs"]
    # a slightly irregular rig: the synthetic one is exactly symmetric (level
    # cameras, principal points on pixel centres), which puts many samples on
    # the kinks of the bilinear cells, where one-sided derivatives differ
    gen = torch.Generator().manual_seed(100 + seed)
    cams6 = camera_parameters(calibs)
    cams6[:, 2:4] += 0.1 + 0.3 * torch.rand((len(calibs), 2), generator=gen, dtype=torch.float64)
    rts = torch.stack([rt @ pose_exp(0.01 * torch.randn(6, generator=gen, dtype=torch.float64))
                       for rt in camera_poses(calibs)])
    imgs32 = np.stack([t.numpy() for t in f["to_match"]]).astype(np.float32)
    masks = np.stack([m[0].numpy() for m in f["masks"]]).astype(np.float32)
    rel = relative_poses(rts, ref)
    sel = O32.select_camera(cams6.numpy(), rel.reshape(len(calibs), 12).numpy(), masks, ref, 0.55, 100)
    cands = np.array(distance_candidates(0.55, 100, N), np.float64)
    return dict(cams6=cams6, rts=rts, rel=rel, imgs32=imgs32, masks=masks, sel=sel, cands=cands, ref=ref, N=N,
                sigma_i=10.0, sigma_s=12.5, beta=1 / (500 * 0.05), inv0=1 / 0.55, inv1=1 / 100.0)


def oracle_forward(o, p, cams6=None, rel=None):
    cams6 = p["cams6"].numpy() if cams6 is None else cams6
    rel = p["rel"].reshape(-1, 12).numpy() if rel is None else rel
    cost = o.sweep(p["imgs32"], p["sel"], p["cands"], cams6, rel, p["ref"])
    g = O32.guide(p["imgs32"][p["ref"]])
    Fv, _ = o.isb_filter(g, cost, p["sigma_i"], p["sigma_s"])
    return o.soft_argmin(Fv, p["beta"], p["inv0"], p["inv1"])


def torch_forward(p, tab, cams6, rel, images):
    return dr.forward(images, cams6, rel, p["ref"], torch.from_numpy(p["sel"]).long(), torch.from_numpy(p["cands"]),
                      tab, p["beta"], p["inv0"], p["inv1"])


def test_torch_reference_forward_matches_oracle():
    p = problem()
    tab = dr.filter_tables(O32.guide(p["imgs32"][p["ref"]]), p["sigma_i"], p["sigma_s"], O32)
    images = [torch.from_numpy(im).double() for im in p["imgs32"]]
    got = torch_forward(p, tab, p["cams6"], p["rel"], images).numpy()
    want = oracle_forward(O64, p)
    # the same method in fp64; only the guide pyramid differs (fp32 bytes here,
    # fp64-rounded bytes in the oracle), which moves a few weights slightly
    err = np.abs(got - want)
    assert np.median(err) < 1e-9, np.median(err)
    assert err.max() < 2e-3 * (p["inv0"] - p["inv1"]), err.max()
    # and the cost stage alone is the oracle's sweep
    cost_t = dr.sweep_cost(images, p["cams6"], p["rel"], p["ref"], torch.from_numpy(p["sel"]).long(),
                           torch.from_numpy(p["cands"])).numpy()
    cost_o = O64.sweep(p["imgs32"], p["sel"], p["cands"], p["cams6"].numpy(), p["rel"].reshape(-1, 12).numpy(),
                       p["ref"])
    np.testing.assert_allclose(cost_t, cost_o, rtol=1e-9, atol=1e-9)


def test_torch_reference_gradient_matches_central_differences():
    """Autograd of the torch reference against float64 central differences
    (h = 1e-6) of the same function, for intrinsics of the sampled and
    reference cameras and entries of a relative transform.  (At this tiny
    resolution steps of 1e-4 in xi already move some samples across the kinks
    of the bilinear cells and of |.|; the full-size check uses h = 1e-4.)"""
    p = problem()
    tab = dr.filter_tables(O32.guide(p["imgs32"][p["ref"]]), p["sigma_i"], p["sigma_s"], O32)
    images = [torch.from_numpy(im).double() for im in p["imgs32"]]
    gen = torch.Generator().manual_seed(3)
    wmap = torch.rand(p["imgs32"].shape[1:3], generator=gen, dtype=torch.float64)
    cams6 = p["cams6"].clone().requires_grad_(True)
    rel = p["rel"].clone().requires_grad_(True)
    loss = (torch_forward(p, tab, cams6, rel, images) * wmap).sum()
    loss.backward()
    k = int(np.bincount(p["sel"][p["sel"] >= 0].ravel()).argmax())  # the most selected camera

    def L(c6, rl):
        with torch.no_grad():
            return float((torch_forward(p, tab, c6, rl, images) * wmap).sum())

    checks = [("cam", k, 0), ("cam", k, 2), ("cam", k, 4), ("cam", k, 5), ("cam", p["ref"], 0), ("cam", p["ref"], 4),
              ("rel", k, (0, 3)), ("rel", k, (1, 0)), ("rel", k, (2, 2))]
    for kind, kk, i in checks:
        if kind == "cam":
            base = p["cams6"]
            h = 1e-6
            cp, cm = base.clone(), base.clone()
            cp[kk, i] += h
            cm[kk, i] -= h
            fd = (L(cp, p["rel"]) - L(cm, p["rel"])) / (2 * h)
            ad = float(cams6.grad[kk, i])
        else:
            base = p["rel"]
            h = 1e-6
            rp, rm = base.clone(), base.clone()
            rp[kk][i] += h
            rm[kk][i] -= h
            fd = (L(p["cams6"], rp) - L(p["cams6"], rm)) / (2 * h)
            ad = float(rel.grad[kk][i])
        assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1.0), (kind, kk, i, ad, fd)


def test_resized_principal_point_follows_the_resize():
    """camera_parameters(pixel_centres="resized"): a point at original pixel
    coordinate c appears at (c + 0.5) * s - 0.5 after the INTER_LINEAR resize
    of read_input_images (utils.py:198-201), which the oracle's resize
    restates: a ramp f(x) = x resized from 1216 to 768 columns reads c there
    (exact for a linear ramp away from the clamped borders); the reference's
    c * s is 0.5 * (1 - s) px off."""
    from sphere_stereo import synth
    from sphere_stereo.utils import parse_json_calib

    rig = synth.SyntheticRig(seed=0, resolution=1216)
    calibs = parse_json_calib(rig.basalt_json()["value0"], [768, 768], "cpu")
    exact = camera_parameters(calibs, "resized")
    ref = camera_parameters(calibs)
    s = 768 / 1216
    assert torch.allclose(exact[:, 2:4], ref[:, 2:4] + 0.5 * s - 0.5, atol=1e-4)
    assert torch.equal(exact[:, :2], ref[:, :2]) and torch.equal(exact[:, 4:], ref[:, 4:])
    ramp = np.tile(np.arange(1216, dtype=np.float32), (3, 1))
    out = O32.resize_linear(ramp, 768, 3)[1].astype(np.float64)

    def at(x):  # the resized ramp read at resized coordinate x
        x0 = int(np.floor(x))
        return out[x0] + (x - x0) * (out[x0 + 1] - out[x0])

    for c in (100.0, 608.0, 1000.25):
        assert abs(at((c + 0.5) * s - 0.5) - c) < 1e-3, c
        assert abs(at(c * s) - c) > 0.25, c  # 0.5 * (1 / s - 1) = 0.29 original pixels off
