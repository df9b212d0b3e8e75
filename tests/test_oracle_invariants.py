"""Closed forms, invariants and independent restatements that pin the oracle
where the reference's CUDA kernels (isb_filter.cu, stitcher.cu) cannot run.

Each check is chosen so that a plausible slip in the oracle (a wrong
neighbour or base weight, a transposed index, a missing clamp, a dropped term)
fails it.
"""
import math

import numpy as np
import pytest
import scipy.ndimage

from oracle import Oracle

O32, O64 = Oracle("f32"), Oracle("f64")


# ---------------------------------------------------------------------------
# exp (Cody-Waite + Cephes polynomial)
# ---------------------------------------------------------------------------
def _ulp32(x):
    return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)


def test_exp_normal_range_within_2ulp():
    x = np.concatenate([np.linspace(-87.3, 88.7, 20001), -np.logspace(-8, 1.9, 3001)]).astype(np.float32)
    got = O32.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    err = np.abs(got - ref) / _ulp32(ref)
    assert err.max() <= 2.0, err.max()


def test_exp_subnormal_and_limits():
    x = np.linspace(-103.9, -87.4, 2001).astype(np.float32)
    got = O32.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    assert np.all(got > 0)
    assert np.abs(got - ref).max() <= 2 * 2.0 ** -149
    assert O32.expf([-105.0])[0] == 0.0
    assert np.isinf(O32.expf([89.0])[0])
    assert O32.expf([0.0])[0] == 1.0


# ---------------------------------------------------------------------------
# double-sphere model
# ---------------------------------------------------------------------------
PINHOLE = [300.0, 290.0, 320.5, 240.25, 0.0, 0.0]  # fs, cs, xi = alpha = 0
DS = [364.8 * 1024 / 1216, 364.8 * 1024 / 1216, 512.0, 512.0, -0.18, 0.59]


@pytest.mark.parametrize("o,tol", [(O64, 1e-9), (O32, 2e-3)])
def test_pinhole_reduction(o, tol):
    rng = np.random.RandomState(0)
    pts = rng.randn(500, 3)
    pts[:, 2] = np.abs(pts[:, 2]) + 0.5
    want = np.stack([300.0 * pts[:, 0] / pts[:, 2] + 320.5, 290.0 * pts[:, 1] / pts[:, 2] + 240.25], -1)
    uv, val = o.project_torch(PINHOLE, pts)
    assert np.abs(uv - want).max() < tol * 10 and val.all()
    uvc = o.project_cuda(PINHOLE, pts)
    assert np.abs(uvc - want).max() < tol * 10
    uvs = rng.rand(500, 2) * [640, 480]
    p, _ = o.unproject_torch(PINHOLE, uvs)
    ray = np.stack([(uvs[:, 0] - 320.5) / 300.0, (uvs[:, 1] - 240.25) / 290.0, np.ones(500)], -1)
    ray /= np.linalg.norm(ray, axis=-1, keepdims=True)
    assert np.abs(p - ray).max() < tol
    assert np.abs(o.unproject_cuda(PINHOLE, uvs) - ray).max() < tol


@pytest.mark.parametrize("o,tol", [(O64, 1e-8), (O32, 5e-3)])
def test_round_trip_and_homogeneity(o, tol):
    rng = np.random.RandomState(1)
    uv = 512 + (rng.rand(2000, 2) - 0.5) * 1000
    p, val = o.unproject_torch(DS, uv)
    back, _ = o.project_torch(DS, p[val])
    assert np.abs(back - uv[val]).max() < tol
    back2, _ = o.project_torch(DS, p[val] * 37.5)
    assert np.abs(back2 - uv[val]).max() < tol
    pc = o.unproject_cuda(DS, uv[val])
    assert np.abs(o.project_cuda(DS, pc * 3.0) - uv[val]).max() < tol
    assert np.abs(np.linalg.norm(p, axis=-1) - 1).max() < 1e-6


def test_projection_validity_boundary():
    # the DS model is valid for z > -w2 * |p| (utils.py:86-92)
    al, xi = DS[5], DS[4]
    w1 = (1 - al) / al
    w2 = (w1 + xi) / math.sqrt(2 * w1 * xi + xi * xi + 1)
    ang = math.acos(-w2)
    th = np.array([ang - 1e-3, ang + 1e-3])
    pts = np.stack([np.sin(th), np.zeros(2), np.cos(th)], -1)
    _, val = O64.project_torch(DS, pts)
    assert val.tolist() == [True, False]
    _, val = O64.project_cuda(DS, pts, with_valid=True)
    assert val.tolist() == [True, False]


# ---------------------------------------------------------------------------
# ISB filter
# ---------------------------------------------------------------------------
def _levels(W, H):
    return int(min(math.log2(W), math.log2(H)) - 1)


def _clampi(i, n):
    return np.clip(i, 0, n - 1)


def stencil_pyramid_np(cost, sigma_s):
    """The filter for a guide whose every pixel is 0 (all range weights equal
    exp(0) + 0.01), written independently as gathers over index grids."""
    N, H, W = cost.shape
    L = _levels(W, H)
    e = 1.01
    vs = 1 / (2 * sigma_s * sigma_s)
    lev = [cost.astype(np.float64)]
    for s in range(1, L):
        src = lev[-1]
        h, w = src.shape[1:]
        ho, wo = math.ceil(H / 2 ** s), math.ceil(W / 2 ** s)
        y, x = np.meshgrid(np.arange(ho), np.arange(wo), indexing="ij")
        acc, ws = np.zeros((N, ho, wo)), 0.0
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                b = 1.0 if (dx, dy) == (0, 0) else (0.5 * e if dx == 0 or dy == 0 else 0.25 * e)
                acc += b * src[:, _clampi(2 * y + dy, h), _clampi(2 * x + dx, w)]
                ws += b
        lev.append(acc / ws)
    for s in range(L - 2, -1, -1):
        fine, coarse = lev[s].copy(), lev[s + 1]
        hc, wc = coarse.shape[1:]
        hf, wf = fine.shape[1:]
        wd = math.exp(-((2 ** s - 0.5) ** 2) * vs)
        wu = 1 - wd
        y, x = np.meshgrid(np.arange(hc), np.arange(wc), indexing="ij")

        def c(dy, dx):
            return coarse[:, _clampi(y + dy, hc), _clampi(x + dx, wc)]

        taps = {
            (0, 0): [((dy, dx), 8.0 if (dy, dx) == (0, 0) else 1.0) for dy in (-1, 0, 1) for dx in (-1, 0, 1)],
            (1, 0): [((0, 0), 8.0), ((1, 0), 8.0), ((0, -1), 1.0), ((1, -1), 1.0), ((0, 1), 1.0), ((1, 1), 1.0)],
            (0, 1): [((0, 0), 8.0), ((0, 1), 8.0), ((-1, 0), 1.0), ((-1, 1), 1.0), ((1, 0), 1.0), ((1, 1), 1.0)],
            (1, 1): [((0, 0), 1.0), ((0, 1), 1.0), ((1, 0), 1.0), ((1, 1), 1.0)],
        }
        out = fine.copy()
        for (oy, ox), tl in taps.items():
            fy, fx = 2 * y + oy, 2 * x + ox
            ok = (fy < hf) & (fx < wf)
            acc = wu * fine[:, np.minimum(fy, hf - 1), np.minimum(fx, wf - 1)]
            ws = wu
            for (dy, dx), b in tl:
                acc = acc + b * wd * e * c(dy, dx)
                ws += b * wd * e
            out[:, fy[ok], fx[ok]] = (acc / ws)[:, ok]
        lev[s] = out
    return lev[0]


@pytest.mark.parametrize("H,W", [(8, 8), (9, 13), (16, 33), (40, 24), (5, 64)])
def test_isb_constant_guide_is_fixed_stencil(H, W):
    rng = np.random.RandomState(H * W)
    cost = rng.rand(3, H, W) * 100
    guide = np.zeros((H, W, 3), np.uint8)
    got, g = O64.isb_filter(guide, cost, 10.0, 25.0)
    want = stencil_pyramid_np(cost, 25.0)
    assert np.abs(got - want).max() < 1e-9 * 100
    assert not g.any()
    got32, _ = O32.isb_filter(guide, cost.astype(np.float32), 10.0, 25.0)
    assert np.abs(got32 - want).max() < 1e-3


def test_isb_permutation_is_exact_and_linearity():
    rng = np.random.RandomState(5)
    H, W, N = 23, 31, 6
    guide = rng.randint(0, 255, (H, W, 3)).astype(np.uint8)
    guide[:, : W // 2] = 40
    c1, c2 = (rng.rand(N, H, W) * 100).astype(np.float32), (rng.rand(N, H, W) * 100).astype(np.float32)
    f1, g1 = O32.isb_filter(guide, c1, 10.0, 25.0)
    perm = rng.permutation(N)
    fp, gp = O32.isb_filter(guide, c1[perm], 10.0, 25.0)
    assert np.array_equal(fp, f1[perm]) and np.array_equal(gp, g1)
    d1, _ = O64.isb_filter(guide, c1, 10.0, 25.0)
    d2, _ = O64.isb_filter(guide, c2, 10.0, 25.0)
    d12, _ = O64.isb_filter(guide, 2.0 * c1.astype(np.float64) - 0.5 * c2.astype(np.float64), 10.0, 25.0)
    assert np.abs(d12 - (2 * d1 - 0.5 * d2)).max() < 1e-10
    const, _ = O32.isb_filter(guide, np.full((N, H, W), 123.25, np.float32), 10.0, 25.0)
    assert np.abs(const - 123.25).max() < 123.25 * 1e-5


def test_isb_edge_preservation():
    # two flat guide regions: a cost step aligned with the guide edge survives,
    # the same step across a flat guide is smoothed away
    H, W = 32, 32
    guide = np.zeros((H, W, 3), np.uint8)
    guide[:, 16:] = 200
    cost = np.zeros((1, H, W))
    cost[0, :, 16:] = 100.0
    edge, _ = O64.isb_filter(guide, cost, 10.0, 25.0)
    flat, _ = O64.isb_filter(np.zeros_like(guide), cost, 10.0, 25.0)
    assert edge[0, :, 15].max() < 5.0 and edge[0, :, 16].min() > 95.0
    assert flat[0, :, 15].min() > 20.0


def test_isb_constant_guide_bytes_truncate_by_at_most_one():
    g = np.full((16, 16, 3), 137, np.uint8)
    _, out = O32.isb_filter(g, np.zeros((1, 16, 16), np.float32), 10.0, 25.0)
    assert out.min() >= 137 - 2 and out.max() <= 137


def test_volume_rounding_matches_numpy_float16_and_torch_bfloat16():
    import torch

    rng = np.random.RandomState(11)
    x = np.concatenate([
        (rng.rand(20000) - 0.5) * 1200,  # the cost range
        rng.rand(5000) * 1e-4,  # binary16 subnormals (below 6.1e-5)
        np.float32(2.0) ** rng.randint(-30, 20, 2000) * (1 + rng.randint(0, 2048, 2000) / 2048.0 + 1 / 4096.0),  # ties
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e6, -1e6, np.inf, -np.inf, 5.9604645e-08, 2.9802322e-08, 3e-8],
    ]).astype(np.float32)
    f16 = O32.round_volume(x, 1)
    with np.errstate(over="ignore"):
        assert np.array_equal(f16, x.astype(np.float16).astype(np.float32))
    bf = O32.round_volume(x, 2)
    assert np.array_equal(bf, torch.from_numpy(x).to(torch.bfloat16).float().numpy())
    assert np.isnan(O32.round_volume(np.array([np.nan], np.float32), 1)).all()
    assert np.isnan(O32.round_volume(np.array([np.nan], np.float32), 2)).all()
    assert np.array_equal(O32.round_volume(x, 0), x)


def test_precision_form_rounds_exactly_the_stored_values():
    rng = np.random.RandomState(3)
    H, W, N = 21, 30, 5
    guide = rng.randint(0, 255, (H, W, 3)).astype(np.uint8)
    cost = (rng.rand(N, H, W) * 500).astype(np.float32)
    for vt, dt in ((1, np.float16), (2, None)):
        c16 = O32.round_volume(cost, vt)
        stored, _ = O32.isb_filter(guide, c16, 10.0, 25.0, vtype=vt, round_level0=True)
        fused, _ = O32.isb_filter(guide, c16, 10.0, 25.0, vtype=vt, round_level0=False)
        # level 0's last pass is the only difference between the two forms
        assert np.array_equal(O32.round_volume(fused, vt), stored)
        if dt is not None:
            assert np.array_equal(stored.astype(dt).astype(np.float32), stored)
        # the same pyramid composed from the single passes: only the level-0
        # store rounds; the coarser levels stay fp32
        L = O32.levels(W, H)
        gs, cs = [guide], [c16]
        for s in range(1, L):
            h, w = -(-H // 2 ** s), -(-W // 2 ** s)
            g, c = O32.isb_down(gs[-1], cs[-1], h, w, 1 / (2 * 10.0 ** 2))
            gs.append(g)
            cs.append(c)
        for s in range(L - 2, -1, -1):
            wd = math.exp(-((2.0 ** s - 0.5) ** 2) / (2 * 25.0 ** 2))
            gs[s], cs[s] = O32.isb_up(gs[s + 1], cs[s + 1], gs[s], cs[s], 1 - wd, wd, 1 / (2 * 10.0 ** 2),
                                      vtype=vt if s == 0 else 0)
        assert np.array_equal(cs[0], stored)
        # ... and with unrounded input the fp32 filter itself
        f32, _ = O32.isb_filter(guide, c16, 10.0, 25.0)
        assert np.array_equal(O32.round_volume(f32, vt), stored)
        # error against the unrounded fp64 method stays at the storage precision
        ref, _ = O64.isb_filter(guide, cost.astype(np.float64), 10.0, 25.0)
        tol = 500 * (2.0 ** -10 if vt == 1 else 2.0 ** -7)
        assert np.abs(stored - ref).max() < 2 * tol
    # a constant volume exactly representable in fp16 stays exactly constant in
    # every level: the stencils are normalised and each pass rounds back to it
    const, _ = O32.isb_filter(guide, np.full((N, H, W), 123.25, np.float32), 10.0, 25.0, vtype=1)
    assert np.all(const == 123.25)


def test_levels_formula():
    for W, H in [(4, 4), (5, 7), (8, 8), (512, 512), (768, 768), (1024, 1024), (1216, 1024), (2048, 1024)]:
        assert O32.levels(W, H) == _levels(W, H)
    with pytest.raises(ValueError):
        O32.isb_filter(np.zeros((3, 3, 3), np.uint8), np.zeros((1, 3, 3), np.float32), 10, 25)


# ---------------------------------------------------------------------------
# argmin + refinement
# ---------------------------------------------------------------------------
def _cands(N, lo=0.55, hi=100.0):
    c = 1 / np.linspace(1 / lo, 1 / hi, N)
    return c, c[0] / c[-1] - 1


def test_argmin_exact_parabola_recovers_subindex():
    N = 32
    c, A = _cands(N)
    i0 = np.linspace(0.6, N - 1.6, 200)
    n = np.arange(N)[:, None]
    vol = 3.0 * (n - i0[None]) ** 2 + 2.0
    d, idx = O64.argmin_refine(vol[:, None, :], c[0], c[-1], A)
    want = c[0] / (A * i0 / (N - 1) + 1)
    assert np.abs(d[0] - want).max() < 1e-9 * want.max()
    assert np.array_equal(idx[0], np.rint(i0).astype(int))


def test_argmin_ties_constant_and_endpoints():
    N = 8
    c, A = _cands(N)
    vol = np.full((N, 1, 4), 10.0)
    vol[3, 0, 0] = vol[5, 0, 0] = 1.0  # tie: first index wins
    vol[0, 0, 1] = 0.5  # minimum at 0: no refinement
    vol[N - 1, 0, 2] = 0.5  # minimum at N-1: no refinement
    d, idx = O64.argmin_refine(vol, c[0], c[-1], A)
    assert idx[0].tolist()[:3] == [3, 0, N - 1]
    assert d[0, 1] == c[0]
    assert abs(d[0, 2] - c[-1]) < 1e-12 * c[-1]
    assert d[0, 3] == c[-1]  # constant volume (max == min) -> farthest candidate


# ---------------------------------------------------------------------------
# stitcher
# ---------------------------------------------------------------------------
INTR = [0.3 * 256, 0.3 * 256, 128.0, 128.0, -0.18, 0.59]


def test_reproject_without_translation_is_identity():
    H = W = 256
    rng = np.random.RandomState(3)
    d = (1 + rng.rand(H, W) * 20).astype(np.float32)
    out = O32.reproject_zmin(d, INTR, [0, 0, 0])
    inner = np.s_[20:-20, 20:-20]
    close = np.abs(out[inner] - d[inner]) <= 2e-6 * d[inner]
    assert close.mean() > 0.999


def test_reproject_keeps_the_nearest():
    # a translation along the optical axis maps several source pixels to one
    # target; every target holds the minimum of the distances sent to it
    H = W = 64
    intr = [0.3 * 64, 0.3 * 64, 32.0, 32.0, -0.18, 0.59]
    d = np.linspace(2, 30, H * W).reshape(H, W)
    t = [0.05, -0.02, 0.3]
    out = O64.reproject_zmin(d, intr, t)
    u, v = np.meshgrid(np.arange(W), np.arange(H))
    p = O64.unproject_cuda(intr, np.stack([u, v], -1).reshape(-1, 2)) * d.reshape(-1, 1) - np.array(t)
    uv = O64.project_cuda(intr, p)
    dist = np.linalg.norm(p, axis=-1)
    idx = np.rint(uv[:, 1]).astype(np.int64) * W + np.rint(uv[:, 0]).astype(np.int64)
    want = np.full(H * W, 1e8)
    ok = (idx >= 0) & (idx < H * W)
    np.minimum.at(want, idx[ok], dist[ok])
    assert np.array_equal(out.reshape(-1), want)


def _thr(max_dist=100):
    return float(max_dist) + 0.1, float(np.float32(max_dist) + np.float32(0.1))


def test_inpaint_fills_isolated_hole_with_lut_mean_and_keeps_full_maps():
    H = W = 64
    lut = O32.inpaint_lut(INTR[:2] + [32.0, 32.0] + INTR[4:], [0.1, 0.0, 0.0], H, W, 0.55, 100)
    rng = np.random.RandomState(4)
    d = (1 + rng.rand(H, W) * 10).astype(np.float32)
    te, ts = _thr()
    assert np.array_equal(O32.inpaint_step(d, lut, te, ts), d)
    d2 = d.copy()
    y, x = 30, 33
    d2[y, x] = 1e8
    out = O32.inpaint_step(d2, lut, te, ts)
    num = den = 0.0
    for s in range(2):
        b = int(lut[y, x, s])
        w = float(b & 240)
        val = d2[y + ((b & 12) >> 2) - 1, x + (b & 3) - 1]
        num, den = num + w * val, den + w
    assert den > 0 and abs(out[y, x] - num / den) <= 1e-6 * out[y, x]
    d3 = d.copy()
    d3[0, 5] = 1e8
    assert O32.inpaint_step(d3, lut, te, ts)[0, 5] == np.float32(1e8)  # border never filled


def test_inpaint_lut_picks_the_two_best_aligned_neighbours():
    H = W = 48
    intr = [0.3 * 48, 0.3 * 48, 24.0, 24.0, -0.18, 0.59]
    t = [0.12, -0.03, 0.05]
    lut = O64.inpaint_lut(intr, t, H, W, 0.55, 100)
    u, v = np.meshgrid(np.arange(W), np.arange(H))
    unit = O64.unproject_cuda(intr, np.stack([u, v], -1).reshape(-1, 2))
    near = O64.project_cuda(intr, unit * 0.55 - np.array(t))
    far = O64.project_cuda(intr, unit * 100 - np.array(t))
    dv = far - near
    dv /= np.linalg.norm(dv, axis=-1, keepdims=True)
    dirs = np.array([(n, m) for n in (-1, 0, 1) for m in (-1, 0, 1) if n or m], float)
    cosv = dv @ (dirs / np.linalg.norm(dirs, axis=-1, keepdims=True)).T
    order = np.argsort(-cosv, axis=-1)
    for i in range(H * W):
        top = cosv[i, order[i]]
        if top[1] - top[2] < 1e-9 or top[0] - top[1] < 1e-9 or top[1] <= 0:
            continue
        b = lut.reshape(-1, 2)[i]
        got = {((int(bb) & 3) - 1, ((int(bb) & 12) >> 2) - 1) for bb in b}
        want = {tuple(dirs[order[i, 0]].astype(int)), tuple(dirs[order[i, 1]].astype(int))}
        assert got == want, i
        assert int(b[1]) & 240 == (int(top[0] * 255) & 240)


@pytest.mark.parametrize("pad", [0.0, 1.0])
def test_box_filter_matches_uniform_filter(pad):
    rng = np.random.RandomState(6)
    x = rng.rand(2, 37, 53)
    got = O64.box_filter(x, 7, pad)
    want = np.stack([scipy.ndimage.uniform_filter(x[b], size=15, mode="constant", cval=pad) for b in range(2)])
    assert np.abs(got - want).max() < 1e-12
    got32 = O32.box_filter(x.astype(np.float32), 7, pad)
    assert np.abs(got32 - want).max() < 1e-5


def _rig2():
    intr = np.array([[0.3 * 96, 0.3 * 96, 48.0, 48.0, -0.18, 0.59]] * 2)
    c, s = math.cos(math.pi / 4), math.sin(math.pi / 4)
    r0 = np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])  # inverse of a +45 deg yaw
    rot = np.stack([r0.reshape(-1), r0.T.reshape(-1)])
    trans = np.array([[0.0, 0.0, -0.05], [0.0, 0.0, -0.05]])
    return intr, rot, trans


def test_blend_lut_weights_sum_to_one_and_sampling_projects():
    intr, rot, trans = _rig2()
    masks = np.ones((2, 96, 96))
    PH, PW = 24, 48
    samp, raw = O64.blend_lut(intr, rot, trans, masks, PH, PW, 0.55, 100)
    assert np.abs(raw.sum(0) - 1).max() < 1e-12
    w = O64.normalize_weights(O64.box_filter(raw, 3, 0.0))
    assert np.abs(w.sum(0) - 1).max() < 1e-12
    y, x = np.meshgrid(np.arange(PH), np.arange(PW), indexing="ij")
    phi = (y + 0.5) * 3.14159265 / PH - 3.14159265 / 2
    th = (x + 0.5) * 2 * 3.14159265 / PW + 3.14159265
    unit = np.stack([np.cos(phi) * np.sin(th), np.sin(phi), np.cos(phi) * np.cos(th)], -1).reshape(-1, 3)
    for r in range(2):
        uv = O64.project_cuda(intr[r], unit @ rot[r].reshape(3, 3).T)
        uv = np.stack([np.clip(uv[:, 0], 0.1, 96 - 1.1), np.clip(uv[:, 1], 0.1, 96 - 1.1)], -1)
        assert np.abs(samp[r].reshape(-1, 2) - uv).max() < 1e-9


def test_merge_of_constant_maps_and_images():
    intr, rot, trans = _rig2()
    PH, PW = 24, 48
    samp, raw = O32.blend_lut(intr, rot, trans, np.ones((2, 96, 96), np.float32), PH, PW, 0.55, 100)
    w = O32.normalize_weights(O32.box_filter(raw, 3, 0.0))
    maps = np.full((2, 96, 96), 5.0, np.float32)
    imgs = np.full((2, 120, 120, 3), 100, np.uint8)
    dist, rgb = O32.merge(samp, w, maps, maps, imgs, trans, intr)
    assert np.abs(dist - 5.0).max() < 5e-6
    assert rgb.min() >= 99 and rgb.max() <= 100


def test_resize_linear_closed_forms():
    """The input stage's cv2 INTER_LINEAR resize (utils.py:198-201): same size
    is the identity; 2:1 takes means of 2 x 2 blocks (weights 1/2 exactly);
    a constant stays constant; upscaled borders repeat the edge samples; a
    ramp along x is reproduced at the mapped source coordinates."""
    rng = np.random.RandomState(0)
    x = (rng.rand(17, 23, 3) * 255).astype(np.float32)
    assert np.array_equal(O32.resize_linear(x, 23, 17), x)
    y = (rng.rand(12, 16, 3) * 255).astype(np.float32)
    h = y[:, 0::2] * np.float32(0.5) + y[:, 1::2] * np.float32(0.5)
    want = h[0::2] * np.float32(0.5) + h[1::2] * np.float32(0.5)
    assert np.array_equal(O32.resize_linear(y, 8, 6), want)
    c = np.full((31, 29), 87.0, np.float32)
    for cols, rows in ((12, 17), (50, 40), (29, 31)):
        out = O32.resize_linear(c, cols, rows)
        assert np.abs(out - 87.0).max() <= 87.0 * 2 ** -22
    up = O32.resize_linear(y, 64, 48)
    assert np.array_equal(up[:, 0], up[:, 1]) and np.array_equal(up[0], up[1])  # (d+0.5)/4-0.5 < 0 for d <= 1
    assert np.array_equal(up[:, -1], up[:, -2])
    ramp = np.tile(np.arange(40, dtype=np.float32) * 3.0, (5, 1))
    out = O64.resize_linear(ramp, 25, 5)
    src_x = (np.arange(25) + 0.5) * (40 / 25) - 0.5
    assert np.abs(out[2] - 3.0 * np.clip(src_x, 0, 39)).max() < 1e-6


def test_resize_linear_matches_torch_bilinear():
    """torch's bilinear interpolation (align_corners=False) maps coordinates
    the same way; it forms the scale and source coordinate in float32 instead
    of cv2's double, which moves a coordinate by up to ~1e-5 pixel, i.e. the
    values by up to ~1e-5 x the local gradient (<= 255 here).  A wrong mapping
    (pixel corners instead of centres, no clamp) is off by tens."""
    import torch
    import torch.nn.functional as F

    rng = np.random.RandomState(1)
    for (h, w), (rows, cols) in (((96, 96), (80, 80)), ((61, 47), (23, 37)), ((20, 30), (53, 71))):
        x = (rng.rand(h, w, 3) * 255).astype(np.float32)
        t = F.interpolate(torch.from_numpy(x).permute(2, 0, 1)[None], size=(rows, cols), mode="bilinear",
                          align_corners=False, antialias=False)[0].permute(1, 2, 0).numpy()
        d = np.abs(O32.resize_linear(x, cols, rows) - t)
        assert d.max() < 1e-2 and d.mean() < 1e-3


def test_oracle_results_do_not_depend_on_the_thread_count():
    """The oracle's per-pixel loops run on OpenMP threads; every output is
    computed by one thread, so 1 and 4 threads give the same bits."""
    import ctypes

    gomp = ctypes.CDLL("libgomp.so.1")
    before = gomp.omp_get_max_threads()
    o = Oracle("f32")
    rng = np.random.default_rng(9)
    cost = (rng.random((16, 37, 29)) * 500).astype(np.float32)
    guide = rng.integers(0, 256, (37, 29, 3), dtype=np.uint8)
    try:
        outs = []
        for n in (1, 4):
            gomp.omp_set_num_threads(n)
            f, g = o.isb_filter(guide, cost, 10.0, 12.5)
            outs.append((f, g, o.soft_argmin(f, 0.04, 1 / 0.55, 0.01)))
    finally:
        gomp.omp_set_num_threads(before)
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


# ---------------------------------------------------------------------------
# stitcher: blend weights and merge against hand-derived values
# ---------------------------------------------------------------------------
PI_REF = 3.14159265  # the reference's PI literal (stitcher.cu:325)


def _ds_project(intr, p):
    """Double-sphere projection from the model's definition (Usenko et al.,
    arXiv:1807.08957, eqs. 40-45), float64: d1 = |p|, c = xi d1 + z,
    d2 = |(x, y, c)|, uv = f xy / (alpha d2 + (1 - alpha) c) + c_xy; and the
    validity z > -w2 d1 with w1 = alpha / (1 - alpha) for alpha <= 0.5, else
    (1 - alpha) / alpha, w2 = (w1 + xi) / sqrt(2 w1 xi + xi^2 + 1)."""
    fx, fy, cx, cy, xi, al = [float(v) for v in intr]
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    d1 = np.sqrt(x * x + y * y + z * z)
    c = xi * d1 + z
    d2 = np.sqrt(x * x + y * y + c * c)
    den = al * d2 + (1 - al) * c
    w1 = al / (1 - al) if al <= 0.5 else (1 - al) / al
    w2 = (w1 + xi) / math.sqrt(2 * w1 * xi + xi * xi + 1)
    return np.stack([fx * x / den + cx, fy * y / den + cy], -1), z > -w2 * d1


def _bilinear(img, u, v):
    """Bilinear read of img (rows x cols [x ch]) at (u, v) from the four
    surrounding samples."""
    u0, v0 = np.floor(u).astype(int), np.floor(v).astype(int)
    fu, fv = u - u0, v - v0
    if img.ndim == 3:
        fu, fv = fu[..., None], fv[..., None]
    return (img[v0, u0] * (1 - fu) * (1 - fv) + img[v0, u0 + 1] * fu * (1 - fv) + img[v0 + 1, u0] * (1 - fu) * fv
            + img[v0 + 1, u0 + 1] * fu * fv)


def _stitch_rig():
    """Two references of a 64 x 64 matching resolution, yawed +-50 degrees,
    with different translations from the panorama centre."""
    W = H = 64
    intr = np.array([[0.3 * W, 0.31 * W, 32.5, 31.0, -0.18, 0.59], [0.29 * W, 0.3 * W, 31.0, 32.5, -0.2, 0.6]])
    rot = []
    for yaw in (math.radians(50), math.radians(-50)):
        c, s = math.cos(yaw), math.sin(yaw)
        rot.append(np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]]))
    trans = np.array([[0.12, -0.02, -0.05], [-0.09, 0.03, 0.07]])
    return H, W, intr, np.stack(rot), trans


def test_blend_lut_weight_closed_form():
    """createBlendingLutKernel (stitcher.cu:345-398): reference r's raw weight
    at a panorama pixel is exp(-|p_far - p_near|^2 / (1e-4 * H * W)), H x W
    the matching resolution (not the panorama's), where p_near / p_far are
    the pixels of camera r that see the panorama ray at min_dist / max_dist
    from the panorama centre (ray - translation, each coordinate clamped to
    [0.1, size - 1.1], the row offset by r * H into the stacked masks); it
    is 1e-8 where a projection is invalid or the stacked mask read at either
    pixel is <= 0.99; the raw weights are divided by their sum over the
    references.  Every expected value here is derived in float64 from the
    projection's definition, so a wrong normaliser, a flipped translation or
    a dropped mask test fails."""
    H, W, intr, rot, trans = _stitch_rig()
    trans = trans * 0.1  # near-far displacements of a pixel or less: weights far from 0 and 1
    PH, PW = 96, 192  # PH * PW = 4.5 * H * W
    masks = np.ones((2, H, W))
    masks[1, :, 40:] = 0.0  # part of camera 1 unreliable: weight 1e-8 there
    masks[1, :, 38:40] = 0.5
    mn, mx = 0.55, 100.0
    samp, raw = O64.blend_lut(intr, rot.reshape(2, 9), trans, masks, PH, PW, mn, mx)
    y, x = np.meshgrid(np.arange(PH), np.arange(PW), indexing="ij")
    phi = (y + 0.5) * PI_REF / PH - PI_REF / 2
    th = (x + 0.5) * 2 * PI_REF / PW + PI_REF
    unit = np.stack([np.cos(phi) * np.sin(th), np.sin(phi), np.cos(phi) * np.cos(th)], -1)
    stacked = masks.reshape(2 * H, W)
    bw = []
    for r in range(2):
        ray = unit @ rot[r].T
        uv, ok0 = _ds_project(intr[r], ray)
        want_uv = np.stack([np.clip(uv[..., 0], 0.1, W - 1.1), np.clip(uv[..., 1], 0.1, H - 1.1)], -1)
        assert np.abs(samp[r] - want_uv).max() < 1e-9
        pn, ok1 = _ds_project(intr[r], ray * mn - trans[r])
        pf, ok2 = _ds_project(intr[r], ray * mx - trans[r])
        pn = np.stack([np.clip(pn[..., 0], 0.1, W - 1.1), r * H + np.clip(pn[..., 1], 0.1, H - 1.1)], -1)
        pf = np.stack([np.clip(pf[..., 0], 0.1, W - 1.1), r * H + np.clip(pf[..., 1], 0.1, H - 1.1)], -1)
        ok = ok0 & ok1 & ok2 & (_bilinear(stacked, pn[..., 0], pn[..., 1]) > 0.99) & \
            (_bilinear(stacked, pf[..., 0], pf[..., 1]) > 0.99)
        d2 = ((pf - pn) ** 2).sum(-1)
        bw.append(np.where(ok, np.exp(-d2 / (1e-4 * H * W)), 1e-8))
    bw = np.stack(bw)
    want = bw / bw.sum(0)
    keep = bw.sum(0) > 1e-250
    assert keep.mean() > 0.9
    both = keep & (bw[0] > 1e-3) & (bw[1] > 1e-3)
    assert both.sum() > 200, "pixels where both references carry a real weight"
    assert np.any(keep & (bw[1] == 1e-8) & (bw[0] > 1e-6)), "pixels where camera 1's mask rules it out"
    assert np.abs(raw[:, keep] - want[:, keep]).max() < 1e-9
    # with the panorama's size as the normaliser the weights would differ
    bw_pano = np.where(bw == 1e-8, 1e-8, bw ** (H * W / (PH * PW)))
    assert np.abs((bw_pano / bw_pano.sum(0))[:, both] - want[:, both]).max() > 0.05
    # the fp32 oracle, the bit-exact target of the GPU kernel, agrees to fp32 rounding
    samp32, raw32 = O32.blend_lut(intr, rot.reshape(2, 9), trans, masks.astype(np.float32), PH, PW, mn, mx)
    assert np.abs(raw32[:, both] - want[:, both]).max() < 1e-4


def test_merge_closed_form():
    """mergeRGBDPanoramaKernel (stitcher.cu:436-471) on non-constant inputs
    and non-zero translations.  Each panorama pixel samples a direction d_r
    of reference r at uv_r = project(d_r); the reprojected and original
    distance maps are linear in (u, v), so their bilinear reads are known;
    the colour comes from the point min(reprojected, original) * d_r +
    t_r, projected into camera r, scaled to the stitch image's size (no
    half-pixel shift), clamped to [0.1, size - 1.1] and read from a stitch
    image that is linear in (x, y); distance = 1 / sum_r w_r / reprojected,
    colour = sum_r w_r * colour_r truncated to uint8.  All of it derived
    here in float64 from the projection's definition."""
    H, W, intr, rot, trans = _stitch_rig()
    trans = trans * 2.5  # large enough that a sign slip moves the colour samples by pixels
    Hs, Ws = 96, 80
    PH, PW = 12, 20
    rng = np.random.RandomState(5)
    dirs, samp = [], np.zeros((2, PH, PW, 2))
    for r in range(2):
        v = rng.randn(PH, PW, 3) * [0.3, 0.3, 0.2] + [0, 0, 1.0]  # well inside the field of view
        v /= np.linalg.norm(v, axis=-1, keepdims=True)
        uv, ok = _ds_project(intr[r], v)
        assert ok.all() and (uv > 2).all() and (uv < W - 3).all()
        dirs.append(v)
        samp[r] = uv
    w0 = rng.uniform(0.1, 0.9, (PH, PW))
    weights = np.stack([w0, 1 - w0])
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    reproj = np.stack([1.0 + 0.02 * xx + 0.01 * yy, 1.2 + 0.01 * xx + 0.015 * yy])
    dists = np.stack([2.2 - 0.02 * xx + 0.005 * yy, 0.9 + 0.03 * xx - 0.01 * yy])  # crosses reproj: min switches
    sy, sx = np.meshgrid(np.arange(Hs), np.arange(Ws), indexing="ij")
    lin = [lambda X, Y: 10 + X + Y, lambda X, Y: 20 + 2 * X, lambda X, Y: 250 - 2 * Y]
    img = np.stack([np.stack([f(sx, sy) for f in lin], -1)] * 2).astype(np.uint8)
    dist, rgb = O64.merge(samp, weights, reproj, dists, img, trans, intr)
    inv, col = 0.0, 0.0
    switched = 0
    for r in range(2):
        u, v = samp[r][..., 0], samp[r][..., 1]
        rd = _bilinear(reproj[r], u, v)
        dc = np.minimum(rd, _bilinear(dists[r], u, v))
        switched += (dc < rd).sum()
        inv = inv + weights[r] / rd
        q, _ = _ds_project(intr[r], dirs[r] * dc[..., None] + trans[r])
        qx = np.clip(q[..., 0] * (Ws / W), 0.1, Ws - 1.1)
        qy = np.clip(q[..., 1] * (Hs / H), 0.1, Hs - 1.1)
        col = col + weights[r][..., None] * np.stack([f(qx, qy) for f in lin], -1)
    assert 0 < switched < 2 * PH * PW
    assert np.abs(dist - 1 / inv).max() < 1e-12 * np.abs(dist).max()
    near_int = np.abs(col - np.round(col)) < 1e-6
    assert np.array_equal(rgb[~near_int], np.floor(col[~near_int]).astype(np.uint8))
    # a translation of the wrong sign moves every colour sample by pixels
    _, flipped = O64.merge(samp, weights, reproj, dists, img, -trans, intr)
    assert np.abs(flipped.astype(int) - rgb.astype(int)).mean() > 3
    # the fp32 oracle (the GPU kernel's bit-exact target) within its rounding
    dist32, rgb32 = O32.merge(samp, weights, reproj, dists, img, trans, intr)
    assert np.abs(dist32 - 1 / inv).max() < 1e-5 * np.abs(dist).max()
    assert np.abs(rgb32.astype(int) - np.floor(col)).max() <= 1
