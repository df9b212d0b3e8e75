"""Parity of every HIP stage with the CPU oracle (bit-exact), through the C ABI.

Sizes are small enough for the oracle to finish in seconds; the odd shapes
exercise the border clamps and the setPlusX/setPlusY guards of the filter.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import first_mismatch, frame, rendered, to_np
from oracle import Oracle
from pipeline import OraclePipeline
from sphere_stereo import FrameBatch, ISB_Filter, RGBD_Estimator, _lib
from sphere_stereo.utils import camera_array, distance_candidates, relative_transforms

pytestmark = pytest.mark.gpu

O32 = Oracle("f32")


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def assert_same(a, b, what):
    msg = first_mismatch(a, b)
    assert msg is None, "%s: %s" % (what, msg)


def assert_matches_oracle(calibs, kw, to_match, to_stitch, rgb, dist, what, **extra):
    """A single-frame result against the oracle's run of the same frame (so
    that a batched result compared with it is compared with the oracle)."""
    want_rgb, want_dist = OraclePipeline(calibs, **kw, **extra).estimate_RGBD_panorama(to_match, to_stitch)
    assert_same(to_np(dist), want_dist, "%s: distance panorama" % what)
    assert_same(to_np(rgb), want_rgb, "%s: rgb panorama" % what)


def test_guide_bit_exact(gpu):
    g = torch.Generator().manual_seed(3)
    img = (torch.rand([37, 53, 3], generator=g) * 255).to(gpu)
    out = torch.empty([37, 53], dtype=torch.int32, device=gpu)
    _lib.check(_lib.lib.ss_guide(img.data_ptr(), 37, 53, out.data_ptr(), _stream()))
    packed = to_np(out).view(np.uint32)
    got = np.stack([(packed >> (8 * c)) & 255 for c in range(3)], -1).astype(np.uint8)
    assert_same(got, O32.guide(to_np(img)), "guide")


@pytest.mark.parametrize("shape,N", [((8, 8), 1), ((9, 13), 3), ((33, 17), 8), ((48, 64), 32), ((31, 45), 48),
                                     ((16, 20), 64), ((24, 40), 128)])
def test_isb_filter_planar_bit_exact(gpu, shape, N):
    H, W = shape
    g = torch.Generator().manual_seed(H * 131 + W + N)
    guide = torch.randint(0, 256, [H, W, 3], generator=g, dtype=torch.uint8)
    # piecewise-smooth guide so that the edge weights vary
    guide[: H // 2] = guide[: H // 2] // 8 + 100
    cost = torch.rand([N, H, W], generator=g) * 500
    want_c, want_g = O32.isb_filter(guide.numpy(), cost.numpy(), 10.0, 25.0)
    f = ISB_Filter(N, (W, H), gpu)
    gd, cd = guide.to(gpu), cost.to(gpu)
    rc, rg = f.apply(gd, cd, 10.0, 25.0)
    assert rc.data_ptr() == cd.data_ptr() and rg.data_ptr() == gd.data_ptr()
    assert_same(to_np(gd), want_g, "guide")
    assert_same(to_np(cd), want_c, "cost")


@pytest.mark.parametrize("iters", [1, 5, 8, 13, 16, 17, 32])
def test_inpaint_bit_exact(gpu, iters):
    H, W = 97, 150
    intr = [0.3 * 150, 0.3 * 150, 75.0, 48.0, -0.18, 0.59]
    t = [0.08, -0.01, 0.03]
    rec = _lib.ss_intrinsics(*[float(np.float32(v)) for v in intr])
    lut = torch.empty([H, W, 2], dtype=torch.uint8, device=gpu)
    _lib.check(_lib.lib.ss_inpaint_lut(rec, _lib.float_array(t), H, W, 0.55, 100, lut.data_ptr(), _stream()))
    want_lut = O32.inpaint_lut([float(np.float32(v)) for v in intr], [float(np.float32(v)) for v in t], H, W,
                               0.55, 100)
    assert_same(to_np(lut), want_lut, "inpaint LUT")
    rng = np.random.RandomState(iters)
    d = (1 + rng.rand(H, W) * 50).astype(np.float32)
    for _ in range(12):  # holes of up to ~40 px, some touching the border
        y, x, h, w = rng.randint(0, H), rng.randint(0, W), rng.randint(1, 40), rng.randint(1, 25)
        d[y:y + h, x:x + w] = 1e8
    d[rng.rand(H, W) < 0.02] = 1e8
    te, ts = 100.1, float(np.float32(100) + np.float32(0.1))
    want = O32.inpaint(d, want_lut, iters, te, ts)
    m = torch.from_numpy(d).to(gpu)
    tmp = torch.empty_like(m)
    _lib.check(_lib.lib.ss_inpaint(m.data_ptr(), tmp.data_ptr(), lut.data_ptr(), H, W, iters, te, ts, _stream()))
    assert_same(to_np(m), want, "inpainted map (%d passes)" % iters)


@pytest.mark.parametrize("iters,double_form", [(16, False), (32, True), (40, False)])
def test_inpaint_large_holes_bit_exact(gpu, iters, double_form):
    """Holes far wider than the passes can fill (the interior stays empty), a
    band of NaNs, and maps spanning many workgroup blocks."""
    H, W = 300, 420
    intr = [0.3 * W, 0.3 * W, W / 2, H / 2, -0.18, 0.59]
    t = [0.05, 0.02, -0.03]
    rec = _lib.ss_intrinsics(*[float(np.float32(v)) for v in intr])
    lut = torch.empty([H, W, 2], dtype=torch.uint8, device=gpu)
    _lib.check(_lib.lib.ss_inpaint_lut(rec, _lib.float_array(t), H, W, 0.55, 100, lut.data_ptr(), _stream()))
    want_lut = O32.inpaint_lut([float(np.float32(v)) for v in intr], [float(np.float32(v)) for v in t], H, W,
                               0.55, 100)
    assert_same(to_np(lut), want_lut, "inpaint LUT")
    rng = np.random.RandomState(100 + iters)
    d = (1 + rng.rand(H, W) * 50).astype(np.float32)
    d[20:160, 30:250] = 1e8  # wider than 2 * iters: its middle is never reached
    d[200:230, 300:410] = 1e8
    d[250:256, :] = np.nan
    for _ in range(20):
        y, x, h, w = rng.randint(0, H), rng.randint(0, W), rng.randint(1, 60), rng.randint(1, 60)
        d[y:y + h, x:x + w] = 1e8
    d[rng.rand(H, W) < 0.03] = 1e8
    if double_form:
        te, ts = 100.1, 100.0 + float(np.float32(0.1))
    else:
        te, ts = 100.1, float(np.float32(100) + np.float32(0.1))
    want = O32.inpaint(d, want_lut, iters, te, ts)
    assert np.isnan(want).any() and (want == np.float32(1e8)).any() and (want != d).sum() > 1000
    m = torch.from_numpy(d).to(gpu)
    tmp = torch.empty_like(m)
    _lib.check(_lib.lib.ss_inpaint(m.data_ptr(), tmp.data_ptr(), lut.data_ptr(), H, W, iters, te, ts, _stream()))
    assert_same(to_np(m), want, "inpainted map (%d passes)" % iters)


def test_reproject_bit_exact(gpu):
    H, W = 97, 150
    intr = [float(np.float32(v)) for v in [0.3 * 150, 0.3 * 150, 75.0, 48.0, -0.18, 0.59]]
    t = [float(np.float32(v)) for v in [0.2, -0.05, 0.1]]
    rng = np.random.RandomState(1)
    d = (0.6 + rng.rand(H, W) * 30).astype(np.float32)
    d[:, 70:75] = 0.7  # a near strip that occludes its neighbourhood
    want = O32.reproject_zmin(d, intr, t)
    out = torch.empty([H, W], device=gpu)
    _lib.check(_lib.lib.ss_reproject(torch.from_numpy(d).to(gpu).data_ptr(), out.data_ptr(), H, W,
                                     _lib.ss_intrinsics(*intr), _lib.float_array(t), _stream()))
    assert_same(to_np(out), want, "reprojected map")
    assert (want < 1e8).mean() > 0.5 and (want >= 1e8).any()


def test_isb_filter_rejects_tiny(gpu):
    with pytest.raises(ValueError):
        ISB_Filter(4, (3, 3), gpu)


@pytest.fixture(scope="module")
def small(gpu):
    fr = frame(seed=1, resolution=160, matching=(80, 72), stitch=(96, 96), refs=(0, 2), device=gpu, baseline=0.1)
    kw = dict(min_dist=0.55, max_dist=100, candidate_count=32, references_indices=fr["refs"],
              reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
              rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(128, 64), sigma_i=10, sigma_s=12.5)
    est = RGBD_Estimator(fr["calibs"], device=gpu, **kw)
    ora = OraclePipeline(fr["calibs"], **kw)
    return fr, est, ora


def test_select_camera_bit_exact(small):
    fr, est, ora = small
    for slot, r in enumerate(fr["refs"]):
        got = to_np(est.selected_cameras[slot][0])
        assert_same(got, ora.sel[r], "selected camera ref %d" % r)
        assert not np.any(got == r), "the reference camera is never selected"


VOLUME_DTYPES = {0: torch.float32, 1: torch.float16, 2: torch.bfloat16}


@pytest.mark.parametrize("vtype", [0, 1, 2])
def test_sweep_bit_exact(small, gpu, vtype):
    fr, est, ora = small
    H, W = 72, 80
    N = 32
    NP = _lib.lib.ss_padded_candidates(N)
    for slot, r in enumerate(fr["refs"]):
        cost = torch.empty([H, W, NP], dtype=VOLUME_DTYPES[vtype], device=gpu)
        guide = torch.empty([H, W], dtype=torch.int32, device=gpu)
        imgs = [i.contiguous() for i in fr["to_match"]]
        _lib.check(_lib.lib.ss_sweep(est._raylut[slot].data_ptr(), _lib.ptr_array(imgs), len(imgs), est._cams,
                                     est._rts[r], r, est._cands, N, H, W, vtype, cost.data_ptr(), guide.data_ptr(),
                                     _stream()))
        want = O32.sweep(np.stack([to_np(i) for i in imgs]), ora.sel[r], ora.cands, ora.cams, ora.rts[r], r, vtype)
        assert_same(np.transpose(to_np(cost.float())[..., :N], (2, 0, 1)), want, "cost volume ref %d" % r)
        packed = to_np(guide).view(np.uint32)
        got_g = np.stack([(packed >> (8 * c)) & 255 for c in range(3)], -1).astype(np.uint8)
        assert_same(got_g, O32.guide(to_np(imgs[r])), "guide written by the sweep, ref %d" % r)


def _fma32(a, b, c):
    """fmaf in float32 (the fp64 sum of an exact fp32 product and an fp32
    value, rounded to fp32; double rounding is too rare to matter here)."""
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


@pytest.mark.parametrize("case", ["point_at_camera_centre", "parameters_out_of_fast_bounds"])
def test_sweep_projection_edge_cases(small, gpu, case):
    """The sweep projects with exact fast square roots and divisions where the
    projection's denominator is not tiny and the camera parameters are within
    the bounds the host checks.  Here one sample lands exactly on the other
    camera's centre (a zero point: the full forms' special cases), or a
    camera's xi leaves the bounds (every projection takes the full forms)."""
    fr, est, ora = small
    H, W, N = 72, 80, 32
    NP = _lib.lib.ss_padded_candidates(N)
    r, slot = fr["refs"][0], 0
    rts_gpu = _lib.float_array(list(est._rts[r]))
    rts_ora = np.array(ora.rts[r], dtype=np.float64).copy()
    cams_gpu = (_lib.ss_camera * len(est._cams))()
    C.memmove(cams_gpu, est._cams, C.sizeof(cams_gpu))
    cams_ora = np.array(ora.cams, dtype=np.float64).copy()
    if case == "point_at_camera_centre":
        lut = to_np(est._raylut[slot]).reshape(H, W, 4)
        ys, xs = np.nonzero(ora.sel[r] >= 0)
        y, x = int(ys[len(ys) // 2]), int(xs[len(xs) // 2])
        k = int(ora.sel[r][y, x])
        d = np.float32(est._cands_list[5])
        p = [np.float32(d * np.float32(lut[y, x, i])) for i in range(3)]
        for row in range(3):
            m = [np.float32(rts_gpu[12 * k + 4 * row + i]) for i in range(3)]
            q = _fma32(p[2], m[2], _fma32(p[1], m[1], np.float32(p[0] * m[0])))
            rts_gpu[12 * k + 4 * row + 3] = float(-q)
            rts_ora[12 * k + 4 * row + 3] = float(-q)
    else:
        k = [c for c in range(len(est._cams)) if c != r][0]
        cams_gpu[k].xi = 20.0
        cams_ora[k][4] = 20.0
    cost = torch.empty([H, W, NP], dtype=torch.float32, device=gpu)
    imgs = [i.contiguous() for i in fr["to_match"]]
    _lib.check(_lib.lib.ss_sweep(est._raylut[slot].data_ptr(), _lib.ptr_array(imgs), len(imgs), cams_gpu, rts_gpu, r,
                                 est._cands, N, H, W, 0, cost.data_ptr(), None, _stream()))
    want = O32.sweep(np.stack([to_np(i) for i in imgs]), ora.sel[r], ora.cands, cams_ora, rts_ora, r, 0)
    assert_same(np.transpose(to_np(cost)[..., :N], (2, 0, 1)), want, "cost volume (%s)" % case)


@pytest.mark.parametrize("vtype", [1, 2])
@pytest.mark.parametrize("shape,N", [((33, 17), 8), ((48, 64), 32), ((31, 45), 48), ((24, 40), 128)])
def test_isb_filter_precision_form_bit_exact(gpu, vtype, shape, N):
    H, W = shape
    NP = _lib.lib.ss_padded_candidates(N)
    g = torch.Generator().manual_seed(H * 7 + W + N + vtype)
    guide = torch.randint(0, 256, [H, W, 3], generator=g, dtype=torch.uint8)
    guide[: H // 2] = guide[: H // 2] // 8 + 100
    cost = torch.rand([N, H, W], generator=g) * 500
    cost16 = O32.round_volume(cost.numpy(), vtype)
    want_c, want_g = O32.isb_filter(guide.numpy(), cost16, 10.0, 25.0, vtype=vtype, round_level0=True)
    vol = torch.zeros([H, W, NP], dtype=VOLUME_DTYPES[vtype])
    vol[..., :N] = torch.from_numpy(cost16).permute(1, 2, 0).to(VOLUME_DTYPES[vtype])
    vol = vol.to(gpu)
    g32 = guide.to(torch.int32)
    packed = (g32[..., 0] | (g32[..., 1] << 8) | (g32[..., 2] << 16)).contiguous().to(gpu)
    nbytes = _lib.lib.ss_isb_workspace_bytes(H, W, N, vtype, 0)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=gpu)
    _lib.check(_lib.lib.ss_isb_filter(packed.data_ptr(), vol.data_ptr(), H, W, N, vtype, 10.0, 25.0, ws.data_ptr(),
                                      nbytes, _stream()))
    got = np.transpose(to_np(vol.float())[..., :N], (2, 0, 1))
    assert_same(got, want_c, "filtered %s volume" % VOLUME_DTYPES[vtype])
    pk = to_np(packed).view(np.uint32)
    assert_same(np.stack([(pk >> (8 * c)) & 255 for c in range(3)], -1).astype(np.uint8), want_g, "guide")


@pytest.mark.parametrize("volume_dtype", ["float16", "bfloat16"])
def test_panorama_precision_form_bit_exact(small, gpu, volume_dtype):
    fr = small[0]
    kw = dict(min_dist=0.55, max_dist=100, candidate_count=32, references_indices=fr["refs"],
              reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
              rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(128, 64), sigma_i=10, sigma_s=12.5)
    est = RGBD_Estimator(fr["calibs"], device=gpu, volume_dtype=volume_dtype, **kw)
    ora = OraclePipeline(fr["calibs"], volume_dtype=volume_dtype, **kw)
    keep = {}
    want_rgb, want_dist = ora.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"], keep=keep)
    rgb, dist = est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])
    for i in range(len(fr["refs"])):
        assert_same(to_np(est._distance_maps[i]), keep["dmaps"][i], "%s distance map %d" % (volume_dtype, i))
    assert_same(to_np(dist), want_dist, "%s distance panorama" % volume_dtype)
    assert_same(to_np(rgb), want_rgb, "%s rgb panorama" % volume_dtype)


def test_estimate_distance_bit_exact(small):
    fr, est, ora = small
    for slot, r in enumerate(fr["refs"]):
        keep = {}
        want = ora.estimate_fisheye_distance(fr["to_match"], r, keep=keep)
        calib = fr["calibs"][r]
        got = est.estimate_fisheye_distance(None, None, calib, None, fr["to_match"])
        assert_same(to_np(got), want, "distance map ref %d" % r)


def test_estimate_distance_with_given_selection(small, gpu):
    """estimate_fisheye_distance matches each pixel against the camera the
    caller's ``selected_camera`` names (depth_estimation.py:155-170), not the
    selection made at construction: here a selection with every camera index
    (the reference itself included), -1 and indices that name no camera."""
    fr, est, ora = small
    H, W = 72, 80
    rng = np.random.RandomState(11)
    for slot, r in enumerate(fr["refs"]):
        base = ora.sel[r]
        given = np.where(rng.rand(H, W) < 0.3, rng.randint(-1, 4, size=(H, W)), base).astype(np.int64)
        given[rng.rand(H, W) < 0.05] = 9  # no camera of the rig
        given[:5] = -7
        assert (given != base).mean() > 0.2
        want = ora.estimate_fisheye_distance(fr["to_match"], r, sel=given)
        got = est.estimate_fisheye_distance(None, None, fr["calibs"][r],
                                            torch.from_numpy(given).unsqueeze(0).to(gpu), fr["to_match"])
        assert_same(to_np(got), want, "distance map ref %d with a given selection" % r)
        # the selection made at construction takes the cached table, same result as the oracle's own
        got0 = est.estimate_fisheye_distance(None, None, fr["calibs"][r], est.selected_cameras[slot],
                                             fr["to_match"])
        assert_same(to_np(got0), ora.estimate_fisheye_distance(fr["to_match"], r), "distance map ref %d" % r)
        assert not np.array_equal(to_np(got0), want)


def test_stitch_tables_bit_exact(small):
    fr, est, ora = small
    st = est.fishey_stitcher
    for i in range(len(fr["refs"])):
        assert_same(to_np(st.inpainting_weights_list[i]), ora.luts[i], "inpaint LUT %d" % i)
    assert_same(to_np(st.blending_sampling), ora.sampling, "blend sampling")
    assert_same(to_np(st.blending_weights), ora.weights, "blend weights")


def test_panorama_bit_exact(small):
    fr, est, ora = small
    keep = {}
    want_rgb, want_dist = ora.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"], keep=keep)
    rgb, dist = est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])
    for i in range(len(fr["refs"])):
        assert_same(to_np(est._distance_maps[i]), keep["dmaps"][i], "distance map %d" % i)
        assert_same(to_np(est.fishey_stitcher.reprojected_distances[i]), keep["reproj"][i], "reprojected %d" % i)
    assert_same(to_np(dist), want_dist, "distance panorama")
    assert_same(to_np(rgb), want_rgb, "rgb panorama")


def test_panorama_deterministic(small):
    fr, est, ora = small
    rgb1, dist1 = [t.clone() for t in est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])]
    rgb2, dist2 = est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])
    assert torch.equal(rgb1, rgb2)
    assert torch.equal(dist1.view(torch.int32), dist2.view(torch.int32))


def test_inverse_distance_and_error_stats(gpu):
    from sphere_stereo.evaluation import error_stats, inverse_distance

    rng = np.random.RandomState(4)
    d = (0.5 + rng.rand(333, 517) * 99.5).astype(np.float32)
    inv = to_np(inverse_distance(torch.from_numpy(d).to(gpu)))
    assert_same(inv, (np.float32(1) / d).astype(np.float32), "inverse distance")
    a = rng.rand(300000).astype(np.float32)
    b = (a + (rng.rand(300000).astype(np.float32) - 0.5) * 0.9).astype(np.float32)
    b[::997] = np.nan
    a[5::1009] = np.inf
    for finite_only in (True, False):
        got = error_stats(torch.from_numpy(a).to(gpu), torch.from_numpy(b).to(gpu), (0.1, 0.4, 0.25), finite_only)
        keep = np.isfinite(a) & np.isfinite(b) if finite_only else np.ones(a.shape, bool)
        err = np.abs(a[keep] - b[keep])  # fp32, as utils.py:233 forms it
        assert got["count"] == keep.sum() and got["skipped"] == (~keep).sum()
        if finite_only:
            assert abs(got["mae"] - err.astype(np.float64).mean()) < 1e-12
            assert abs(got["rmse"] - np.sqrt((err * err).astype(np.float64).mean())) < 1e-12
            for t, ratio in zip((0.1, 0.4, 0.25), got["bad_px_ratios"]):
                assert ratio == 100.0 * np.sum(err > np.float32(t)) / err.size
        else:
            assert np.isnan(got["mae"]) and np.isnan(got["rmse"])


def test_graph_replay_matches_eager(small):
    fr, est, ora = small
    rgb0, dist0 = [t.clone() for t in est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])]
    imgs = [t.contiguous() for t in fr["to_match"]]
    stitch = [t.contiguous() for t in fr["to_stitch"]]
    est.use_graph = True
    try:
        for rep in range(4):
            rgb, dist = est.estimate_RGBD_panorama(imgs, stitch)
            torch.cuda.synchronize()
            assert torch.equal(rgb, rgb0), "replay %d" % rep
            assert torch.equal(dist.view(torch.int32), dist0.view(torch.int32)), "replay %d" % rep
            # eager work and a synchronisation between replays
            junk = torch.randn(1 << 20, device=rgb.device).sum().item()
            assert junk == junk
        # new input tensors: the frame is captured again
        imgs2 = [t.clone() for t in imgs]
        rgb, dist = est.estimate_RGBD_panorama(imgs2, stitch)
        torch.cuda.synchronize()
        assert torch.equal(rgb, rgb0) and torch.equal(dist.view(torch.int32), dist0.view(torch.int32))
    finally:
        est.use_graph = False


@pytest.mark.parametrize("volume_dtype", ["float32", "float16"])
def test_frame_batch_matches_single_frames(gpu, volume_dtype):
    """Nine rigs (own seeds and baselines, one configuration) through the
    batching form: two chunks (8 + 1 frames of 2 references), eager and
    replayed from a graph, each frame bit-exact with its single-frame run."""
    ests, frames, want, kws = [], [], [], []
    for seed in range(9):
        fr = frame(seed=10 + seed, resolution=160, matching=(80, 72), stitch=(96, 96), refs=(0, 2), device=gpu,
                   baseline=0.1 + 0.02 * seed)
        kw = dict(min_dist=0.55, max_dist=100, candidate_count=32, references_indices=fr["refs"],
                  reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
                  rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(128, 64), sigma_i=10, sigma_s=12.5)
        est = RGBD_Estimator(fr["calibs"], device=gpu, volume_dtype=volume_dtype, **kw)
        rgb, dist = est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])
        assert_matches_oracle(fr["calibs"], kw, fr["to_match"], fr["to_stitch"], rgb, dist, "frame %d" % seed,
                              volume_dtype=volume_dtype)
        want.append((rgb.clone(), dist.clone()))
        kws.append((fr["calibs"], kw))
        ests.append(est)
        frames.append(([t.contiguous() for t in fr["to_match"]], [t.contiguous() for t in fr["to_stitch"]]))
    for e in ests:  # clear the outputs so that stale values cannot pass
        e.fishey_stitcher.RGB_panorama.zero_()
        e.fishey_stitcher.distance_panorama.zero_()
    batch = FrameBatch(ests)
    for mode in ("eager", "graph", "graph", "graph"):
        batch.use_graph = mode == "graph"
        out = batch.estimate_RGBD_panoramas(frames)
        torch.cuda.synchronize()
        for i, ((rgb, dist), (wr, wd)) in enumerate(zip(out, want)):
            assert torch.equal(rgb, wr), "%s rgb panorama of frame %d" % (mode, i)
            assert torch.equal(dist.view(torch.int32), wd.view(torch.int32)), "%s distance panorama %d" % (mode, i)
        junk = torch.randn(1 << 20, device=gpu).sum().item()  # eager work between replays
        assert junk == junk
        for e in ests:
            e.fishey_stitcher.distance_panorama.fill_(-1.0)
    # the frames of one native batch share one configuration
    calibs, kw = kws[1]
    other = RGBD_Estimator(calibs, device=gpu, volume_dtype=volume_dtype, **dict(kw, candidate_count=16))
    descs = (_lib.ss_frame_desc * 2)(ests[0]._frame_desc(), other._frame_desc())
    handle = _lib.C.c_void_p()
    assert _lib.lib.ss_frame_batch_create(descs, 2, _lib.C.byref(handle)) == 1  # SS_ERR_ARG
    assert b"differ" in _lib.lib.ss_last_error()


def _lane_frames(gpu, count, refs, seed0):
    ests, frames = [], []
    for i in range(count):
        fr = frame(seed=seed0 + i, resolution=160, matching=(64, 48), stitch=(64, 64), refs=refs, device=gpu,
                   baseline=0.12 + 0.03 * i)
        kw = dict(min_dist=0.55, max_dist=100, candidate_count=16, references_indices=fr["refs"],
                  reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
                  rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(128, 64), sigma_i=10, sigma_s=12.5)
        ests.append((RGBD_Estimator(fr["calibs"], device=gpu, **kw), fr, kw))
        frames.append(([t.contiguous() for t in fr["to_match"]], [t.contiguous() for t in fr["to_stitch"]]))
    return ests, frames


def test_frame_lanes_identical(gpu):
    """ss_frame_set_lanes: the chunks of a run spread over 2 to 4 streams
    (sweeps one after the other, each chunk's later passes beside the next
    chunk's sweeps) give the bits of the one-stream run, eager and replayed
    from graphs: nine frames of four references (three chunks of up to four
    frames, so a lane takes two chunks) and two frames of three references
    over four lanes (one chunk per frame).  The one-stream results are
    checked against the oracle for two frames."""
    ests, frames = _lane_frames(gpu, 9, (0, 1, 2, 3), 70)
    want = []
    for i, ((est, fr, kw), (tm, ts)) in enumerate(zip(ests, frames)):
        rgb, dist = est.estimate_RGBD_panorama(tm, ts)
        if i in (0, 8):
            assert_matches_oracle(fr["calibs"], kw, fr["to_match"], fr["to_stitch"], rgb, dist, "frame %d" % i)
        want.append((rgb.clone(), dist.clone()))

    def check(out, wants, what):
        torch.cuda.synchronize()
        for i, ((rgb, dist), (wr, wd)) in enumerate(zip(out, wants)):
            assert torch.equal(rgb, wr), "%s: rgb panorama of frame %d" % (what, i)
            assert torch.equal(dist.view(torch.int32), wd.view(torch.int32)), "%s: distance panorama %d" % (what, i)

    def clear(estimators):
        for e in estimators:
            e.fishey_stitcher.RGB_panorama.zero_()
            e.fishey_stitcher.distance_panorama.fill_(-1.0)

    batch = FrameBatch([e for e, _, _ in ests])
    for lanes in (2, 3, 4, 1):
        batch.lanes = lanes
        assert batch.lanes == lanes and all(_lib.lib.ss_frame_get_lanes(h) == lanes for h in batch._frames)
        for mode in ("eager", "graph", "graph"):
            clear(e for e, _, _ in ests)
            batch.use_graph = mode == "graph"
            check(batch.estimate_RGBD_panoramas(frames), want, "9 frames, %d lanes, %s" % (lanes, mode))
            junk = torch.randn(1 << 20, device=gpu).sum().item()  # eager work between replays
            assert junk == junk
    batch.use_graph = False
    del batch

    ests3, frames3 = _lane_frames(gpu, 2, (0, 1, 3), 90)
    want3 = []
    for (est, _, _), (tm, ts) in zip(ests3, frames3):
        rgb, dist = est.estimate_RGBD_panorama(tm, ts)
        want3.append((rgb.clone(), dist.clone()))
    batch3 = FrameBatch([e for e, _, _ in ests3], lanes=4)
    for mode in ("eager", "graph", "graph"):
        clear(e for e, _, _ in ests3)
        batch3.use_graph = mode == "graph"
        check(batch3.estimate_RGBD_panoramas(frames3), want3, "2 frames of 3 references, 4 lanes, %s" % mode)
    with pytest.raises(_lib.SphereStereoError, match="lanes outside"):
        batch3.lanes = 5
    batch3.use_graph = False


def test_frame_batch_ragged_matches_single_frames(gpu):
    """A ragged batch: rigs of different matching resolutions, candidate
    counts, reference sets and volume types (FrameBatch groups them by
    configuration); every frame bit-exact with its single-frame run, eager and
    replayed from graphs."""
    cases = [((80, 72), 32, (0, 2), "float32"), ((64, 64), 16, (0, 1, 2), "float16"), ((80, 72), 32, (0, 2), "float32"),
             ((96, 80), 48, (0, 1, 2, 3), "float32"), ((64, 64), 16, (0, 1, 2), "float16"), ((80, 72), 24, (1, 3),
                                                                                         "bfloat16")]
    ests, frames, want = [], [], []
    for i, (matching, N, refs, vt) in enumerate(cases):
        fr = frame(seed=30 + i, resolution=160, matching=matching, stitch=(96, 96), refs=refs, device=gpu,
                   baseline=0.15 + 0.05 * i)
        kw = dict(min_dist=0.55, max_dist=100, candidate_count=N, references_indices=fr["refs"],
                  reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
                  rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(128, 64), sigma_i=10, sigma_s=12.5)
        est = RGBD_Estimator(fr["calibs"], device=gpu, volume_dtype=vt, **kw)
        rgb, dist = est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])
        assert_matches_oracle(fr["calibs"], kw, fr["to_match"], fr["to_stitch"], rgb, dist, "frame %d" % i,
                              volume_dtype=vt)
        want.append((rgb.clone(), dist.clone()))
        ests.append(est)
        frames.append(([t.contiguous() for t in fr["to_match"]], [t.contiguous() for t in fr["to_stitch"]]))
    for e in ests:
        e.fishey_stitcher.RGB_panorama.zero_()
        e.fishey_stitcher.distance_panorama.zero_()
    batch = FrameBatch(ests)
    assert sorted(len(g) for g in batch.groups) == [1, 1, 2, 2]
    for mode in ("eager", "graph", "graph"):
        batch.use_graph = mode == "graph"
        out = batch.estimate_RGBD_panoramas(frames)
        torch.cuda.synchronize()
        for i, ((rgb, dist), (wr, wd)) in enumerate(zip(out, want)):
            assert torch.equal(rgb, wr), "%s rgb panorama of frame %d" % (mode, i)
            assert torch.equal(dist.view(torch.int32), wd.view(torch.int32)), "%s distance panorama %d" % (mode, i)
        for e in ests:
            e.fishey_stitcher.distance_panorama.fill_(-1.0)


@pytest.mark.parametrize("vt,schedule", [("float32", "persistent"), ("float16", "persistent"), ("bfloat16", "persistent"),
                                         ("float32", "dispatched"), ("float16", "dispatched"), ("float16", "auto"),
                                         ("bfloat16", "auto")])
def test_frame_batch_single_executor_matches_single_frames(gpu, vt, schedule):
    """The ragged persistent form: frames of different matching resolutions
    (pyramids of different depths, with and without a coarse tail),
    candidate counts (V = 2, 3, 4, 6, 8), reference and camera sets, sigmas,
    inpainting pass counts (an odd and an even number of 16-pass launches,
    and none), stitch sizes and panoramas in one executor, one launch per
    pass; every frame bit-exact with its own single-frame run, eager and over
    graph replays, in each scheduling form."""
    from sphere_stereo import _lib

    cases = [((80, 72), 32, (0, 2), 12.5, 32, (96, 96), (128, 64)),
             ((64, 64), 16, (0, 1, 2), 10.0, 16, (96, 96), (96, 48)),
             ((96, 80), 48, (0, 1, 2, 3), 15.0, 40, (80, 88), (128, 64)),
             ((32, 32), 24, (1, 3), 12.5, 0, (96, 96), (64, 32)),
             ((128, 112), 64, (0, 2), 25.0, 20, (112, 112), (160, 80)),
             ((80, 72), 32, (0, 2), 12.5, 32, (96, 96), (128, 64))]
    ests, frames, want = [], [], []
    for i, (matching, N, refs, sig_s, iters, stitch, pano) in enumerate(cases):
        fr = frame(seed=50 + i, resolution=160, matching=matching, stitch=stitch, refs=refs, device=gpu,
                   baseline=0.15 + 0.04 * i)
        kw = dict(min_dist=0.55, max_dist=100, candidate_count=N, references_indices=fr["refs"],
                  reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
                  rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=pano, sigma_i=10, sigma_s=sig_s)
        est = RGBD_Estimator(fr["calibs"], device=gpu, volume_dtype=vt, **kw)
        if iters != est.fishey_stitcher.inpainting_iterations:
            est.fishey_stitcher.inpainting_iterations = iters
            _lib.lib.ss_frame_destroy(est._frame)
            est._frame = est._make_frame()
        print("persistent case %d: %s N=%d refs=%s iters=%d %s" % (i, matching, N, refs, iters, vt), flush=True)
        rgb, dist = est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])
        assert_matches_oracle(fr["calibs"], kw, fr["to_match"], fr["to_stitch"], rgb, dist, "case %d" % i,
                              volume_dtype=vt, inpainting_iterations=iters)
        want.append((rgb.clone(), dist.clone()))
        ests.append(est)
        frames.append(([t.contiguous() for t in fr["to_match"]], [t.contiguous() for t in fr["to_stitch"]]))
    for e in ests:
        e.fishey_stitcher.RGB_panorama.zero_()
        e.fishey_stitcher.distance_panorama.zero_()
    batch = FrameBatch(ests, single_executor=True, schedule=schedule)
    assert batch.groups == [list(range(len(cases)))]
    for mode in ("eager", "eager", "graph", "graph", "graph"):
        batch.use_graph = mode == "graph"
        out = batch.estimate_RGBD_panoramas(frames)
        torch.cuda.synchronize()
        for i, ((rgb, dist), (wr, wd)) in enumerate(zip(out, want)):
            assert torch.equal(rgb, wr), "%s rgb panorama of frame %d" % (mode, i)
            assert torch.equal(dist.view(torch.int32), wd.view(torch.int32)), "%s distance panorama %d" % (mode, i)
        for e in ests:
            e.fishey_stitcher.distance_panorama.fill_(-1.0)
            e.fishey_stitcher.RGB_panorama.fill_(7)


@pytest.mark.parametrize("schedule", ["dispatched", "persistent"])
def test_frame_batch_single_executor_many_pairs(gpu, schedule):
    """Frame-reference pairs of 17 rigs in one executor (13 rigs of 4
    cameras x 4 references, 4 rigs of 3 cameras x 3 references: 64 pairs,
    and each frame's images at its own offset of the run's image table): the
    passes have more than 63 entries, where an item's entry is found by
    binary search instead of the wavefront ballot; bit-exact against the
    single-frame runs."""
    ests, frames, want = [], [], []
    for i in range(17):
        cams = 3 if i % 4 == 1 else 4
        refs = tuple(range(cams))
        fr = frame(seed=80 + i, resolution=160, matching=(32, 32) if i % 2 else (40, 32), stitch=(48, 48),
                   refs=refs, device=gpu, baseline=0.2 + 0.01 * i)
        calibs, masks, to_match = fr["calibs"][:cams], fr["masks"][:cams], fr["to_match"][:cams]
        kw = dict(min_dist=0.55, max_dist=100, candidate_count=16 if i % 3 else 24, references_indices=list(refs),
                  reprojection_viewpoint=fr["viewpoint"], masks=masks, matching_resolution=fr["matching"],
                  rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(64, 32), sigma_i=10, sigma_s=12.5)
        est = RGBD_Estimator(calibs, device=gpu, volume_dtype="float16", **kw)
        rgb, dist = est.estimate_RGBD_panorama(to_match, fr["to_stitch"])
        assert_matches_oracle(calibs, kw, to_match, fr["to_stitch"], rgb, dist, "rig %d" % i, volume_dtype="float16")
        want.append((rgb.clone(), dist.clone()))
        ests.append(est)
        frames.append(([t.contiguous() for t in to_match], [t.contiguous() for t in fr["to_stitch"]]))
    print("many pairs: 17 frames x 4 references, %s" % schedule, flush=True)
    for e in ests:
        e.fishey_stitcher.distance_panorama.fill_(-1.0)
    batch = FrameBatch(ests, single_executor=True, schedule=schedule)
    for mode in ("eager", "graph", "graph"):
        batch.use_graph = mode == "graph"
        out = batch.estimate_RGBD_panoramas(frames)
        torch.cuda.synchronize()
        for i, ((rgb, dist), (wr, wd)) in enumerate(zip(out, want)):
            assert torch.equal(rgb, wr), "%s rgb panorama of frame %d" % (mode, i)
            assert torch.equal(dist.view(torch.int32), wd.view(torch.int32)), "%s distance panorama %d" % (mode, i)
        for e in ests:
            e.fishey_stitcher.distance_panorama.fill_(-1.0)


def test_frame_batch_single_executor_rejects(gpu):
    """The single executor needs one volume type and N >= 2."""
    ests = []
    for i, (N, vt) in enumerate([(32, "float32"), (16, "float16")]):
        fr = frame(seed=70 + i, resolution=160, matching=(64, 64), refs=(0, 2), device=gpu)
        kw = dict(min_dist=0.55, max_dist=100, candidate_count=N, references_indices=fr["refs"],
                  reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
                  rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(64, 32), sigma_i=10, sigma_s=12.5)
        ests.append(RGBD_Estimator(fr["calibs"], device=gpu, volume_dtype=vt, **kw))
    with pytest.raises(ValueError):
        FrameBatch(ests, single_executor=True)
    FrameBatch(ests)  # the grouped form takes them


@pytest.mark.parametrize("dtype", [torch.uint8, torch.uint16, torch.float32])
@pytest.mark.parametrize("shape,size", [((1216, 1216, 3), (1024, 1024)), ((1216, 1216, 3), (512, 512)),
                                        ((100, 90, 3), (137, 121)), ((7, 9, 1), (3, 4)), ((64, 48), (64, 48))])
def test_resize_linear_bit_exact(gpu, dtype, shape, size):
    """The input stage (utils.py:157-204) on the GPU against the oracle: the
    conversion to float32 in [0, 255] and the INTER_LINEAR resize."""
    from sphere_stereo.utils import resize_linear

    rows, cols = size
    g = torch.Generator().manual_seed(sum(shape) + rows)
    hi = {torch.uint8: 256, torch.uint16: 65536, torch.float32: 256}[dtype]
    img = torch.randint(0, hi, shape, generator=g, dtype=torch.int64)
    img = img.to(torch.float32) if dtype == torch.float32 else img.to(dtype)
    got = resize_linear(img.to(gpu), cols, rows)
    f = img.to(torch.float32).numpy()
    if dtype == torch.uint16:
        f = f / np.float32(255)  # utils.py:176, a float32 division
    assert_same(to_np(got), O32.resize_linear(f.astype(np.float32), cols, rows), "resized %s %s" % (dtype, shape))


def test_streaming_driver_matches_frame_by_frame(gpu):
    """The streaming driver (double-buffered uploads, GPU input stage, graph
    replay, double-buffered downloads) over five frames of decoded uint8
    fisheyes gives, in order, exactly the panoramas of the frame-by-frame
    path (input stage then estimate_RGBD_panorama, then 1 / distance)."""
    from sphere_stereo import StreamingDriver
    from sphere_stereo.evaluation import inverse_distance
    from sphere_stereo.utils import prepare_inputs

    fr = frame(seed=3, resolution=160, matching=(80, 72), stitch=(96, 96), refs=(0, 2), device=gpu, baseline=0.12)
    _, imgs, _ = rendered(3, 160, 0.12)
    kw = dict(min_dist=0.55, max_dist=100, candidate_count=32, references_indices=fr["refs"],
              reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
              rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(128, 64), sigma_i=10, sigma_s=12.5)
    est = RGBD_Estimator(fr["calibs"], device=gpu, **kw)
    raws = [[np.ascontiguousarray(np.roll(imgs[k].numpy(), 5 * i, axis=1)) for k in range(imgs.shape[0])]
            for i in range(5)]
    want = []
    for raw in raws:
        tm, ts = prepare_inputs([torch.from_numpy(a).to(gpu) for a in raw], fr["matching"], fr["stitch"], fr["refs"])
        rgb, dist = est.estimate_RGBD_panorama(tm, ts)
        want.append((to_np(rgb).copy(), to_np(inverse_distance(dist)).copy()))
    assert not np.array_equal(want[0][0], want[1][0])
    drv = StreamingDriver(est, raws[0][0].shape, np.uint8, use_graph=True)
    got = list(drv.run(raws))
    assert [g[0] for g in got] == list(range(5))
    for (i, rgb, inv), (wr, wi) in zip(got, want):
        assert np.array_equal(rgb, wr), "rgb panorama of frame %d" % i
        assert np.array_equal(inv.view(np.int32), wi.view(np.int32)), "inverse distance of frame %d" % i
    est.use_graph = False


@pytest.mark.parametrize("N,vtype", [(32, "float32"), (20, "float16"), (48, "bfloat16")])
def test_sweep_variants_identical_and_autotune(gpu, N, vtype):
    """Every scheduling variant of the sweep writes the same volume and guide
    bits; the autotuner times them all, records the fastest for the
    configuration, and frames then run with it unchanged in result."""
    fr = frame(seed=4, resolution=160, matching=(80, 72), stitch=(96, 96), refs=(0, 2), device=gpu)
    kw = dict(min_dist=0.55, max_dist=100, candidate_count=N, references_indices=fr["refs"],
              reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
              rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(128, 64), sigma_i=10, sigma_s=12.5)
    est = RGBD_Estimator(fr["calibs"], device=gpu, volume_dtype=vtype, **kw)
    lib = _lib.lib
    H, W, vt = est.rows, est.cols, _lib.VOLUME_TYPES[vtype]
    NP = lib.ss_padded_candidates(N)
    eb = 4 if vtype == "float32" else 2
    imgs = [t.contiguous() for t in fr["to_match"]]
    r = est.references_indices[0]
    stream = _lib.C.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    nv = lib.ss_sweep_variant_count()
    assert nv >= 2
    outs = []
    for v in range(nv):
        assert lib.ss_sweep_set_variant(H, W, N, vt, v) == 0
        cost = torch.full((H * W * NP * eb,), 0xAB, dtype=torch.uint8, device=gpu)
        guide = torch.full((H * W,), -1, dtype=torch.int32, device=gpu)
        _lib.check(lib.ss_sweep(est._raylut[0].data_ptr(), _lib.ptr_array(imgs), len(imgs), est._cams, est._rts[r], r,
                                est._cands, N, H, W, vt, cost.data_ptr(), guide.data_ptr(), stream))
        torch.cuda.synchronize()
        outs.append((cost, guide))
    for v in range(1, nv):
        chunk, group = _lib.C.c_int(), _lib.C.c_int()
        assert lib.ss_sweep_variant_info(v, _lib.C.byref(chunk), _lib.C.byref(group)) == 0
        assert torch.equal(outs[v][0], outs[0][0]), "volume of variant %d (chunk %d, group %d)" % (v, chunk.value,
                                                                                                   group.value)
        assert torch.equal(outs[v][1], outs[0][1]), "guide of variant %d" % v
    assert lib.ss_sweep_set_variant(H, W, N, vt, -1) == 0
    assert lib.ss_sweep_get_variant(H, W, N, vt) == 0
    want = [t.clone() for t in est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])]
    res = est.autotune(imgs, repeats=2)
    try:
        assert 0 <= res["variant"] < nv and len(res["ms"]) == nv and all(m > 0 for m in res["ms"]), res
        assert lib.ss_sweep_get_variant(H, W, N, vt) == res["variant"]
        nf = lib.ss_isb_variant_count()
        assert 0 <= res["filter_variant"] < nf and len(res["filter_ms"]) == nf, res
        fm = res["filter_ms"]
        assert all(m > 0 for m in fm)
        # the fastest, unless it is within 2 % of the default (variant 0)
        fv = res["filter_variant"]
        assert (fv == 0 and min(fm) >= 0.98 * fm[0] - 1e-4) or (fm[fv] == min(fm) and fm[fv] < 0.98 * fm[0] + 1e-4)
        assert lib.ss_isb_get_variant(H, W, N, vt) == res["filter_variant"]
        ni = lib.ss_inpaint_variant_count()
        assert 0 <= res["inpaint_variant"] < ni and len(res["inpaint_ms"]) == ni and min(res["inpaint_ms"]) > 0, res
        assert lib.ss_inpaint_get_variant(H, W) == res["inpaint_variant"]
        # the estimator's own frame decides the inpainting tile, with the same margin
        ff = res["inpaint_frame_ms"]
        iv = res["inpaint_variant"]
        assert len(ff) == ni and min(ff) > 0
        assert (iv == 0 and min(ff) >= 0.98 * ff[0] - 1e-4) or (ff[iv] == min(ff) and ff[iv] < 0.98 * ff[0] + 1e-4)
        got = est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])
        torch.cuda.synchronize()
        assert torch.equal(got[0], want[0])
        assert torch.equal(got[1].view(torch.int32), want[1].view(torch.int32))
    finally:
        lib.ss_sweep_set_variant(H, W, N, vt, -1)
        lib.ss_isb_set_variant(H, W, N, vt, -1)
        lib.ss_inpaint_set_variant(H, W, -1)


@pytest.mark.parametrize("vtype", [0, 1])
@pytest.mark.parametrize("shape,N", [((33, 17), 8), ((48, 64), 32), ((31, 45), 48), ((40, 24), 64), ((24, 40), 128)])
def test_isb_filter_variants_identical(gpu, vtype, shape, N):
    """The cost filter's scheduling variants (4 or 8 lanes per pixel, XCD or
    launch tile order) give the oracle's bits; a variant that does not exist
    for the candidate count (4 lanes past 64) is refused."""
    H, W = shape
    lib = _lib.lib
    NP = lib.ss_padded_candidates(N)
    g = torch.Generator().manual_seed(H * 13 + W + N + vtype)
    guide = torch.randint(0, 256, [H, W, 3], generator=g, dtype=torch.uint8)
    guide[: H // 2] = guide[: H // 2] // 8 + 100
    cost = O32.round_volume((torch.rand([N, H, W], generator=g) * 500).numpy(), vtype)
    want_c, want_g = O32.isb_filter(guide.numpy(), cost, 10.0, 25.0, vtype=vtype, round_level0=True)
    vol0 = torch.zeros([H, W, NP], dtype=VOLUME_DTYPES[vtype])
    vol0[..., :N] = torch.from_numpy(cost).permute(1, 2, 0).to(VOLUME_DTYPES[vtype])
    g32 = guide.to(torch.int32)
    packed0 = (g32[..., 0] | (g32[..., 1] << 8) | (g32[..., 2] << 16)).contiguous()
    nbytes = lib.ss_isb_workspace_bytes(H, W, N, vtype, 0)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=gpu)
    try:
        for v in range(lib.ss_isb_variant_count()):
            rc = lib.ss_isb_set_variant(H, W, N, vtype, v)
            if NP > 64 and v >= 2:
                assert rc == 1
                continue
            assert rc == 0
            vol, packed = vol0.clone().to(gpu), packed0.clone().to(gpu)
            _lib.check(lib.ss_isb_filter(packed.data_ptr(), vol.data_ptr(), H, W, N, vtype, 10.0, 25.0,
                                         ws.data_ptr(), nbytes, _stream()))
            got = np.transpose(to_np(vol.float())[..., :N], (2, 0, 1))
            assert_same(got, want_c, "filtered volume, variant %d" % v)
            pk = to_np(packed).view(np.uint32)
            got_g = np.stack([(pk >> (8 * c)) & 255 for c in range(3)], -1).astype(np.uint8)
            assert_same(got_g, want_g, "guide, variant %d" % v)
    finally:
        lib.ss_isb_set_variant(H, W, N, vtype, -1)


@pytest.mark.parametrize("volume_dtype", ["float32", "float16"])
def test_frame_filter_variants_identical(small, gpu, volume_dtype):
    """Whole frames (cost filter with the fused argmin, the distance filter,
    the stitch) in every cost-filter variant: the oracle's panorama."""
    fr, _, _ = small
    kw = dict(min_dist=0.55, max_dist=100, candidate_count=32, references_indices=fr["refs"],
              reprojection_viewpoint=fr["viewpoint"], masks=fr["masks"], matching_resolution=fr["matching"],
              rgb_to_stitch_resolution=fr["stitch"], panorama_resolution=(128, 64), sigma_i=10, sigma_s=12.5)
    want_rgb, want_dist = OraclePipeline(fr["calibs"], volume_dtype=volume_dtype, **kw).estimate_RGBD_panorama(
        fr["to_match"], fr["to_stitch"])
    lib = _lib.lib
    vt = _lib.VOLUME_TYPES[volume_dtype]
    try:
        for v in range(lib.ss_isb_variant_count()):
            assert lib.ss_isb_set_variant(72, 80, 32, vt, v) == 0
            est = RGBD_Estimator(fr["calibs"], device=gpu, volume_dtype=volume_dtype, **kw)
            rgb, dist = est.estimate_RGBD_panorama(fr["to_match"], fr["to_stitch"])
            assert_same(to_np(dist), want_dist, "distance panorama, filter variant %d" % v)
            assert_same(to_np(rgb), want_rgb, "rgb panorama, filter variant %d" % v)
    finally:
        lib.ss_isb_set_variant(72, 80, 32, vt, -1)


@pytest.mark.parametrize("shape,iters", [((97, 150), 13), ((300, 420), 32)])
def test_inpaint_tiles_identical(gpu, shape, iters):
    """Every inpainting tile (48, 32, 16 pixels a side, 16-pixel halo) gives
    the oracle's map: holes wider than the passes reach, holes at the
    border, a band of NaNs."""
    H, W = shape
    lib = _lib.lib
    intr = [0.3 * W, 0.3 * W, W / 2, H / 2, -0.18, 0.59]
    t = [0.05, 0.02, -0.03]
    rec = _lib.ss_intrinsics(*[float(np.float32(v)) for v in intr])
    lut = torch.empty([H, W, 2], dtype=torch.uint8, device=gpu)
    _lib.check(lib.ss_inpaint_lut(rec, _lib.float_array(t), H, W, 0.55, 100, lut.data_ptr(), _stream()))
    want_lut = to_np(lut)
    rng = np.random.RandomState(H + iters)
    d = (1 + rng.rand(H, W) * 50).astype(np.float32)
    d[H // 8:H // 2, W // 10:W // 2] = 1e8
    d[H // 2:H // 2 + 4, :] = np.nan
    for _ in range(20):
        y, x, h, w = rng.randint(0, H), rng.randint(0, W), rng.randint(1, 40), rng.randint(1, 40)
        d[y:y + h, x:x + w] = 1e8
    d[rng.rand(H, W) < 0.03] = 1e8
    te, ts = 100.1, float(np.float32(100) + np.float32(0.1))
    want = O32.inpaint(d, want_lut, iters, te, ts)
    try:
        for v in range(lib.ss_inpaint_variant_count()):
            assert lib.ss_inpaint_set_variant(H, W, v) == 0
            m = torch.from_numpy(d).to(gpu)
            tmp = torch.empty_like(m)
            _lib.check(lib.ss_inpaint(m.data_ptr(), tmp.data_ptr(), lut.data_ptr(), H, W, iters, te, ts, _stream()))
            assert_same(to_np(m), want, "inpainted map, tile variant %d" % v)
    finally:
        lib.ss_inpaint_set_variant(H, W, -1)
