"""The HIP path against the oracle at the workloads' full sizes, bit for bit.

c0: 512^2 matching, 32 candidates, references 0 and 2, sigma_s 12.5,
1024 x 512 panorama.  c1: 1024^2, 32 candidates, references 0 and 2,
sigma_s 25, 1216^2 stitch input, 2048 x 1024 panorama (main.py's defaults,
the measured workload).  c2: 1024^2, 64 candidates, all four cameras as
references, in the fp32, fp16 and bf16 volume forms, and the precision
form's acceptance: each 16-bit form's inverse-distance MAE against the fp64
oracle within 0.002 of the fp32 form's.  c3: the 16 ragged rigs (512^2 to
1024^2, 16 to 64 candidates, 2 to 4 references, fp16 volume) through the
batching form grouped by configuration and through the single executor in
both schedules, every frame against its own oracle run.  Also the README
demo through the dataset loop (python -m sphere_stereo.main), against the
oracle and the reference's own panorama.

At these sizes the pyramid has 8-9 levels and a coarse tail, the stitcher
maps span hundreds of tiles, and the persistent schedules give a workgroup
many items in a row.  Each comparison covers the camera selections, the
per-reference distance maps, the reprojected (inpainted) maps and both
panoramas.
"""
import os
import shutil

import numpy as np
import pytest
import torch

from helpers import DEMO, DEMO_EXPECTED, demo_frame, first_mismatch, oracle_frame, psnr, to_np
from oracle import Oracle
from sphere_stereo import FrameBatch, RGBD_Estimator, synth

pytestmark = pytest.mark.gpu


def assert_same(a, b, what):
    msg = first_mismatch(a, b)
    assert msg is None, "%s: %s" % (what, msg)


def say(msg):
    print(msg, flush=True)


def compare_frame(est, ora, keep, want_rgb, want_dist, rgb, dist, tag):
    for slot, r in enumerate(ora.refs):
        assert_same(to_np(est.selected_cameras[slot][0]), ora.sel[r], "%s: selected cameras of ref %d" % (tag, r))
        assert_same(to_np(est._distance_maps[slot]), keep["dmaps"][slot], "%s: distance map of ref %d" % (tag, r))
        assert_same(to_np(est.fishey_stitcher.reprojected_distances[slot]), keep["reproj"][slot],
                    "%s: reprojected map of ref %d" % (tag, r))
    assert_same(to_np(dist), want_dist, "%s: distance panorama" % tag)
    assert_same(to_np(rgb), want_rgb, "%s: rgb panorama" % tag)


@pytest.mark.parametrize("name", ["c0", "c1"])
def test_workload_bit_exact(gpu, name):
    wl = synth.WORKLOADS[name]
    say("full size %s: %s" % (name, wl))
    calibs, kw, tm, ts, _ = synth.workload_frame(wl, gpu)
    est = RGBD_Estimator(calibs, device=gpu, **kw)
    rgb, dist = est.estimate_RGBD_panorama(tm, ts)
    torch.cuda.synchronize()
    ora, keep, want_rgb, want_dist = oracle_frame(calibs, kw, tm, ts)
    compare_frame(est, ora, keep, want_rgb, want_dist, rgb, dist, name)
    assert np.isfinite(want_dist).mean() > 0.8


@pytest.fixture(scope="module")
def c2(gpu):
    calibs, kw, tm, ts, _ = synth.workload_frame(synth.WORKLOADS["c2"], gpu)
    return calibs, kw, tm, ts


@pytest.mark.parametrize("volume_dtype", ["float32", "float16", "bfloat16"])
def test_c2_precision_forms_bit_exact(c2, gpu, volume_dtype):
    say("full size c2 (1024^2, 64 candidates, 4 references), %s volume" % volume_dtype)
    calibs, kw, tm, ts = c2
    est = RGBD_Estimator(calibs, device=gpu, volume_dtype=volume_dtype, **kw)
    rgb, dist = est.estimate_RGBD_panorama(tm, ts)
    torch.cuda.synchronize()
    ora, keep, want_rgb, want_dist = oracle_frame(calibs, kw, tm, ts, volume_dtype=volume_dtype)
    compare_frame(est, ora, keep, want_rgb, want_dist, rgb, dist, "c2 " + volume_dtype)


def test_c2_precision_form_error_against_fp64(c2, gpu):
    """The precision form's acceptance: inverse-distance MAE of the fp16 and
    bf16 volume forms against the fp64 oracle of the same method within
    0.002 (1/m) of the fp32 form's, at c2 (errors as utils.py:233-241 forms
    them, over the pixels where the fp64 panorama is finite)."""
    from sphere_stereo.evaluation import error_stats, inverse_distance

    calibs, kw, tm, ts = c2
    say("c2: fp64 oracle frame")
    _, _, _, dist64 = oracle_frame(calibs, kw, tm, ts, precision="f64")
    inv64 = torch.from_numpy((1.0 / dist64).astype(np.float32)).to(gpu)
    stats = {}
    for dt in ("float32", "float16", "bfloat16"):
        est = RGBD_Estimator(calibs, device=gpu, volume_dtype=dt, **kw)
        _, dist = est.estimate_RGBD_panorama(tm, ts)
        stats[dt] = error_stats(inverse_distance(dist.clone()), inv64, (0.1, 0.4), finite_only=True)
        say("c2 %s: MAE %.3e RMSE %.3e bad px %s" % (dt, stats[dt]["mae"], stats[dt]["rmse"],
                                                     stats[dt]["bad_px_ratios"]))
    assert stats["float32"]["skipped"] < 0.2 * dist64.size
    assert stats["float32"]["mae"] < 0.005
    for dt in ("float16", "bfloat16"):
        assert stats[dt]["mae"] <= stats["float32"]["mae"] + 0.002, (dt, stats[dt]["mae"], stats["float32"]["mae"])


@pytest.fixture(scope="module")
def c3(gpu):
    """The 16 ragged rigs, their fp16 single-frame oracle runs, and the
    estimators (fp16 volume)."""
    say("c3: rendering 16 ragged rigs")
    rigs = synth.ragged_rigs(gpu)
    say("c3: %s" % [s for (_, _, _, _, s) in rigs])
    ests, frames, want = [], [], []
    for i, (calibs, kw, tm, ts, shape) in enumerate(rigs):
        ests.append(RGBD_Estimator(calibs, device=gpu, volume_dtype="float16", **kw))
        frames.append((tm, ts))
        want.append(oracle_frame(calibs, kw, tm, ts, volume_dtype="float16"))
        say("c3: oracle frame %d of %d (%d^2, N %d, R %d)" % (i + 1, len(rigs), *shape))
    return ests, frames, want


def _check_batch(c3, batch, tag, graph_replays=1):
    ests, frames, want = c3
    for mode in ["eager"] + ["graph"] * graph_replays:
        for e in ests:  # stale outputs cannot pass
            e.fishey_stitcher.distance_panorama.fill_(-1.0)
            e.fishey_stitcher.RGB_panorama.fill_(7)
        batch.use_graph = mode == "graph"
        out = batch.estimate_RGBD_panoramas(frames)
        torch.cuda.synchronize()
        for i, ((rgb, dist), e, (ora, keep, wr, wd)) in enumerate(zip(out, ests, want)):
            compare_frame(e, ora, keep, wr, wd, rgb, dist, "%s %s frame %d" % (tag, mode, i))
    batch.use_graph = False


def test_c3_grouped_batch_against_oracle(c3):
    say("c3: grouped batching form")
    batch = FrameBatch(c3[0])
    assert len(batch.groups) > 4
    _check_batch(c3, batch, "c3 grouped")


@pytest.mark.parametrize("schedule", ["persistent", "dispatched"])
def test_c3_single_executor_against_oracle(c3, schedule):
    """One executor for all 16 rigs.  In the persistent schedule every
    workgroup of the inpainting pass runs many 48 x 48 tiles in a row (the
    maps hold thousands of tiles, the grid a few hundred workgroups)."""
    say("c3: single executor, %s schedule" % schedule)
    batch = FrameBatch(c3[0], single_executor=True, schedule=schedule)
    _check_batch(c3, batch, "c3 single executor " + schedule)


def test_demo_dataset_loop(gpu, tmp_path):
    """README.md:27-30: python main.py --dataset_path resources
    --references_indices 2 3 with main.py's defaults, through this package's
    dataset loop (PIL decoding, GPU input stage and mask resize, the
    estimator, PNG / EXR writers): bit for bit the oracle's panorama on the
    same decoded frames, and within 45 dB PSNR of the panorama the reference
    wrote (rgb_0.png)."""
    from sphere_stereo.main import parse_args, read_image, run

    data = os.path.join(str(tmp_path), "demo")
    shutil.copytree(DEMO, data)
    say("demo: dataset loop on %s" % data)
    panos = run(parse_args(["--dataset_path", data, "--references_indices", "2", "3", "--device", "cuda:0"]))
    assert list(panos) == ["0.jpg"]
    rgb, inv = panos["0.jpg"]["rgb"], panos["0.jpg"]["inv_distance"]
    calibs, kw, tm, ts = demo_frame(Oracle("f32"))
    from pipeline import OraclePipeline

    want_rgb, want_dist = OraclePipeline(calibs, **kw).estimate_RGBD_panorama(tm, ts)
    assert_same(rgb, want_rgb, "demo rgb panorama")
    assert_same(inv, (np.float32(1) / want_dist).astype(np.float32), "demo inverse distance panorama")
    assert psnr(rgb, read_image(DEMO_EXPECTED)) >= 45.0
    assert np.array_equal(read_image(os.path.join(data, "output", "rgb_0.png")), rgb)
    assert os.path.getsize(os.path.join(data, "output", "inv_distance_0.exr")) > 2048 * 1024 * 4
