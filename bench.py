#!/usr/bin/env python3
"""RGB-D panoramas/s of the sphere-sweeping stereo per-frame path on one MI355X.

Workload (default, "c1"): four seeded synthetic 1216x1216 double-sphere
fisheyes (uint8 RGB, ray-cast on the GPU), matched at 1024x1024, references 0
and 2, 32 candidates uniform in inverse distance over [1/100, 1/0.55],
sigma_i 10, sigma_s 25, 1216x1216 stitch input, 2048x1024 RGB + float32
distance panorama, fp32 (the reference's type).  One step = one
estimate_RGBD_panorama call on inputs already resident in HBM.  The same
frames are then timed with the cost volume stored in fp16 and bf16
("precision_forms"), with HIP-graph replay ("graph_replay"), and in the
batching form: --batch frames (default 8) from as many seeded rigs of the
same configuration, fp32 and fp16 volumes ("batching_form"), and a ragged
batch of 16 rigs of mixed resolutions, candidate counts and reference sets
against the same frames run one at a time ("batching_form_ragged").  Last, the
differentiation form (4 frames of one rig at 768^2, 48 candidates): forward
and forward + backward time per batch and a 50-step Gauss-Newton calibration
refinement ("differentiation_form"; --no-diff-form skips it).
--diff-report checks its gradient against fp64 central differences instead;
--issue-rates prints the measured per-instruction-class issue costs instead.

Prints one JSON line (see the roofline / cpu_baseline fields) and a per-stage
split on stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sphere-stereo_amd"))

HBM_PEAK_GBS = 8000.0

from sphere_stereo.synth import WORKLOADS  # noqa: E402
METRIC = "RGB-D panoramas/s (4×1024² fisheye, 32 cands) + % HBM peak; FP16 vs FP32"

def levels(H, W):
    return int(min(math.log2(W), math.log2(H)) - 1)


def algorithmic_bytes(stage, wl, C=4):
    """Algorithmic HBM bytes per frame of a stage: each stage reads its inputs
    once and writes its outputs once, in the reference's stage order (fp32
    volume e = 4 B, guide g = 3 B)."""
    W, H = wl["matching"]
    N, R = wl["N"], len(wl["refs"])
    e, g = 4, 3
    L = levels(H, W)
    P = [math.ceil(H / 2 ** s) * math.ceil(W / 2 ** s) for s in range(L)]
    if stage == "sweep":
        # the volume, images and guide, plus the stitch images' uint8 cast,
        # which runs as trailing workgroups of the frame's sweep launch
        sc, sr = wl["stitch"]
        return R * (N * P[0] * e + P[0] * (12 * C + 1)) + R * sc * sr * 3 * 5
    if stage == "isb_down":
        return R * sum((P[s - 1] + P[s]) * (g + N * e) for s in range(1, L))
    if stage == "isb_up":
        return R * sum((P[s + 1] + 2 * P[s]) * (g + N * e) for s in range(1, L - 1))
    if stage == "isb_up_argmin":
        return R * ((P[1] + 2 * P[0]) * (g + N * e) + N * P[0] * e + 4 * P[0])
    if stage == "dist_down":
        return R * sum((P[s - 1] + P[s]) * (g + 4) for s in range(1, L))
    if stage == "dist_up":
        # its level-0 pass also scatters the reprojection (4 B per pixel)
        return R * sum((P[s + 1] + 2 * P[s]) * (g + 4) for s in range(0, L - 1)) + R * 4 * P[0]
    if stage == "reproject":
        return R * (12 * P[0])  # read the map, scatter, plus the fill
    if stage == "inpaint":
        return R * 32 * 10 * P[0]
    if stage == "merge":
        PW, PH = wl["pano"]
        sc, sr = wl["stitch"]
        return PW * PH * (12 * R + 7) + R * (8 * P[0] + 3 * sc * sr)
    if stage == "guide":
        return R * P[0] * (12 + 4)
    if stage == "to_u8":
        sc, sr = wl["stitch"]
        return R * sc * sr * 3 * 5
    return 0


def setup(wl, device, frames_cache=None, build_estimator=True):
    import torch

    from sphere_stereo import RGBD_Estimator, synth

    rendered = None
    if frames_cache and os.path.exists(frames_cache):
        saved = torch.load(frames_cache, weights_only=True)
        rendered = (saved["images"].to(device), saved["masks"].to(device))
    calibs, kw, to_match, to_stitch, rendered = synth.workload_frame(wl, device, rendered)
    if frames_cache and not os.path.exists(frames_cache):
        torch.save({"images": rendered[0].cpu(), "masks": rendered[1].cpu()}, frames_cache)
    print("[bench] rig rendered", file=sys.stderr, flush=True)
    est = RGBD_Estimator(calibs, device=device, **kw) if build_estimator else None
    return est, to_match, to_stitch, calibs, kw


def batching_form(args, wl, device, plain_fp32):
    """The batching form: args.batch frames of the workload's configuration,
    each from its own seeded rig (seed, seed + 1, ...), through FrameBatch
    (every pass after the sweep one launch per chunk of frames, the
    variants picked by the autotuner first), timed like the plain form, on one
    stream and on two lanes (value); fp32 and fp16 volumes.  value =
    panoramas/s over all frames of the batch."""
    import torch

    from sphere_stereo import FrameBatch, RGBD_Estimator

    sets = []
    for b in range(args.batch):
        wb = dict(wl, seed=wl["seed"] + b)
        cache = "%s.rig%d" % (args.frames_cache, b) if args.frames_cache and b else args.frames_cache
        _, tm, ts, calibs, kw = setup(wb, device, cache, build_estimator=False)
        sets.append(([t.contiguous() for t in tm], [t.contiguous() for t in ts], calibs, kw))
    out = {"frames_per_batch": args.batch, "rigs": "seeds %d..%d" % (wl["seed"], wl["seed"] + args.batch - 1),
           "lanes": 2}
    for dt in ("float32", "float16"):
        ests = [RGBD_Estimator(c, device=device, volume_dtype=dt, **kw) for (_, _, c, kw) in sets]
        batch = FrameBatch(ests)
        frames = [(tm, ts) for (tm, ts, _, _) in sets]
        tuned = batch.autotune(frames)  # the sweep autotuner, per configuration

        def timed(n):
            for _ in range(args.warmup):
                batch.estimate_RGBD_panoramas(frames)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(n):
                batch.estimate_RGBD_panoramas(frames)
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        # one stream, then two lanes (chunks of 4 frames: the second chunk's
        # sweeps beside the first chunk's filters and stitching)
        v1 = args.steps * args.batch / timed(args.steps)
        batch.lanes = 2
        el = timed(args.steps)
        v = args.steps * args.batch / el
        # the same batches replayed from a HIP graph (captured once)
        batch.use_graph = True
        for _ in range(2):
            batch.estimate_RGBD_panoramas(frames)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.estimate_RGBD_panoramas(frames)
        torch.cuda.synchronize()
        gv = args.steps * args.batch / (time.perf_counter() - t0)
        out[dt] = {"value": round(v, 3), "unit": "panoramas/s", "ms_per_batch": round(1e3 * el / args.steps, 4),
                   "vs_plain_fp32": round(v / plain_fp32, 4), "autotune": tuned[0],
                   "one_lane": {"value": round(v1, 3), "unit": "panoramas/s", "vs_plain_fp32": round(v1 / plain_fp32, 4)},
                   "graph_replay": {"value": round(gv, 3), "unit": "panoramas/s",
                                    "vs_plain_fp32": round(gv / plain_fp32, 4)}}
        print("[bench] batching form, %d frames, %s volume: %.1f panoramas/s on 2 lanes (one lane %.1f, graph replay "
              "%.1f)" % (args.batch, dt, v, v1, gv), file=sys.stderr, flush=True)
        del batch, ests
        torch.cuda.empty_cache()
    return out


def sweep_filter_summary(wl, per_frame):
    """Sweep + cost filter as one stage against the HBM peak: their
    algorithmic bytes per frame (the sweep; every down pass; every up pass
    including level 0's filter part, without the argmin the last pass fuses)
    over the summed per-frame time of the sweep, down, up and fused
    up + argmin launches (the fused pass counted whole)."""
    W, H = wl["matching"]
    N, R = wl["N"], len(wl["refs"])
    L = levels(H, W)
    P = [math.ceil(H / 2 ** s) * math.ceil(W / 2 ** s) for s in range(L)]
    up0 = R * (P[1] + 2 * P[0]) * (3 + N * 4) if L > 1 else 0
    b = algorithmic_bytes("sweep", wl) + algorithmic_bytes("isb_down", wl) + algorithmic_bytes("isb_up", wl) + up0
    ms = sum(per_frame.get(k, 0.0) for k in ("sweep", "isb_down", "isb_up", "isb_up_argmin"))
    if ms <= 0:
        return None
    gbs = b / (ms * 1e-3) / 1e9
    return {"algorithmic_bytes": int(b), "ms_per_frame": round(ms, 4), "achieved_gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def batching_ragged(args, device):
    """The ragged batching form: 16 rigs, each a seeded scene with its own
    baseline drawn from U(0.15, 0.40) m, matching resolution from {512, 768,
    1024}^2, candidate count from {16, 32, 48, 64} and 2 to 4 reference
    cameras (1216^2 stitch input, 1024x512 panorama), fp16 volume: one
    FrameBatch grouped by configuration (each group one launch per pass and
    chunk), and the single-executor form (one launch per pass for all rigs)
    in its persistent and dispatched schedules, timed against the same 16
    frames run one at a time in the plain fp32 form."""
    import torch

    from sphere_stereo import FrameBatch, RGBD_Estimator, synth

    sets = [(tm, ts, calibs, kw, shape) for (calibs, kw, tm, ts, shape) in synth.ragged_rigs(device)]
    print("[bench] ragged batch rendered: %s" % [c for (_, _, _, _, c) in sets], file=sys.stderr, flush=True)
    frames = [(tm, ts) for (tm, ts, _, _, _) in sets]
    plain = [RGBD_Estimator(c, device=device, **kw) for (_, _, c, kw, _) in sets]
    for e, (tm, _) in zip(plain, frames):  # tuned like the batched form below
        e.autotune(tm)
    for _ in range(args.warmup):
        for e, (tm, ts) in zip(plain, frames):
            e.estimate_RGBD_panorama(tm, ts)
    torch.cuda.synchronize()
    reps = max(1, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        for e, (tm, ts) in zip(plain, frames):
            e.estimate_RGBD_panorama(tm, ts)
    torch.cuda.synchronize()
    v_plain = reps * len(frames) / (time.perf_counter() - t0)
    del plain
    torch.cuda.empty_cache()
    ests = [RGBD_Estimator(c, device=device, volume_dtype="float16", **kw) for (_, _, c, kw, _) in sets]
    batch = FrameBatch(ests)
    tuned = batch.autotune(frames)
    out = {"rigs": len(sets), "configurations": len(batch.groups),
           "sweep_variants": [t["variant"] for t in tuned],
           "filter_variants": [t.get("filter_variant") for t in tuned],
           "inpaint_variants": [t.get("inpaint_variant") for t in tuned],
           "shapes": ["%d^2 N%d R%d" % c for (_, _, _, _, c) in sets],
           "plain_fp32_sequential": {"value": round(v_plain, 3), "unit": "panoramas/s"}}
    for mode in ("eager", "graph"):
        batch.use_graph = mode == "graph"
        for _ in range(args.warmup):
            batch.estimate_RGBD_panoramas(frames)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            batch.estimate_RGBD_panoramas(frames)
        torch.cuda.synchronize()
        v = reps * len(frames) / (time.perf_counter() - t0)
        out["float16_batched_" + mode] = {"value": round(v, 3), "unit": "panoramas/s",
                                          "vs_plain_fp32": round(v / v_plain, 4)}
        print("[bench] ragged batch (%d rigs, %d configurations), fp16 %s: %.1f panoramas/s (plain fp32 one at a "
              "time: %.1f)" % (len(sets), len(batch.groups), mode, v, v_plain), file=sys.stderr, flush=True)
    del batch
    # the single-executor ragged form: one launch per pass (per candidate-lane count for the
    # filter levels) for all 16 rigs, in its two schedules
    for schedule in ("persistent", "dispatched"):
        batch = FrameBatch(ests, single_executor=True, schedule=schedule)
        batch.autotune(frames)  # the variants tuned above: rebuilds with one sweep launch per variant
        for mode in ("eager", "graph"):
            batch.use_graph = mode == "graph"
            for _ in range(args.warmup):
                batch.estimate_RGBD_panoramas(frames)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                batch.estimate_RGBD_panoramas(frames)
            torch.cuda.synchronize()
            v = reps * len(frames) / (time.perf_counter() - t0)
            out["float16_ragged_%s_%s" % (schedule, mode)] = {"value": round(v, 3), "unit": "panoramas/s",
                                                               "vs_plain_fp32": round(v / v_plain, 4)}
            print("[bench] ragged batch (%d rigs), fp16, one executor, %s schedule, %s: %.1f panoramas/s"
                  % (len(sets), schedule, mode, v), file=sys.stderr, flush=True)
        del batch
    del ests
    torch.cuda.empty_cache()
    return out


DIFF_FORM = {"matching": (768, 768), "N": 48, "frames": 4, "refs": (0, 2), "sigma_i": 10.0, "sigma_s": 12.5,
             "temperature": 0.05}


def diff_setup(device):
    """The differentiation form's batch: 4 frames (scene seeds 0..3) of the
    config-0 rig at 768^2 matching, rendered on the GPU."""
    from sphere_stereo import synth
    from sphere_stereo.differentiable import DifferentiableDepth
    from sphere_stereo.utils import parse_json_calib

    d = DIFF_FORM
    frames = []
    for seed in range(d["frames"]):
        srig = synth.SyntheticRig(seed=seed, resolution=1216)
        imgs, masks = srig.render(device)
        tm, _, mk = synth.make_frame(srig, d["matching"], (1216, 1216), list(d["refs"]), device=device,
                                     rendered=(imgs, masks))
        frames.append([t.contiguous() for t in tm])
        calibs = parse_json_calib(srig.basalt_json()["value0"], list(d["matching"]), device)
    rig = DifferentiableDepth(calibs, 0.55, 100, d["N"], list(d["refs"]), mk, d["matching"], d["sigma_i"],
                              d["sigma_s"], device, temperature=d["temperature"])
    return rig, frames


def differentiation_form(args, device):
    """The differentiation form: soft-argmin inverse distance of both
    references of 4 frames, forward and forward + backward (gradients to the
    images, all intrinsics and all poses; and to intrinsics and poses only)
    timed per batch; then the
    Gauss-Newton refinement of camera 1 from a perturbation of 0.5 degrees,
    5 mm and 1 % of fx over 50 steps."""
    import numpy as np
    import torch

    from sphere_stereo.differentiable import pose_exp, rotation_angle_deg

    rig, frames = diff_setup(device)
    gen = torch.Generator(device="cpu").manual_seed(7)
    wmap = torch.rand((rig.rows, rig.cols), generator=gen).to(device)
    req = [[t.clone().requires_grad_(True) for t in f] for f in frames]

    def fwd(grad, images=True):
        cams6 = rig.cams6.clone().requires_grad_(grad)
        rts = rig.rts.clone().requires_grad_(grad)
        loss = 0.0
        for imgs in (req if grad and images else frames):
            for inv in rig(imgs, cams6, rts):
                loss = loss + (inv * wmap).sum()
        if grad:
            loss.backward()
        return loss

    out = {"frames_per_batch": DIFF_FORM["frames"], "matching": list(DIFF_FORM["matching"]), "candidates": DIFF_FORM["N"],
           "references": list(DIFF_FORM["refs"]), "temperature": DIFF_FORM["temperature"], "dtype": "f32"}
    # forward; forward + backward to images, intrinsics and poses; forward +
    # backward to intrinsics and poses only (what the calibration refinement
    # needs: no image-gradient scatter)
    for name, grad, images in (("forward", False, False), ("forward_backward", True, True),
                               ("forward_backward_calibration", True, False)):
        ctx = torch.no_grad() if not grad else torch.enable_grad()
        with ctx:
            for _ in range(2):
                fwd(grad, images)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 5
            for _ in range(reps):
                fwd(grad, images)
            torch.cuda.synchronize()
            el = (time.perf_counter() - t0) / reps
        out[name] = {"ms_per_batch": round(1e3 * el, 3), "frames_per_s": round(DIFF_FORM["frames"] / el, 2)}
        if grad:
            out[name]["vs_forward"] = round(el / (out["forward"]["ms_per_batch"] * 1e-3), 3)
        print("[bench] differentiation form %s: %.2f ms per batch of %d frames" % (name, 1e3 * el,
              DIFF_FORM["frames"]), file=sys.stderr, flush=True)
    ct, rt = rig.cams6.clone(), rig.rts.clone()
    ax = torch.tensor([0.3, 0.8, -0.5], dtype=torch.float64)
    ax = ax / ax.norm() * np.radians(0.5)
    tr = torch.tensor([0.004, -0.002, 0.0022], dtype=torch.float64)
    tr = tr / tr.norm() * 0.005
    r0, c0 = rt.clone(), ct.clone()
    r0[1] = rt[1] @ pose_exp(torch.cat([ax, tr]))
    c0[1, 0] *= 1.01

    def errors(c6, rts):
        D = torch.linalg.inv(rt[1]) @ rts[1]
        return {"rotation_deg": round(rotation_angle_deg(D[:3, :3]), 4), "translation_mm": round(float(
            D[:3, 3].norm()) * 1e3, 3), "fx_rel": round(float(c6[1, 0] / ct[1, 0]) - 1, 5)}

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cams, rts, hist = rig.refine_calibration(frames, 1, steps=50, cams6=c0, rts=r0)
    el = time.perf_counter() - t0
    out["calibration_refinement"] = {"camera": 1, "steps": len(hist), "seconds": round(el, 3),
                                     "before": errors(c0, r0), "after": errors(cams, rts),
                                     "loss_first": round(hist[0]["loss"], 3), "loss_last": round(hist[-1]["loss"], 3)}
    print("[bench] calibration refinement: %s -> %s in %.2f s" % (errors(c0, r0), errors(cams, rts), el),
          file=sys.stderr, flush=True)
    return out


def diff_report(args):
    """Gradient check at the differentiation form's size: the GPU gradient of
    L = sum_p w_p / d_p (reference 0 of frame 0, soft-argmin readout) with
    respect to the camera parameters that L depends on most, against fp64
    central differences (h = 1e-4, and h = 1e-6, where fewer samples cross
    the kinks of the bilinear cells and of |.| within the step) of the
    oracle's forward (sweep, ISB filter, soft-argmin) with the same camera
    selection."""
    import numpy as np
    import torch

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    from sphere_stereo.differentiable import relative_poses
    from sphere_stereo.utils import distance_candidates

    device = torch.device("cuda:0")
    rig, frames = diff_setup(device)
    images = frames[0]
    gen = torch.Generator(device="cpu").manual_seed(7)
    wmap = torch.rand((rig.rows, rig.cols), generator=gen, dtype=torch.float64)
    cams6 = rig.cams6.clone().requires_grad_(True)
    inv = rig(images, cams6, rig.rts)[0]
    (inv.double() * wmap.to(device)).sum().backward()
    g = cams6.grad.numpy()
    o64, o32 = Oracle("f64"), Oracle("f32")
    imgs32 = np.stack([t.cpu().numpy() for t in images]).astype(np.float32)
    sel = rig._sel[0].cpu().numpy().astype(np.int32).reshape(rig.rows, rig.cols)
    rel = relative_poses(rig.rts, 0).reshape(-1, 12).numpy()
    cands = np.array(distance_candidates(0.55, 100, rig.N), np.float64)
    guide = o32.guide(imgs32[0])
    wnp = wmap.numpy()

    def L(c6):
        cost = o64.sweep(imgs32, sel, cands, c6, rel, 0)
        Fv, _ = o64.isb_filter(guide, cost, rig.sigma_i, rig.sigma_s)
        return float((o64.soft_argmin(Fv, rig.beta, rig.inv0, rig.inv1) * wnp).sum())

    rows = []
    names = ["fs0", "fs1", "cs0", "cs1", "xi", "alpha"]
    for flat in np.argsort(-np.abs(g).ravel())[:args.diff_params]:
        k, i = divmod(int(flat), 6)
        row = {"camera": k, "parameter": names[i], "gpu": round(float(g[k, i]), 5)}
        for h in (1e-4, 1e-6):
            cp, cm = rig.cams6.numpy().copy(), rig.cams6.numpy().copy()
            cp[k, i] += h
            cm[k, i] -= h
            fd = (L(cp) - L(cm)) / (2 * h)
            tag = "h%.0e" % h
            row["fd_" + tag] = round(fd, 5)
            row["rel_err_" + tag] = round(abs(float(g[k, i]) - fd) / max(abs(fd), 1e-12), 6)
        rows.append(row)
        print("[bench] d/d %s[%d]: %s" % (names[i], k, row), file=sys.stderr, flush=True)
    print(json.dumps({"differentiation_gradient_check": {"workload": "frame 0, reference 0, 768^2, 48 candidates",
                                                         "loss": "sum_p w_p / d_p, w ~ U(0,1) seed 7",
                                                         "rows": rows}}), flush=True)


def profile(lib, est, to_match, to_stitch, frames, mask):
    import ctypes as C

    import torch

    lib.ss_profile_enable(1, mask)
    for _ in range(frames):
        est.estimate_RGBD_panorama(to_match, to_stitch)
    torch.cuda.synchronize()
    lib.ss_profile_enable(0, 0)
    ms = (C.c_double * 11)()
    n = (C.c_int * 11)()
    lib.ss_profile_collect(ms, n)
    return {lib.ss_stage_name(i).decode(): (ms[i], n[i]) for i in range(11)}


def cpu_baseline(wl, calibs, kw, to_match, to_stitch, budget_s):
    """The fp32 oracle (a C port of the reference path, its per-pixel loops
    on OpenMP threads) on the host: full frames of the same workload, timed
    after their set-up."""
    import torch

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pipeline import OraclePipeline

    cpu_cal = []
    for c in calibs:
        cc = type(c)(c.original_resolution, c.principal.cpu(), c.fl.cpu(), c.xi, c.alpha, c.rt.cpu(),
                     c.matching_scale.cpu())
        cpu_cal.append(cc)
    kwc = dict(kw)
    kwc["reprojection_viewpoint"] = kw["reprojection_viewpoint"].cpu()
    kwc["masks"] = [m.cpu() for m in kw["masks"]]
    ora = OraclePipeline(cpu_cal, **kwc)
    tm = [t.cpu() for t in to_match]
    ts = [t.cpu() for t in to_stitch]
    frames, t0 = 0, time.perf_counter()
    while True:
        ora.estimate_RGBD_panorama(tm, ts)
        frames += 1
        el = time.perf_counter() - t0
        if el > budget_s or frames >= 40:
            break
    import ctypes

    threads = ctypes.CDLL("libgomp.so.1").omp_get_max_threads()
    return dict(value=frames / el, unit="panoramas/s", cores=int(threads), kind="port",
                sample="%d full %s frame(s) (%dx%d matching, %d candidates, %d refs, %dx%d panorama) through the "
                       "fp32 oracle (C, per-pixel loops on %d OpenMP threads; the z-min reprojection serial), "
                       "set-up excluded; %.1f s" %
                       (frames, "c1" if wl["matching"][0] == 1024 else "c0", wl["matching"][0], wl["matching"][1],
                        wl["N"], len(wl["refs"]), wl["pano"][0], wl["pano"][1], threads, el))


def precision_report(args):
    """Inverse-distance error of the fp32 / fp16 / bf16 volume forms against the
    fp64 oracle of the same method on one frame of the workload (MAE, RMSE,
    bad-pixel ratios at 0.1 and 0.4 as utils.py:233-241 computes them), plus
    each form's panoramas/s.  Pixels where the fp64 panorama is not finite (the
    reference's own 0/0 blend weights) are left out and counted."""
    import torch

    from sphere_stereo import RGBD_Estimator
    from sphere_stereo.evaluation import error_stats, inverse_distance

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pipeline import OraclePipeline

    wl = WORKLOADS[args.workload]
    device = torch.device("cuda:0")
    print("[precision] setting up %s" % args.workload, file=sys.stderr, flush=True)
    _, to_match, to_stitch, calibs, kw = setup(wl, device, args.frames_cache, build_estimator=False)
    forms = {}
    for dt in ("float32", "float16", "bfloat16"):
        est = RGBD_Estimator(calibs, device=device, volume_dtype=dt, **kw)
        for _ in range(args.warmup):
            est.estimate_RGBD_panorama(to_match, to_stitch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            est.estimate_RGBD_panorama(to_match, to_stitch)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        _, dist = est.estimate_RGBD_panorama(to_match, to_stitch)
        forms[dt] = {"inv": inverse_distance(dist.clone()), "panoramas_per_s": round(args.steps / el, 2)}
        print("[precision] %s: %.1f panoramas/s" % (dt, args.steps / el), file=sys.stderr, flush=True)
        del est
    # the fp64 oracle on the host, same inputs
    cpu_cal = [type(c)(c.original_resolution, c.principal.cpu(), c.fl.cpu(), c.xi, c.alpha, c.rt.cpu(),
                       c.matching_scale.cpu()) for c in calibs]
    kwc = dict(kw)
    kwc["reprojection_viewpoint"] = kw["reprojection_viewpoint"].cpu()
    kwc["masks"] = [m.cpu() for m in kw["masks"]]
    t0 = time.perf_counter()
    ora = OraclePipeline(cpu_cal, precision="f64", **kwc)
    print("[precision] fp64 oracle set up (%.1f s)" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    _, dist64 = ora.estimate_RGBD_panorama([t.cpu() for t in to_match], [t.cpu() for t in to_stitch])
    t_ora = time.perf_counter() - t0
    print("[precision] fp64 oracle frame done (%.1f s)" % t_ora, file=sys.stderr, flush=True)
    inv64 = torch.from_numpy((1.0 / dist64).astype("float32")).to(device)
    out = {}
    for dt, f in forms.items():
        st = error_stats(f["inv"], inv64, (0.1, 0.4), finite_only=True)
        out[dt] = {"mae": st["mae"], "rmse": st["rmse"], "bad_px_ratio_0.1": st["bad_px_ratios"][0],
                   "bad_px_ratio_0.4": st["bad_px_ratios"][1], "pixels": st["count"], "skipped": st["skipped"],
                   "panoramas_per_s": f["panoramas_per_s"]}
    line = {"precision_report": args.workload, "metric": "inverse-distance error vs the fp64 oracle (1/m)",
            "config": {"matching": list(wl["matching"]), "candidates": wl["N"], "refs": list(wl["refs"]),
                       "panorama": list(wl["pano"]), "sigma_i": wl["sigma_i"], "sigma_s": wl["sigma_s"]},
            "forms": out, "fp64_oracle_seconds": round(t_ora, 1),
            "mae_fp16_minus_fp32": out["float16"]["mae"] - out["float32"]["mae"],
            "mae_bf16_minus_fp32": out["bfloat16"]["mae"] - out["float32"]["mae"],
            "fp16_speedup": round(out["float16"]["panoramas_per_s"] / out["float32"]["panoramas_per_s"], 4)}
    print(json.dumps(line), flush=True)


def issue_rates():
    """Runs the issue-rate probe (built by __graft_entry__.build() from
    sphere-stereo_amd/csrc/probes/issue_rates.hip) as a child process and
    passes its markdown tables through."""
    import subprocess
    exe = os.path.join(ROOT, "sphere-stereo_amd", "sphere_stereo", "issue_rates")
    if not os.path.exists(exe):
        sys.exit("the issue-rate probe is not built (%s); run __graft_entry__.build()" % exe)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    sys.stdout.write(res.stdout)
    if res.returncode != 0:
        sys.stderr.write(res.stderr)
        sys.exit(res.returncode)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--frames-cache", default=None, help="reuse rendered fisheyes saved at this path")
    ap.add_argument("--tuning", default=None, help="autotuner results of the plain and precision forms: loaded from "
                    "this JSON file when it exists (no tuning run), else tuned and saved there (counter runs, which "
                    "serialise kernels, then time the variants of a normal run)")
    ap.add_argument("--batch", type=int, default=8,
                    help="frames per batch of the batching form (0: skip it); each from its own seeded rig")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged batching form (16 rigs of mixed "
                    "resolutions, candidate counts and reference sets)")
    ap.add_argument("--no-diff-form", action="store_true", help="skip the differentiation form's timing and "
                    "calibration refinement")
    ap.add_argument("--precision-report", action="store_true",
                    help="instead of the throughput line: inverse-distance error of the fp32/fp16/bf16 volume forms "
                         "against the fp64 oracle (use --workload c2)")
    ap.add_argument("--diff-report", action="store_true",
                    help="instead of the throughput line: the differentiation form's GPU gradient against fp64 "
                         "central differences (h = 1e-4) of the oracle at full size")
    ap.add_argument("--diff-params", type=int, default=6, help="parameters checked by --diff-report")
    ap.add_argument("--issue-rates", action="store_true",
                    help="instead of the throughput line: the issue-rate probe's table (SIMD cycles per wave64 "
                         "instruction by class, LDS atomic throughput) that the kernels' rooflines are priced with")
    args = ap.parse_args()
    if args.issue_rates:
        issue_rates()
        return
    if args.diff_report:
        diff_report(args)
        return
    if args.precision_report:
        precision_report(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would run its own frames "
              "or references independently); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from sphere_stereo import _lib

    lib = _lib.lib
    wl = WORKLOADS[args.workload]
    device = torch.device("cuda:0")
    print("[bench] setting up %s" % args.workload, file=sys.stderr, flush=True)
    est, to_match, to_stitch, calibs, kw = setup(wl, device, args.frames_cache)
    # the autotuner's variants (sweep, cost filter, inpainting tile: identical
    # bits) for this configuration, fp32 and the 16-bit volume types
    from sphere_stereo.depth_estimation import load_sweep_tuning, save_sweep_tuning

    tuning_loaded = bool(args.tuning) and os.path.exists(args.tuning)
    if tuning_loaded:
        tuned_plain = {"loaded": args.tuning, "records": load_sweep_tuning(args.tuning)}
    else:
        tuned_plain = est.autotune(to_match)
    for _ in range(args.warmup):
        est.estimate_RGBD_panorama(to_match, to_stitch)
    torch.cuda.synchronize()

    # untimed pass: per-stage split, to find the dominant kernel
    split = profile(lib, est, to_match, to_stitch, 2, (1 << 11) - 1)
    per_frame = {k: v[0] / 2 for k, v in split.items()}
    dominant = max(per_frame, key=per_frame.get)
    dom_bit = [lib.ss_stage_name(i).decode() for i in range(11)].index(dominant)
    for k, (ms, n) in sorted(split.items(), key=lambda kv: -kv[1][0]):
        if n:
            b = algorithmic_bytes(k, wl)
            print("[bench] %-14s %8.3f ms/frame %4d launches/frame  %8.1f GB/s algorithmic" %
                  (k, ms / 2, n // 2, b / (ms / 2 * 1e-3) / 1e9 if ms else 0), file=sys.stderr)

    # timed region: only the dominant stage is bracketed by events
    lib.ss_profile_enable(1, 1 << dom_bit)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        est.estimate_RGBD_panorama(to_match, to_stitch)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib.ss_profile_enable(0, 0)
    import ctypes as C

    ms = (C.c_double * 11)()
    n = (C.c_int * 11)()
    lib.ss_profile_collect(ms, n)
    elapsed = t1 - t0
    value = args.steps / elapsed
    launches_per_frame = n[dom_bit] / args.steps
    avg_launch_ms = ms[dom_bit] / max(n[dom_bit], 1)
    bytes_per_launch = algorithmic_bytes(dominant, wl) / launches_per_frame
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9

    # the same frames replayed from a HIP graph captured once (same kernels,
    # same inputs; only the per-launch host work goes away)
    est.use_graph = True
    for _ in range(2):
        est.estimate_RGBD_panorama(to_match, to_stitch)
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    for _ in range(args.steps):
        est.estimate_RGBD_panorama(to_match, to_stitch)
    torch.cuda.synchronize()
    g_elapsed = time.perf_counter() - g0
    est.use_graph = False

    # precision forms: the same frames with the cost volume stored in fp16 /
    # bf16 (fp32 arithmetic), timed the same way
    from sphere_stereo import RGBD_Estimator

    forms = {}
    for dt in ("float16", "bfloat16"):
        est_h = RGBD_Estimator(calibs, device=device, volume_dtype=dt, **kw)
        tuned_h = {"loaded": args.tuning} if tuning_loaded else est_h.autotune(to_match)
        for _ in range(args.warmup):
            est_h.estimate_RGBD_panorama(to_match, to_stitch)
        torch.cuda.synchronize()
        split_h = profile(lib, est_h, to_match, to_stitch, 2, (1 << 11) - 1)
        print("[bench] %s stages (ms/frame): %s" % (dt, ", ".join("%s %.3f" % (k, v[0] / 2) for k, v in sorted(
            split_h.items(), key=lambda kv: -kv[1][0]) if v[1])), file=sys.stderr, flush=True)
        h0 = time.perf_counter()
        for _ in range(args.steps):
            est_h.estimate_RGBD_panorama(to_match, to_stitch)
        torch.cuda.synchronize()
        h_el = time.perf_counter() - h0
        forms[dt] = {"value": round(args.steps / h_el, 3), "unit": "panoramas/s",
                     "ms_per_step": round(1e3 * h_el / args.steps, 4),
                     "vs_fp32": round((args.steps / h_el) / value, 4), "autotune": tuned_h}
        print("[bench] %s volume: %.1f panoramas/s" % (dt, args.steps / h_el), file=sys.stderr, flush=True)
        del est_h

    if args.tuning and not tuning_loaded:
        save_sweep_tuning(args.tuning)
    batching = batching_form(args, wl, device, value) if args.batch > 0 else None
    ragged = batching_ragged(args, device) if not args.no_ragged else None
    diff_form = differentiation_form(args, device) if not args.no_diff_form else None

    # HBM traffic and VALU counters of the dominant kernel: rocprofv3 counter
    # runs (--pmc) cannot share a run with this timing, so they are read from
    # the committed profiles of an earlier counter run of this workload
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dominant)
        if traffic is not None:
            traffic = dict(traffic, source="profiles/traffic_%s.json (rocprofv3 FETCH_SIZE / WRITE_SIZE runs, not "
                                           "this run)" % args.workload)
    # the dominant kernel's VALU counters (rocprofv3 SQ_INSTS_VALU by class,
    # VALUBusy, profiles/valu_<workload>.json) and its VALU issue time priced
    # at measured per-class rates: the sweep is bound neither by HBM nor by
    # VALU issue (~2/3 of its time), but by its bilinear gathers
    valu = None
    vpath = os.path.join(ROOT, "profiles", "valu_%s.json" % args.workload)
    if os.path.exists(vpath):
        with open(vpath) as f:
            valu = json.load(f).get(dominant)
        if valu is not None:
            valu = dict(valu, source="profiles/valu_%s.json (rocprofv3 SQ_INSTS_VALU / VALUBusy runs, not this run)"
                        % args.workload)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "panoramas/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ray-cast double-sphere rig, 4 x 1216^2 uint8 fisheyes)",
        "config": {"workload": args.workload + ": 4 fisheyes 1216^2 -> %dx%d matching, refs %s, %d candidates "
                                                "[0.55, 100] m, sigma_i %g, sigma_s %g, %dx%d stitch, %dx%d RGB-D "
                                                "panorama" % (wl["matching"][0], wl["matching"][1],
                                                              list(wl["refs"]), wl["N"], wl["sigma_i"],
                                                              wl["sigma_s"], wl["stitch"][0], wl["stitch"][1],
                                                              wl["pano"][0], wl["pano"][1]),
                   "global_batch": 1, "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_ms": round(avg_launch_ms, 5), "algorithmic_bytes_per_launch": int(bytes_per_launch),
                     "valu": valu},
        "autotune": tuned_plain,
        "stage_ms_per_frame": {k: round(v, 4) for k, v in per_frame.items() if v > 0},
        "sweep_filter": sweep_filter_summary(wl, per_frame),
        "graph_replay": {"value": round(args.steps / g_elapsed, 3), "unit": "panoramas/s",
                         "ms_per_step": round(1e3 * g_elapsed / args.steps, 4)},
        "precision_forms": forms,
        "batching_form": batching,
        "batching_form_ragged": ragged,
        "differentiation_form": diff_form,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(wl, calibs, kw, to_match, to_stitch, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
