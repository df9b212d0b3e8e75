"""The reference's per-frame pipeline composed from the oracle stages.

TEST INFRASTRUCTURE ONLY (see oracle/oracle.py).  ``OraclePipeline`` follows
RGBD_Estimator (depth_estimation.py:39-243) and Stitcher (stitcher.py:49-220)
step by step with the CPU oracle; the 4x4 pose algebra the reference does with
torch stays torch on the CPU, in fp32 for the fp32 oracle and in fp64 for the
fp64 one.  Inputs are duck-typed calibrations (fl, principal, xi, alpha, rt,
matching_scale as the reference's Calibration) and numpy / torch arrays.
"""
import numpy as np
import torch

from oracle import Oracle


def _np(t, dtype=np.float32):
    if isinstance(t, torch.Tensor):
        t = t.detach().cpu().numpy()
    return np.ascontiguousarray(t, dtype=dtype)


def _t(x, dt):
    return torch.as_tensor(x).detach().to("cpu", dt)


def thresholds(max_dist):
    """inpaintKernel's empty / source tests (stitcher.cu:287, 311) as doubles."""
    thr_empty = float(max_dist) + 0.1
    if isinstance(max_dist, (int, np.integer)):
        thr_src = float(np.float32(max_dist) + np.float32(0.1))
    else:
        thr_src = float(max_dist) + float(np.float32(0.1))
    return thr_empty, thr_src


class OraclePipeline:
    def __init__(self, calibrations, min_dist, max_dist, candidate_count, references_indices, reprojection_viewpoint,
                 masks, matching_resolution, rgb_to_stitch_resolution, panorama_resolution, sigma_i, sigma_s,
                 precision="f32", smoothing_radius=15, inpainting_iterations=32, volume_dtype="float32"):
        self.o = Oracle(precision)
        # storage rounding of the cost volume (precision form); the fp64
        # oracle ignores it
        self.vtype = {"float32": 0, "float16": 1, "bfloat16": 2}[volume_dtype]
        dt = torch.float32 if precision == "f32" else torch.float64
        self.dt = dt
        self.calibs = calibrations
        self.min_dist, self.max_dist = min_dist, max_dist
        self.N = int(candidate_count)
        self.refs = list(references_indices)
        self.W, self.H = int(matching_resolution[0]), int(matching_resolution[1])
        self.PW, self.PH = int(panorama_resolution[0]), int(panorama_resolution[1])
        self.sigma_i, self.sigma_s = sigma_i, sigma_s
        self.iters = int(inpainting_iterations)
        self.masks = np.stack([_np(m).reshape(self.H, self.W) for m in masks])

        # camera records: fl*s, principal*s as products in the working precision
        self.cams = np.array([[*(_t(c.fl, dt) * _t(c.matching_scale, dt)).tolist(),
                               *(_t(c.principal, dt) * _t(c.matching_scale, dt)).tolist(),
                               float(c.xi), float(c.alpha)] for c in calibrations])
        self.rts = {}
        for r in range(len(calibrations)):
            ref_rt = _t(calibrations[r].rt, dt)
            self.rts[r] = np.concatenate([torch.matmul(torch.inverse(_t(c.rt, dt)), ref_rt)[:3].reshape(-1).numpy()
                                          for c in calibrations]).astype(np.float64)
        self.cands = (1 / torch.linspace(1 / min_dist, 1 / max_dist, self.N, dtype=dt)).numpy().astype(np.float64)
        d0, dN1 = self.cands[0], self.cands[-1]
        if precision == "f32":
            self.A = float(np.float32(np.float32(d0) / np.float32(dN1)) - np.float32(1))
        else:
            self.A = d0 / dN1 - 1
        self.sel = {r: self.o.select_camera(self.cams, self.rts[r], self.masks, r, min_dist, max_dist)
                    for r in self.refs}

        # stitcher set-up (stitcher.py:126-164)
        vp = torch.cat([_t(reprojection_viewpoint, dt), torch.ones([1], dtype=dt)])
        sc = [calibrations[r] for r in self.refs]
        self.intr = np.array([[*(_t(c.fl, torch.float32) * _t(c.matching_scale, torch.float32)).tolist(),
                               *(_t(c.principal, torch.float32) * _t(c.matching_scale, torch.float32)).tolist(),
                               float(np.float32(c.xi)), float(np.float32(c.alpha))] for c in sc])
        if precision == "f64":
            self.intr = np.array([[*(_t(c.fl, dt) * _t(c.matching_scale, dt)).tolist(),
                                   *(_t(c.principal, dt) * _t(c.matching_scale, dt)).tolist(),
                                   float(c.xi), float(c.alpha)] for c in sc])
        self.trans = np.array([torch.matmul(torch.inverse(_t(c.rt, dt)), vp)[:3].numpy() for c in sc], np.float64)
        self.rot = np.array([torch.inverse(_t(c.rt, dt)[:3, :3]).reshape(-1).numpy() for c in sc], np.float64)
        self.luts = [self.o.inpaint_lut(self.intr[i], self.trans[i], self.H, self.W, min_dist, max_dist)
                     for i in range(len(sc))]
        m_ref = np.stack([self.masks[r] for r in self.refs])
        sm = self.o.box_filter(m_ref, smoothing_radius, 1.0)
        self.sampling, raw = self.o.blend_lut(self.intr, self.rot, self.trans, sm, self.PH, self.PW, min_dist,
                                              max_dist)
        self.weights = self.o.normalize_weights(self.o.box_filter(raw, smoothing_radius, 0.0))
        self.thr_empty, self.thr_src = thresholds(max_dist)

    def estimate_fisheye_distance(self, images_to_match, r, keep=None, sel=None):
        """depth_estimation.py:137-203 for camera r; ``sel`` (H x W camera
        indices) replaces the selection made at set-up, as the reference's
        ``selected_camera`` argument does."""
        imgs = np.stack([_np(i) for i in images_to_match])
        guide = self.o.guide(imgs[r])
        if sel is None:
            sel = self.sel[r]
        else:  # a value that is no camera index matches no camera (depth_estimation.py:155-170)
            sel = np.asarray(sel).reshape(self.H, self.W)
            sel = np.ascontiguousarray(np.where((sel >= 0) & (sel < len(self.calibs)), sel, -1), np.int32)
        cost = self.o.sweep(imgs, sel, self.cands, self.cams, self.rts[r], r, self.vtype)
        if keep is not None:
            keep["cost_raw"] = cost.copy()
            keep["guide"] = guide.copy()
        # the last up pass feeds the argmin without a stored (rounded) level 0
        cost, _ = self.o.isb_filter(guide, cost, self.sigma_i, self.sigma_s, self.vtype, round_level0=False)
        dist, idx = self.o.argmin_refine(cost, self.cands[0], self.cands[-1], self.A)
        if keep is not None:
            keep["cost_filtered"] = cost
            keep["idx"] = idx
            keep["dist_raw"] = dist.copy()
        dist, _ = self.o.isb_filter(guide, dist[None], self.sigma_i / 2, self.sigma_s / 2)
        return dist[0]

    def stitch(self, images_u8, distance_maps, keep=None):
        reproj = []
        for i in range(len(self.refs)):
            rep = self.o.reproject_zmin(distance_maps[i], self.intr[i], self.trans[i])
            if keep is not None:
                keep.setdefault("reproj_raw", []).append(rep.copy())
            reproj.append(self.o.inpaint(rep, self.luts[i], self.iters, self.thr_empty, self.thr_src))
        reproj = np.stack(reproj)
        dists = np.stack([np.asarray(d, self.o.real) for d in distance_maps])
        imgs = np.stack([_np(i, np.uint8) for i in images_u8])
        if keep is not None:
            keep["reproj"] = reproj
        dist, rgb = self.o.merge(self.sampling, self.weights, reproj, dists, imgs, self.trans, self.intr)
        return rgb, dist

    def estimate_RGBD_panorama(self, images_to_match, images_to_stitch, keep=None):
        dmaps = []
        for r in self.refs:
            k = {} if keep is not None else None
            dmaps.append(self.estimate_fisheye_distance(images_to_match, r, keep=k))
            if keep is not None:
                keep.setdefault("refs", []).append(k)
        u8 = [_np(i).astype(np.uint8) for i in images_to_stitch]  # truncation, depth_estimation.py:239-240
        if keep is not None:
            keep["dmaps"] = dmaps
        return self.stitch(u8, dmaps, keep=keep)
