"""RGBD_Estimator with the reference's interface (python/depth_estimation.py).

The estimator keeps the reference's constructor, ``select_camera``,
``estimate_fisheye_distance`` and ``estimate_RGBD_panorama``.  Each reference
camera's chain - guide, sweep, cost filter with the final up-pass fused with
the argmin, distance filter - is one ss_estimate_distance call that enqueues
the gfx950 kernels on torch's current stream; the panorama is produced by the
Stitcher.  Init-time work (camera selection, unit-ray tables, stitching
tables) also runs on the GPU.
"""
import torch

from . import _lib
from .isb_filter import ISB_Filter
from .stitcher import Stitcher
from .utils import camera_array, distance_candidates, relative_transforms


def _hwc(image):
    """[1,3,1,H,W] permuted views (depth_estimation.py:222-223) back to their
    contiguous HWC storage; HWC tensors pass through."""
    if image.dim() == 5:
        image = image[0, :, 0].permute(1, 2, 0)
    return image.contiguous()


class RGBD_Estimator:
    def __init__(self, calibrations, min_dist, max_dist, candidate_count, references_indices, reprojection_viewpoint,
                 masks, matching_resolution, rgb_to_stitch_resolution, panorama_resolution, sigma_i, sigma_s,
                 device):
        """RGB-D estimation set-up (depth_estimation.py:39-81): camera selection
        for adaptive matching, filters and stitcher.  Arguments as the
        reference: resolutions are (cols, rows), masks are [1, rows, cols]."""
        self.calibrations = calibrations
        self.min_dist = min_dist
        self.max_dist = max_dist
        self.candidate_count = int(candidate_count)
        self.references_indices = list(references_indices)
        self.reprojection_viewpoint = reprojection_viewpoint
        self.matching_resolution = matching_resolution
        self.device = torch.device(device)
        self.sigma_i = sigma_i
        self.sigma_s = sigma_s
        self.cols, self.rows = int(matching_resolution[0]), int(matching_resolution[1])
        C = len(calibrations)
        if not 1 <= C <= _lib.SS_MAX_CAMERAS:
            raise ValueError("between 1 and %d cameras supported" % _lib.SS_MAX_CAMERAS)
        if not 1 <= self.candidate_count <= _lib.SS_MAX_CANDIDATES:
            raise ValueError("candidate_count must be in [1, %d]" % _lib.SS_MAX_CANDIDATES)

        self.cost_filter = ISB_Filter(candidate_count, matching_resolution, device)
        self.distance_filter = ISB_Filter(1, matching_resolution, device)
        calibrations_for_stitch = [calibrations[r] for r in self.references_indices]
        masks_for_stitching = [masks[r] for r in self.references_indices]
        self.fishey_stitcher = Stitcher(calibrations_for_stitch, reprojection_viewpoint, masks_for_stitching,
                                        min_dist, max_dist, matching_resolution, rgb_to_stitch_resolution,
                                        panorama_resolution, device)

        self._cams = camera_array(calibrations)
        self._rts = {r: _lib.float_array(relative_transforms(calibrations, calibrations[r]))
                     for r in range(C)}
        self._cands_list = distance_candidates(min_dist, max_dist, self.candidate_count)
        self._cands = _lib.float_array(self._cands_list)
        nbytes = _lib.lib.ss_estimate_workspace_bytes(self.rows, self.cols, self.candidate_count)
        if nbytes == 0:
            raise ValueError("unsupported matching resolution / candidate count")
        R = len(self.references_indices)
        # one workspace and one stream per reference: the references' chains
        # are independent until the merge and run concurrently
        self._workspaces = [torch.empty(nbytes, dtype=torch.uint8, device=self.device) for _ in range(R)]
        self._workspace = self._workspaces[0]
        self._streams = [torch.cuda.Stream(device=self.device) for _ in range(R)]
        self._fork = torch.cuda.Event()
        self._joins = [torch.cuda.Event() for _ in range(R)]
        self._distance_maps = torch.zeros([R, self.rows, self.cols], device=self.device)
        self._stitch_u8 = None
        self.select_camera(masks)

    def _stream(self):
        return _lib.stream_of(self._workspace)

    def select_camera(self, masks):
        """Adaptive camera selection (depth_estimation.py:83-135).  Fills
        ``selected_cameras`` ([1, rows, cols] int64, -1 = none, as the
        reference) and the per-reference unit-ray tables of the sweep."""
        stream = self._stream()
        stacked = torch.cat([m.to(self.device, torch.float32).reshape(1, self.rows, self.cols) for m in masks],
                            dim=0).contiguous()
        self.selected_cameras, self._sel8, self._raylut = [], [], []
        for r in self.references_indices:
            sel = torch.empty([self.rows, self.cols], dtype=torch.int8, device=self.device)
            _lib.check(_lib.lib.ss_select_camera(self._cams, self._rts[r], len(self.calibrations), r,
                                                 stacked.data_ptr(), self.rows, self.cols, float(self.min_dist),
                                                 float(self.max_dist), sel.data_ptr(), stream))
            lut = torch.empty([self.rows, self.cols, 4], dtype=torch.float32, device=self.device)
            _lib.check(_lib.lib.ss_ray_lut(self._cams[r], sel.data_ptr(), self.rows, self.cols, lut.data_ptr(),
                                           stream))
            self._sel8.append(sel)
            self._raylut.append(lut)
            self.selected_cameras.append(sel.to(torch.int64).unsqueeze(0))
        torch.cuda.synchronize(self.device)  # stacked is released on return

    def _reference_slot(self, reference_calibration):
        for slot, r in enumerate(self.references_indices):
            if self.calibrations[r] is reference_calibration:
                return slot, r
        raise ValueError("reference_calibration is not one of the estimator's references")

    def _estimate(self, slot, r, images, guide, out, workspace=None, stream=None):
        ptrs = _lib.ptr_array(images)
        ws = self._workspace if workspace is None else workspace
        st = self._stream() if stream is None else _lib.C.c_void_p(stream.cuda_stream)
        _lib.check(_lib.lib.ss_estimate_distance(
            self._raylut[slot].data_ptr(), ptrs, len(images), self._cams, self._rts[r], r, self._cands,
            self.candidate_count, self.rows, self.cols, float(self.sigma_i), float(self.sigma_s),
            guide.data_ptr() if guide is not None else None, out.data_ptr(), ws.data_ptr(), ws.numel(), st))
        return out

    def estimate_fisheye_distance(self, reference_image, guide, reference_calibration, selected_camera, images):
        """Distance map of one reference fisheye (depth_estimation.py:137-203):
        sweep, cost, cost filter, argmin with parabola refinement, distance
        filter.  The camera selection of ``reference_calibration`` made at
        construction is used; ``selected_camera`` is accepted for interface
        compatibility."""
        slot, r = self._reference_slot(reference_calibration)
        imgs = [_hwc(i) for i in images]
        g = guide.contiguous() if guide is not None else None
        out = torch.empty([self.rows, self.cols], device=self.device)
        return self._estimate(slot, r, imgs, g, out)

    def estimate_RGBD_panorama(self, images_to_match, images_to_stitch):
        """Per-frame entry point (depth_estimation.py:205-243).

        Args:
            images_to_match: C x float32 [rows, cols, 3] in [0, 255]
            images_to_stitch: R x float32 [stitch rows, stitch cols, 3] in [0, 255]
        Returns:
            rgb: uint8 [pano rows, pano cols, 3]; distance: float32 [pano rows, pano cols]
            (buffers reused by the next call)
        """
        imgs = [_hwc(i) for i in images_to_match]
        stitch = [i.contiguous() for i in images_to_stitch]
        R = len(self.references_indices)
        if len(stitch) != R:
            raise ValueError("expected %d images to stitch" % R)
        if self._stitch_u8 is None or tuple(self._stitch_u8.shape[1:]) != tuple(stitch[0].shape):
            self._stitch_u8 = torch.empty([R] + list(stitch[0].shape), dtype=torch.uint8, device=self.device)
        cur = torch.cuda.current_stream(self.device)
        self._fork.record(cur)
        for slot, r in enumerate(self.references_indices):
            s = self._streams[slot]
            s.wait_event(self._fork)
            for t in imgs + stitch:
                t.record_stream(s)
            self._estimate(slot, r, imgs, None, self._distance_maps[slot], self._workspaces[slot], s)
            st = _lib.C.c_void_p(s.cuda_stream)
            _lib.check(_lib.lib.ss_to_u8(stitch[slot].data_ptr(), self._stitch_u8[slot].data_ptr(),
                                         stitch[slot].numel(), st))
            self.fishey_stitcher.reproject_and_inpaint(slot, self._distance_maps[slot], stream=s)
            self._joins[slot].record(s)
        for ev in self._joins:
            cur.wait_event(ev)
        return self.fishey_stitcher.merge([self._stitch_u8[k] for k in range(R)],
                                          [self._distance_maps[k] for k in range(R)])
