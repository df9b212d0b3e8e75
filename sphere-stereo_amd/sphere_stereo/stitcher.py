"""Stitcher with the reference's interface (python/stitcher.py:35-220).

Set-up (inpainting tables, blending tables, their box smoothing) and the
per-frame reprojection, inpainting and merge all run in the gfx950 library.
The 4x4 pose algebra of stitcher.py:130 and 141 stays a torch expression on
the CPU (fp32, as the reference forms it).

Intentional differences, both removing data races of the reference: the
z-test of the reprojection is an exact atomic minimum (the reference runs a
non-atomic test twice), and inpainting passes are Jacobi steps (the reference
updates the map in place while other threads read it).
"""
import numpy as np
import torch

from . import _lib
from .utils import _cpu32


def vectorize_calibration(calibration, device):
    """Intrinsics record [fl*s, principal*s, xi, alpha] (stitcher.py:35-46)."""
    v = torch.zeros([6], device=device)
    v[0:2] = calibration.fl * calibration.matching_scale
    v[2:4] = calibration.principal * calibration.matching_scale
    v[4] = calibration.xi
    v[5] = calibration.alpha
    return v


def _intrinsics(vec):
    vals = _cpu32(vec).tolist()
    rec = _lib.ss_intrinsics()
    rec.fx, rec.fy, rec.cx, rec.cy, rec.xi, rec.alpha = vals
    return rec


def inpaint_thresholds(max_dist):
    """The two tests of inpaintKernel: empty when d >= MAX_DIST + 0.1 (a double
    sum, stitcher.cu:287) and usable when d <= MAX_DIST + 0.1f (stitcher.cu:311),
    which is an fp32 sum when max_dist is an int literal and a double sum when
    it is a float (the constant is substituted as str(max_dist))."""
    thr_empty = float(max_dist) + 0.1
    if isinstance(max_dist, (int, np.integer)):
        thr_src = float(np.float32(max_dist) + np.float32(0.1))
    else:
        thr_src = float(max_dist) + float(np.float32(0.1))
    return thr_empty, thr_src


class Stitcher:
    def __init__(self, calibrations, reprojection_viewpoint, masks, min_dist, max_dist, matching_resolution,
                 rgb_to_stitch_resolution, panorama_resolution, device, smoothing_radius=15,
                 inpainting_iterations=32):
        """RGB-D fisheye -> panorama stitcher (stitcher.py:49-164).

        Args follow the reference: calibrations and masks of the references,
        resolutions as (cols, rows), smoothing_radius of the box filter applied
        to the blending weights, inpainting_iterations passes of hole filling.
        """
        self.device = torch.device(device)
        self.max_dist = max_dist
        self.min_dist = min_dist
        self.inpainting_iterations = int(inpainting_iterations)
        self.cols, self.rows = int(matching_resolution[0]), int(matching_resolution[1])
        self.scols, self.srows = int(rgb_to_stitch_resolution[0]), int(rgb_to_stitch_resolution[1])
        self.pcols, self.prows = int(panorama_resolution[0]), int(panorama_resolution[1])
        R = len(calibrations)
        if not 1 <= R <= _lib.SS_MAX_REFERENCES:
            raise ValueError("between 1 and %d references supported" % _lib.SS_MAX_REFERENCES)
        self.R = R
        self.thr_empty, self.thr_src = inpaint_thresholds(max_dist)
        H, W, dev = self.rows, self.cols, self.device
        stream = _lib.stream_of(torch.empty(0, device=dev))

        viewpoint = torch.cat([_cpu32(reprojection_viewpoint), torch.ones([1])])
        self.calibration_vectors_list = [vectorize_calibration(c, dev) for c in calibrations]
        self.intrinsics = (_lib.ss_intrinsics * R)(*[_intrinsics(v) for v in self.calibration_vectors_list])
        self.translations_list = [torch.matmul(torch.inverse(_cpu32(c.rt)), viewpoint)[:3] for c in calibrations]
        self.translations_host = _lib.float_array(torch.cat(self.translations_list).tolist())
        rotations = [torch.inverse(_cpu32(c.rt)[:3, :3]) for c in calibrations]
        self.rotations_host = _lib.float_array(torch.cat([r.reshape(-1) for r in rotations]).tolist())

        # buffers reused by every stitch() call
        self.reprojected_distances = torch.zeros([R, H, W], device=dev)
        self.inpaint_scratch = torch.zeros([R, H, W], device=dev)  # one per reference: they may run concurrently
        self.RGB_panorama = torch.zeros([self.prows, self.pcols, 3], dtype=torch.uint8, device=dev)
        self.distance_panorama = torch.zeros([self.prows, self.pcols], device=dev)

        # inpainting tables (stitcher.py:126-139)
        self.inpainting_weights_list = []
        for r in range(R):
            lut = torch.zeros([H, W, 2], dtype=torch.uint8, device=dev)
            _lib.check(_lib.lib.ss_inpaint_lut(self.intrinsics[r], _lib.float_array(self.translations_list[r]),
                                               H, W, float(min_dist), float(max_dist), lut.data_ptr(), stream))
            self.inpainting_weights_list.append(lut)

        # blending tables (stitcher.py:141-164)
        r_ = int(smoothing_radius)
        stacked = torch.cat([m.to(dev, torch.float32).reshape(1, H, W) for m in masks], dim=0).contiguous()
        smoothed = torch.empty_like(stacked)
        tmp = torch.empty([max(H, self.prows) + 2 * r_, max(W, self.pcols)], device=dev)
        _lib.check(_lib.lib.ss_box_filter(stacked.data_ptr(), smoothed.data_ptr(), tmp.data_ptr(), R, H, W, r_,
                                          1.0, stream))
        self.blending_sampling = torch.zeros([R, self.prows, self.pcols, 2], device=dev)
        raw = torch.zeros([R, self.prows, self.pcols], device=dev)
        _lib.check(_lib.lib.ss_blend_lut(self.intrinsics, self.rotations_host, self.translations_host, R,
                                         smoothed.data_ptr(), H, W, self.prows, self.pcols, float(min_dist),
                                         float(max_dist), self.blending_sampling.data_ptr(), raw.data_ptr(), stream))
        self.blending_weights = torch.zeros_like(raw)
        _lib.check(_lib.lib.ss_box_filter(raw.data_ptr(), self.blending_weights.data_ptr(), tmp.data_ptr(), R,
                                          self.prows, self.pcols, r_, 0.0, stream))
        _lib.check(_lib.lib.ss_normalize_weights(self.blending_weights.data_ptr(), R, self.prows * self.pcols,
                                                 stream))
        torch.cuda.synchronize(dev)  # tmp and raw are released on return

    def reproject_and_inpaint(self, r, distance_map, stream=None):
        """Reference r's part of stitch(): z-tested reprojection to the
        panorama viewpoint and hole filling (stitcher.py:178-200), enqueued on
        ``stream`` (default: torch's current stream)."""
        if distance_map.dtype != torch.float32 or distance_map.numel() != self.rows * self.cols or \
                not distance_map.is_contiguous():
            raise ValueError("distance maps must be contiguous float32 [%d, %d]" % (self.rows, self.cols))
        st = _lib.C.c_void_p(stream.cuda_stream) if stream is not None else _lib.stream_of(self.distance_panorama)
        rep = self.reprojected_distances[r]
        _lib.check(_lib.lib.ss_reproject(distance_map.data_ptr(), rep.data_ptr(), self.rows, self.cols,
                                         self.intrinsics[r], _lib.float_array(self.translations_list[r]), st))
        _lib.check(_lib.lib.ss_inpaint(rep.data_ptr(), self.inpaint_scratch[r].data_ptr(),
                                       self.inpainting_weights_list[r].data_ptr(), self.rows, self.cols,
                                       self.inpainting_iterations, self.thr_empty, self.thr_src, st))

    def merge(self, images, distance_maps, stream=None):
        """The blend of stitch() (stitcher.py:205-218) over the reprojected maps
        left by reproject_and_inpaint for every reference."""
        for img in images:
            if img.dtype != torch.uint8 or tuple(img.shape) != (self.srows, self.scols, 3) or not img.is_contiguous():
                raise ValueError("images must be contiguous uint8 [%d, %d, 3]" % (self.srows, self.scols))
        st = _lib.C.c_void_p(stream.cuda_stream) if stream is not None else _lib.stream_of(self.distance_panorama)
        reproj = _lib.ptr_array([self.reprojected_distances[r] for r in range(self.R)])
        _lib.check(_lib.lib.ss_merge(
            self.blending_sampling.data_ptr(), self.blending_weights.data_ptr(), reproj,
            _lib.ptr_array(distance_maps), _lib.ptr_array(images), self.srows, self.scols, self.translations_host,
            self.intrinsics, self.R, self.rows, self.cols, self.prows, self.pcols, self.distance_panorama.data_ptr(),
            self.RGB_panorama.data_ptr(), st))
        return self.RGB_panorama, self.distance_panorama

    def stitch(self, images, distance_maps):
        """images: R x uint8 [srows, scols, 3]; distance_maps: R x float32
        [rows, cols].  Returns the (rgb, distance) panoramas, buffers reused
        by the next call (stitcher.py:166-220)."""
        if len(images) != self.R or len(distance_maps) != self.R:
            raise ValueError("expected %d images and distance maps" % self.R)
        for r in range(self.R):
            self.reproject_and_inpaint(r, distance_maps[r])
        return self.merge(images, distance_maps)
