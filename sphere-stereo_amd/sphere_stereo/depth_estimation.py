"""RGBD_Estimator with the reference's interface (python/depth_estimation.py).

The estimator keeps the reference's constructor, ``select_camera``,
``estimate_fisheye_distance`` and ``estimate_RGBD_panorama``.  A frame is one
ss_frame_run call into the native executor, on torch's current stream: one
sweep launch covers every reference (and writes their guides), then every
later pass (each cost filter level, the final up pass fused with the argmin,
the distance filter, the uint8 cast of the stitch images, the reprojection,
the inpainting) is one launch for all references, then the merge.  With ``use_graph = True`` the frame is captured once into a HIP graph and
replayed on later calls with the same input tensors.  Init-time work (camera
selection, unit-ray tables, stitching tables) also runs on the GPU.

``volume_dtype`` (keyword, default "float32" = the reference) selects the
element type the cost volume is stored in: "float16" or "bfloat16" halve its
HBM traffic; sums, filter weights and the refinement stay fp32.
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
                 device, volume_dtype="float32"):
        """RGB-D estimation set-up (depth_estimation.py:39-81): camera selection
        for adaptive matching, filters and stitcher.  Arguments as the
        reference: resolutions are (cols, rows), masks are [1, rows, cols].
        volume_dtype: storage type of the cost volume ("float32", "float16",
        "bfloat16")."""
        if volume_dtype not in _lib.VOLUME_TYPES:
            raise ValueError("volume_dtype must be one of %s" % sorted(_lib.VOLUME_TYPES))
        self.volume_dtype = volume_dtype
        self._vtype = _lib.VOLUME_TYPES[volume_dtype]
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
        nbytes = _lib.lib.ss_estimate_workspace_bytes(self.rows, self.cols, self.candidate_count, self._vtype)
        if nbytes == 0:
            raise ValueError("unsupported matching resolution / candidate count")
        R = len(self.references_indices)
        # one workspace per reference: the references' chains run concurrently
        self._workspaces = [torch.empty(nbytes, dtype=torch.uint8, device=self.device) for _ in range(R)]
        self._workspace = self._workspaces[0]
        self._distance_maps = torch.zeros([R, self.rows, self.cols], device=self.device)
        self._stitch_res = (int(rgb_to_stitch_resolution[1]), int(rgb_to_stitch_resolution[0]))
        self._stitch_u8 = torch.zeros([R, self._stitch_res[0], self._stitch_res[1], 3], dtype=torch.uint8,
                                      device=self.device)
        self.use_graph = False
        self._frame = None
        self.select_camera(masks)
        self._frame = self._make_frame()

    def _make_frame(self):
        d = self._frame_desc()
        handle = _lib.C.c_void_p()
        _lib.check(_lib.lib.ss_frame_create(_lib.C.byref(d), _lib.C.byref(handle)))
        return handle

    def autotune(self, images_to_match, repeats=5, maps=None):
        """Times the scheduling variants of the sweep (ss_sweep_autotune:
        candidate chunking and LDS staging) and of the cost filter
        (ss_estimate_filter_autotune: lanes per pixel and tile order, timed
        with the fused argmin on the first reference's swept volume) on these
        images, and records the fastest of each for this configuration; every
        later frame of the configuration uses them; then the inpainting tile
        (ss_inpaint_autotune) on that reference's distance map reprojected as
        the stitcher reprojects it, ``maps`` copies inpainted together (default:
        the number of references, as a frame runs them), recorded for the map
        size; with ``maps`` left at its default the tile is then picked by
        timing this estimator's whole frame with each tile (zero stitch
        images).  A variant other than the default is taken only when it is at
        least 2 % faster.  Every variant gives the same bits.  Returns
        {"variant": v, "ms": [...], "filter_variant": f, "filter_ms": [...],
        "inpaint_variant": i, "inpaint_ms": [...], "inpaint_frame_ms": [...]}
        (ms per launch, per run or per frame; -1 for a filter variant that
        does not exist for this candidate count).  A captured frame graph is
        dropped (the next graph call captures the tuned kernels), and the
        frames it times overwrite the estimator's output buffers."""
        imgs = [_hwc(i) for i in images_to_match]
        if len(imgs) != len(self.calibrations):
            raise ValueError("expected %d images to match" % len(self.calibrations))
        if self.candidate_count < 2:
            return {"variant": 0, "ms": [], "filter_variant": 0, "filter_ms": []}
        nv = _lib.lib.ss_sweep_variant_count()
        NP = _lib.lib.ss_padded_candidates(self.candidate_count)
        eb = 4 if self.volume_dtype == "float32" else 2
        scratch = torch.empty(self.rows * self.cols * NP * eb, dtype=torch.uint8, device=self.device)
        ms = (_lib.C.c_float * nv)()
        best = _lib.C.c_int(0)
        r = self.references_indices[0]
        _lib.check(_lib.lib.ss_sweep_autotune(self._raylut[0].data_ptr(), _lib.ptr_array(imgs), len(imgs), self._cams,
                                              self._rts[r], r, self._cands, self.candidate_count, self.rows,
                                              self.cols, self._vtype, scratch.data_ptr(), int(repeats),
                                              self._stream(), ms, _lib.C.byref(best)))
        _TUNED[(self.rows, self.cols, self.candidate_count, self._vtype)] = int(best.value)
        # the filter, on the volume and guide a sweep leaves in the workspace
        dist = torch.empty([self.rows, self.cols], device=self.device)
        self._estimate(0, r, imgs, None, dist)
        nf = _lib.lib.ss_isb_variant_count()
        fms = (_lib.C.c_float * nf)()
        fbest = _lib.C.c_int(0)
        ws = self._workspaces[0]
        _lib.check(_lib.lib.ss_estimate_filter_autotune(ws.data_ptr(), ws.numel(), self.rows, self.cols,
                                                        self.candidate_count, self._vtype, self._cands,
                                                        float(self.sigma_i), float(self.sigma_s), dist.data_ptr(),
                                                        int(repeats), self._stream(), fms, _lib.C.byref(fbest)))
        _TUNED_FILTER[(self.rows, self.cols, self.candidate_count, self._vtype)] = int(fbest.value)
        out = {"variant": int(best.value), "ms": [round(float(v), 4) for v in ms],
               "filter_variant": int(fbest.value), "filter_ms": [round(float(v), 4) for v in fms]}
        # the inpainting tile, on that distance map reprojected as the stitcher does
        st = self.fishey_stitcher
        if st.inpainting_iterations > 0:
            nm = len(self.references_indices) if maps is None else int(maps)
            rep = torch.empty([self.rows, self.cols], device=self.device)
            work = torch.empty([nm, self.rows, self.cols], device=self.device)
            tmp = torch.empty_like(work)
            _lib.check(_lib.lib.ss_reproject(dist.data_ptr(), rep.data_ptr(), self.rows, self.cols, st.intrinsics[0],
                                             _lib.float_array(st.translations_list[0]), self._stream()))
            ni = _lib.lib.ss_inpaint_variant_count()
            ims = (_lib.C.c_float * ni)()
            ibest = _lib.C.c_int(0)
            _lib.check(_lib.lib.ss_inpaint_autotune(rep.data_ptr(), nm, work.data_ptr(), tmp.data_ptr(),
                                                    st.inpainting_weights_list[0].data_ptr(), self.rows, self.cols,
                                                    st.inpainting_iterations, st.thr_empty, st.thr_src,
                                                    int(repeats), self._stream(), ims, _lib.C.byref(ibest)))
            _TUNED_INPAINT[(self.rows, self.cols)] = int(ibest.value)
            out.update(inpaint_variant=int(ibest.value), inpaint_ms=[round(float(v), 4) for v in ims])
            if maps is None:
                # for this estimator's own frame, whole frames decide: the
                # inpainting launch meets its maps right after the
                # reprojection wrote them there, which the copies above do
                # not reproduce (a tile faster there by 7 % ran 10 % slower
                # inside the frame)
                fms_f = self._time_frames_per_inpaint_tile(imgs, ni, int(repeats))
                pick = 0
                for v in range(1, ni):
                    if fms_f[v] < 0.98 * fms_f[0] and fms_f[v] < fms_f[pick]:
                        pick = v
                _lib.check(_lib.lib.ss_inpaint_set_variant(self.rows, self.cols, pick))
                _TUNED_INPAINT[(self.rows, self.cols)] = pick
                out.update(inpaint_variant=pick, inpaint_frame_ms=[round(v, 4) for v in fms_f])
        # a captured frame graph keeps the kernels it recorded
        if self._frame:
            _lib.lib.ss_frame_destroy(self._frame)
            self._frame = self._make_frame()
        return out

    def _time_frames_per_inpaint_tile(self, imgs, ni, repeats):
        """ms per frame (eager, zero stitch images) with each inpainting tile
        variant of this map size."""
        st = self.fishey_stitcher
        stitch = [torch.zeros([st.srows, st.scols, 3], device=self.device)
                  for _ in self.references_indices]
        stream = torch.cuda.current_stream(self.device)
        graph, self.use_graph = self.use_graph, False
        out = []
        try:
            for v in range(ni):
                _lib.check(_lib.lib.ss_inpaint_set_variant(self.rows, self.cols, v))
                self.estimate_RGBD_panorama(imgs, stitch)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(max(1, repeats)):
                    self.estimate_RGBD_panorama(imgs, stitch)
                e1.record(stream)
                e1.synchronize()
                out.append(e0.elapsed_time(e1) / max(1, repeats))
        finally:
            self.use_graph = graph
        return out

    def _config_key(self):
        """What the frames of one native batch share (ss_frame_batch_create
        checks the same fields)."""
        st = self.fishey_stitcher
        return (len(self.calibrations), len(self.references_indices), self.candidate_count, self.rows, self.cols,
                st.srows, st.scols, st.prows, st.pcols, st.inpainting_iterations, self._vtype, float(self.sigma_i),
                float(self.sigma_s), st.thr_empty, st.thr_src, tuple(self._cands_list))

    def _frame_desc(self):
        """The ss_frame_desc of this estimator's frame (its rig, tables and
        buffers)."""
        st = self.fishey_stitcher
        d = _lib.ss_frame_desc()
        d.C, d.R, d.N, d.H, d.W = len(self.calibrations), len(self.references_indices), self.candidate_count, \
            self.rows, self.cols
        d.Hs, d.Ws, d.PH, d.PW = st.srows, st.scols, st.prows, st.pcols
        d.inpaint_iterations = st.inpainting_iterations
        d.volume_type = self._vtype
        for k in range(d.C):
            d.cams[k] = self._cams[k]
        for n, c in enumerate(self._cands_list):
            d.cands[n] = c
        d.sigma_i, d.sigma_s = float(self.sigma_i), float(self.sigma_s)
        d.thr_empty, d.thr_src = st.thr_empty, st.thr_src
        d.workspace_bytes = self._workspaces[0].numel()
        d.d_sampling, d.d_weights = st.blending_sampling.data_ptr(), st.blending_weights.data_ptr()
        d.d_dist_pano, d.d_rgb_pano = st.distance_panorama.data_ptr(), st.RGB_panorama.data_ptr()
        for slot, r in enumerate(self.references_indices):
            d.refs[slot] = r
            for i in range(12 * d.C):
                d.rts[slot][i] = self._rts[r][i]
            d.d_raylut[slot] = self._raylut[slot].data_ptr()
            d.d_workspace[slot] = self._workspaces[slot].data_ptr()
            d.d_dist[slot] = self._distance_maps[slot].data_ptr()
            d.intr[slot] = st.intrinsics[slot]
            for i in range(3):
                d.trans[slot][i] = float(st.translations_list[slot][i])
            d.d_inpaint_lut[slot] = st.inpainting_weights_list[slot].data_ptr()
            d.d_reproj[slot] = st.reprojected_distances[slot].data_ptr()
            d.d_scratch[slot] = st.inpaint_scratch[slot].data_ptr()
            d.d_stitch_u8[slot] = self._stitch_u8[slot].data_ptr()
        return d

    def __del__(self):
        # at interpreter shutdown the module globals may already be cleared
        lib = getattr(_lib, "lib", None)
        if getattr(self, "_frame", None) and lib is not None:
            lib.ss_frame_destroy(self._frame)
            self._frame = None

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

    def _estimate(self, slot, r, images, guide, out, workspace=None, stream=None, raylut=None):
        ptrs = _lib.ptr_array(images)
        ws = self._workspace if workspace is None else workspace
        st = self._stream() if stream is None else _lib.C.c_void_p(stream.cuda_stream)
        lut = self._raylut[slot] if raylut is None else raylut
        _lib.check(_lib.lib.ss_estimate_distance(
            lut.data_ptr(), ptrs, len(images), self._cams, self._rts[r], r, self._cands,
            self.candidate_count, self.rows, self.cols, float(self.sigma_i), float(self.sigma_s), self._vtype,
            guide.data_ptr() if guide is not None else None, out.data_ptr(), ws.data_ptr(), ws.numel(), st))
        return out

    def _selection_raylut(self, r, selected_camera):
        """Unit-ray table of reference r with a caller-given camera selection
        ([1, rows, cols] or [rows, cols] integers).  A value outside [0, C)
        selects no camera: the reference's masked copy (depth_estimation.py:
        155-170) compares it with every camera index and matches none."""
        s = torch.as_tensor(selected_camera, device=self.device).reshape(self.rows, self.cols)
        if s.dtype.is_floating_point or s.dtype == torch.bool:
            raise ValueError("selected_camera must hold integer camera indices")
        C = len(self.calibrations)
        sel = torch.where((s >= 0) & (s < C), s, torch.full_like(s, -1)).to(torch.int8).contiguous()
        lut = torch.empty([self.rows, self.cols, 4], dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib.ss_ray_lut(self._cams[r], sel.data_ptr(), self.rows, self.cols, lut.data_ptr(),
                                       self._stream()))
        return lut, sel

    def estimate_fisheye_distance(self, reference_image, guide, reference_calibration, selected_camera, images):
        """Distance map of one reference fisheye (depth_estimation.py:137-203):
        sweep, cost, cost filter, argmin with parabola refinement, distance
        filter.  ``selected_camera`` is the camera each reference pixel is
        matched against (depth_estimation.py:155-170), as the reference uses
        the one it is given: the selection built at construction
        (``selected_cameras[slot]``, or None) runs from its cached ray table;
        any other is turned into a ray table for this call."""
        slot, r = self._reference_slot(reference_calibration)
        imgs = [_hwc(i) for i in images]
        g = guide.contiguous() if guide is not None else None
        out = torch.empty([self.rows, self.cols], device=self.device)
        lut = None
        if selected_camera is not None and selected_camera is not self.selected_cameras[slot]:
            lut, sel = self._selection_raylut(r, selected_camera)
            self._call_selection = sel  # kept alive until the next call (the kernels may still read it)
        self._call_raylut = lut
        return self._estimate(slot, r, imgs, g, out, raylut=lut)

    def estimate_RGBD_panorama(self, images_to_match, images_to_stitch):
        """Per-frame entry point (depth_estimation.py:205-243).

        Args:
            images_to_match: C x float32 [rows, cols, 3] in [0, 255]
            images_to_stitch: R x float32 [stitch rows, stitch cols, 3] in [0, 255]
        Returns:
            rgb: uint8 [pano rows, pano cols, 3]; distance: float32 [pano rows, pano cols]
            (buffers reused by the next call)
        """
        imgs, stitch = self._check_inputs(images_to_match, images_to_stitch)
        self._inputs = (imgs, stitch)  # kept alive until the next call (the frame may still read them)
        _lib.check(_lib.lib.ss_frame_run(self._frame, _lib.ptr_array(imgs), _lib.ptr_array(stitch),
                                         int(bool(self.use_graph)), self._stream()))
        return self.fishey_stitcher.RGB_panorama, self.fishey_stitcher.distance_panorama

    def _check_inputs(self, images_to_match, images_to_stitch):
        """The frame's images as contiguous HWC tensors, after the checks of
        estimate_RGBD_panorama."""
        imgs = [_hwc(i) for i in images_to_match]
        stitch = [i.contiguous() for i in images_to_stitch]
        if len(imgs) != len(self.calibrations) or len(stitch) != len(self.references_indices):
            raise ValueError("expected %d images to match and %d to stitch" % (len(self.calibrations),
                                                                              len(self.references_indices)))
        for t in imgs:
            if t.dtype != torch.float32 or tuple(t.shape) != (self.rows, self.cols, 3) or t.device != self.device:
                raise ValueError("images to match must be float32 [%d, %d, 3] on %s" % (self.rows, self.cols,
                                                                                       self.device))
        for t in stitch:
            if t.dtype != torch.float32 or tuple(t.shape) != self._stitch_res + (3,) or t.device != self.device:
                raise ValueError("images to stitch must be float32 [%d, %d, 3] on %s" % (self._stitch_res + (
                    self.device,)))
        return imgs, stitch


# variants chosen by RGBD_Estimator.autotune in this process: (rows, cols,
# candidates, volume type) -> variant, for the sweep and for the cost filter
_TUNED = {}
_TUNED_FILTER = {}
_TUNED_INPAINT = {}  # (rows, cols) -> inpainting tile variant


def save_sweep_tuning(path):
    """Writes the sweep, cost-filter and inpainting variants tuned in this process
    (RGBD_Estimator.autotune, FrameBatch.autotune) to a JSON file, with the
    device they were measured on.  Returns the number of records written."""
    import json

    dev = torch.cuda.get_device_name() if torch.cuda.is_available() else ""

    def entries(tab):
        return [{"rows": k[0], "cols": k[1], "candidates": k[2], "volume_type": k[3], "variant": v}
                for k, v in sorted(tab.items())]

    data = {"device": dev, "sweep_variants": entries(_TUNED), "filter_variants": entries(_TUNED_FILTER),
            "inpaint_variants": [{"rows": k[0], "cols": k[1], "variant": v} for k, v in sorted(_TUNED_INPAINT.items())]}
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    return len(data["sweep_variants"]) + len(data["filter_variants"]) + len(data["inpaint_variants"])


def load_sweep_tuning(path):
    """Applies the variants of a save_sweep_tuning file (ss_sweep_set_variant,
    ss_isb_set_variant, ss_inpaint_set_variant) when it was written on a device of the same name;
    every later frame of those configurations uses them.  Returns the number
    of records applied."""
    import json

    with open(path) as f:
        data = json.load(f)
    dev = torch.cuda.get_device_name() if torch.cuda.is_available() else ""
    if data.get("device", "") != dev:
        return 0
    n = 0
    for name, setter, tab in (("sweep_variants", _lib.lib.ss_sweep_set_variant, _TUNED),
                              ("filter_variants", _lib.lib.ss_isb_set_variant, _TUNED_FILTER)):
        for e in data.get(name, []):
            key = (int(e["rows"]), int(e["cols"]), int(e["candidates"]), int(e["volume_type"]))
            _lib.check(setter(key[0], key[1], key[2], key[3], int(e["variant"])))
            tab[key] = int(e["variant"])
            n += 1
    for e in data.get("inpaint_variants", []):
        key = (int(e["rows"]), int(e["cols"]))
        _lib.check(_lib.lib.ss_inpaint_set_variant(key[0], key[1], int(e["variant"])))
        _TUNED_INPAINT[key] = int(e["variant"])
        n += 1
    return n


class FrameBatch:
    """The batching form (no counterpart in the reference, which runs one frame
    per call): several estimators run their frames together.  Their rigs and
    images differ, and so may their configurations (matching resolution,
    candidate count, references, sigmas, panorama, inpainting settings).

    ``single_executor=True`` (the ragged form): one native executor runs every
    pass once for all the frames, whatever their configurations (the filter
    levels once per padded candidate count present).  ``schedule``:
    "persistent" launches are sized to what the GPU holds resident and their
    workers loop over work items (pixel tiles and blocks); "dispatched"
    launches have a worker per item;
    "auto" picks per pass the form measured faster.  The estimators must
    share the volume type and have at least 2 candidates and an 8 x 8
    matching resolution.  The launch records live in a device table the batch
    allocates once.

    ``single_executor=False``: the batch is split into groups of one
    configuration, and each group is one native batch in which every pass
    after the sweep is one launch for a chunk of frames (up to 16
    frame-reference pairs), so the latency-bound coarse pyramid levels and
    the small stitching passes are paid once per chunk instead of once per
    frame.  The groups are enqueued one after the other on the same stream.

    Each estimator keeps its own buffers: after ``estimate_RGBD_panoramas`` its
    ``fishey_stitcher`` holds its frame's panorama, as after
    ``estimate_RGBD_panorama``.
    """

    def __init__(self, estimators, single_executor=False, schedule="auto", lanes=1):
        self.estimators = list(estimators)
        F = len(self.estimators)
        if not 1 <= F <= _lib.SS_MAX_BATCH_FRAMES:
            raise ValueError("a batch holds 1 to %d estimators" % _lib.SS_MAX_BATCH_FRAMES)
        devs = {e.device for e in self.estimators}
        if len(devs) != 1:
            raise ValueError("the estimators of a batch must share one device")
        self.single_executor = bool(single_executor)
        self.schedule = {"persistent": _lib.SS_RAGGED_PERSISTENT, "dispatched": _lib.SS_RAGGED_DISPATCHED,
                         "auto": _lib.SS_RAGGED_AUTO}[schedule]
        self._table = None
        self.groups = []
        self._frames = []
        if self.single_executor:
            self.groups.append(list(range(F)))
            self._frames.append(self._create_ragged())
        else:
            # group by configuration, in order of first appearance
            groups = {}
            for i, e in enumerate(self.estimators):
                groups.setdefault(e._config_key(), []).append(i)
            for idx in groups.values():
                self._frames.append(self._create_group(idx))
                self.groups.append(idx)
        self.use_graph = False
        self._inputs = None
        if lanes != 1:
            self.lanes = lanes

    def _create_group(self, idx):
        descs = (_lib.ss_frame_desc * len(idx))(*[self.estimators[i]._frame_desc() for i in idx])
        handle = _lib.C.c_void_p()
        _lib.check(_lib.lib.ss_frame_batch_create(descs, len(idx), _lib.C.byref(handle)))
        lanes = getattr(self, "_lanes", 1)
        if lanes != 1:
            _lib.check(_lib.lib.ss_frame_set_lanes(handle, lanes))
        return handle

    @property
    def lanes(self):
        """Streams each configuration group's frames alternate over
        (ss_frame_set_lanes): with L > 1 the group is cut into L or more
        chunks of whole frames and each chunk's filters, reprojection,
        inpainting and merges run beside the next chunk's sweeps.  The
        results are the same.  The single executor has one lane."""
        return getattr(self, "_lanes", 1)

    @lanes.setter
    def lanes(self, n):
        n = int(n)
        if self.single_executor and n != 1:
            raise ValueError("the single executor runs on one lane")
        for h in self._frames:
            _lib.check(_lib.lib.ss_frame_set_lanes(h, n))
        self._lanes = n

    def _create_ragged(self):
        F = len(self.estimators)
        descs = (_lib.ss_frame_desc * F)(*[e._frame_desc() for e in self.estimators])
        nbytes = _lib.lib.ss_frame_ragged_table_bytes(descs, F)
        if nbytes == 0:
            raise ValueError("these estimators cannot share one executor: they must share the volume type "
                             "and have >= 2 candidates and an 8 x 8 or larger matching resolution")
        self._table = torch.empty(int(nbytes), dtype=torch.uint8, device=self.estimators[0].device)
        handle = _lib.C.c_void_p()
        _lib.check(_lib.lib.ss_frame_ragged_create(descs, F, self.schedule, self._table.data_ptr(), int(nbytes),
                                                   self.estimators[0]._stream(), _lib.C.byref(handle)))
        return handle

    def autotune(self, frames, repeats=5):
        """The sweep autotuner (RGBD_Estimator.autotune) for each configuration
        group of the batch, on its first frame; then drops captured graphs so
        that the next graph replay records the chosen variants.  Returns one
        result per group.  In the single-executor form (``single_executor=True``)
        each configuration is tuned on its first frame and the executor is
        rebuilt: its sweep runs one launch per variant the configurations
        picked."""
        frames = list(frames)
        if self.single_executor:
            out, seen = [], set()
            for i, e in enumerate(self.estimators):
                if e._config_key() not in seen:
                    seen.add(e._config_key())
                    out.append(e.autotune(frames[i][0], repeats))
            _lib.lib.ss_frame_destroy(self._frames[0])
            self._frames[0] = self._create_ragged()
            return out
        out = []
        for idx in self.groups:
            e = self.estimators[idx[0]]
            # a native batch runs its frames in chunks of up to SS_BATCH_PAIRS
            # frame-reference pairs: the maps one inpainting launch takes
            R = len(e.references_indices)
            maps = min(len(idx), _lib.SS_BATCH_PAIRS // R) * R
            out.append(e.autotune(frames[idx[0]][0], repeats, maps=maps))
        for i, h in enumerate(self._frames):
            # recreate the native batches: a captured graph keeps its kernels
            _lib.lib.ss_frame_destroy(h)
            self._frames[i] = self._create_group(self.groups[i])
        return out

    def estimate_RGBD_panoramas(self, frames):
        """frames: one (images_to_match, images_to_stitch) pair per estimator,
        as estimate_RGBD_panorama takes them.  Returns one (rgb, distance)
        pair per estimator (its reused buffers)."""
        frames = list(frames)
        if len(frames) != len(self.estimators):
            raise ValueError("expected %d frames" % len(self.estimators))
        checked = [e._check_inputs(m, s) for e, (m, s) in zip(self.estimators, frames)]
        inputs = []
        stream = self.estimators[0]._stream()
        for idx, handle in zip(self.groups, self._frames):
            match, stitch = [], []
            for i in idx:
                match += checked[i][0]
                stitch += checked[i][1]
            inputs.append((match, stitch))
            _lib.check(_lib.lib.ss_frame_batch_run(handle, _lib.ptr_array(match), _lib.ptr_array(stitch),
                                                   int(bool(self.use_graph)), stream))
        self._inputs = inputs  # kept alive until the next call
        return [(e.fishey_stitcher.RGB_panorama, e.fishey_stitcher.distance_panorama) for e in self.estimators]

    def __del__(self):
        lib = getattr(_lib, "lib", None)
        for h in getattr(self, "_frames", None) or []:
            if h and lib is not None:
                lib.ss_frame_destroy(h)
        self._frames = []
