"""Streaming multi-frame driver: the reference's serial frame loop (main.py:87-125:
read and resize every frame on the host, then one estimate_RGBD_panorama per
frame, each followed by a device-to-host copy of the panorama) as a pipeline
on one GPU.

Per frame the host only copies the decoded fisheyes (as cv2.imread gives them:
HWC uint8, uint16 or float32) into a pinned staging buffer.  The upload runs on
a copy stream into one of two device buffers, so frame i+1's upload overlaps
frame i's compute; the input stage (conversion + resize, utils.py:157-204) runs
on the GPU into stable buffers, so the frame itself can be replayed from a HIP
graph; the RGB panorama and the inverse distance main.py:114 saves are copied
to one of two device output buffers and downloaded on the copy stream while
the next frame computes.  Results come back one frame late, in order.  All
streams run at default priority; staging memory is allocated pinned.
"""
import numpy as np
import torch

from .evaluation import inverse_distance
from .utils import prepare_inputs

_TORCH_TYPES = {np.dtype(np.uint8): torch.uint8, np.dtype(np.uint16): torch.uint16,
                np.dtype(np.float32): torch.float32}


class StreamingDriver:
    """Runs ``estimator`` (an RGBD_Estimator) over a stream of frames.

    Args:
        estimator: the RGBD_Estimator (its configuration fixes C, the
            references and all resolutions).
        image_shape: the decoded fisheyes' shape (rows, cols, 3), as
            read_input_images checks it (utils.py:172).
        image_dtype: numpy dtype of the decoded fisheyes (uint8, uint16 or
            float32 already in [0, 255]).
        use_graph: replay each frame from a captured HIP graph.
    """

    def __init__(self, estimator, image_shape, image_dtype=np.uint8, use_graph=True):
        self.est = estimator
        self.device = estimator.device
        self.shape = tuple(image_shape)
        dt = np.dtype(image_dtype)
        if dt not in _TORCH_TYPES:
            raise ValueError("decoded images must be uint8, uint16 or float32")
        self.np_dtype = dt
        tdt = _TORCH_TYPES[dt]
        C = len(estimator.calibrations)
        self.C = C
        self.compute = torch.cuda.Stream(device=self.device)
        self.copy = torch.cuda.Stream(device=self.device)
        self.pinned_in = [[torch.empty(self.shape, dtype=tdt, pin_memory=True) for _ in range(C)] for _ in range(2)]
        self.raw = [[torch.empty(self.shape, dtype=tdt, device=self.device) for _ in range(C)] for _ in range(2)]
        st = estimator.fishey_stitcher
        self.rgb_dev = [torch.empty_like(st.RGB_panorama) for _ in range(2)]
        self.inv_dev = [torch.empty_like(st.distance_panorama) for _ in range(2)]
        self.rgb_host = [torch.empty(tuple(st.RGB_panorama.shape), dtype=torch.uint8, pin_memory=True)
                         for _ in range(2)]
        self.inv_host = [torch.empty(tuple(st.distance_panorama.shape), dtype=torch.float32, pin_memory=True)
                         for _ in range(2)]
        ev = lambda: [torch.cuda.Event() for _ in range(2)]  # noqa: E731
        self.h2d_done, self.raw_free, self.out_ready, self.d2h_done = ev(), ev(), ev(), ev()
        self.use_graph = use_graph
        self._prepared = None

    def _stage(self, i, images):
        """Host side of frame i: its images into pinned staging buffer i % 2."""
        b = i % 2
        if len(images) != self.C:
            raise ValueError("expected %d images per frame" % self.C)
        if i >= 2:
            self.h2d_done[b].synchronize()  # frame i-2's upload from this buffer is over
        for k, img in enumerate(images):
            a = np.asarray(img)
            if a.shape != self.shape or a.dtype != self.np_dtype:
                raise ValueError("frame images must be %s %s" % (self.np_dtype, self.shape))
            self.pinned_in[b][k].copy_(torch.from_numpy(np.ascontiguousarray(a)))

    def _enqueue(self, i):
        """Device side of frame i: upload, input stage, frame, outputs, download."""
        b = i % 2
        est = self.est
        with torch.cuda.stream(self.copy):
            if i >= 2:
                self.copy.wait_event(self.raw_free[b])  # frame i-2's input stage is done with raw[b]
            for k in range(self.C):
                self.raw[b][k].copy_(self.pinned_in[b][k], non_blocking=True)
            self.h2d_done[b].record(self.copy)
        with torch.cuda.stream(self.compute):
            self.compute.wait_event(self.h2d_done[b])
            self._prepared = prepare_inputs(self.raw[b], (est.cols, est.rows), (est._stitch_res[1], est._stitch_res[0]),
                                            est.references_indices, out=self._prepared)
            self.raw_free[b].record(self.compute)
            est.use_graph = self.use_graph
            rgb, dist = est.estimate_RGBD_panorama(*self._prepared)
            if i >= 2:
                self.compute.wait_event(self.d2h_done[b])  # frame i-2's download from out[b] is over
            self.rgb_dev[b].copy_(rgb)
            inverse_distance(dist, out=self.inv_dev[b])
            self.out_ready[b].record(self.compute)
        with torch.cuda.stream(self.copy):
            self.copy.wait_event(self.out_ready[b])
            self.rgb_host[b].copy_(self.rgb_dev[b], non_blocking=True)
            self.inv_host[b].copy_(self.inv_dev[b], non_blocking=True)
            self.d2h_done[b].record(self.copy)

    def _result(self, i, copy):
        b = i % 2
        self.d2h_done[b].synchronize()
        rgb, inv = self.rgb_host[b].numpy(), self.inv_host[b].numpy()
        return (i, rgb.copy(), inv.copy()) if copy else (i, rgb, inv)

    def run(self, frames, copy=True):
        """frames: an iterable of per-frame lists of C decoded fisheyes (host
        arrays).  Yields (index, rgb uint8 [PH, PW, 3] BGR, inverse distance
        float32 [PH, PW]) per frame, as main.py:112-114 keeps them.  With
        copy=False the arrays are views of the pinned download buffers, valid
        until the next frame is requested from the generator."""
        i = -1
        for i, images in enumerate(frames):
            self._stage(i, images)
            self._enqueue(i)
            if i >= 1:
                yield self._result(i - 1, copy)
        if i >= 0:
            yield self._result(i, copy)
