"""ISB_Filter with the reference's interface (python/isb_filter.py:35-136).

``apply(guide, cost, sigma_i, sigma_s)`` filters the caller's tensors in place
and returns ``(cost, guide)``, as the reference does (its level 0 aliases the
caller's tensors).  The work is done by ss_isb_filter_planar: the guide and
the [N][H][W] volume are repacked into the library's candidate-contiguous
layout, filtered by the gfx950 kernels and written back.
"""
import torch

from . import _lib


class ISB_Filter:
    def __init__(self, candidate_count, resolution, device):
        """Fast edge-preserving filter (isb_filter.py:36-70).

        Args:
            candidate_count: number of channels of the filtered quantity
            resolution: (cols, rows) of the guide and the cost
            device: GPU used for processing
        """
        self.candidate_count = int(candidate_count)
        self.cols, self.rows = int(resolution[0]), int(resolution[1])
        self.device = torch.device(device)
        self.scale_count = _lib.lib.ss_isb_levels(self.rows, self.cols)
        if self.scale_count < 1:
            # the reference indexes an empty level list and raises IndexError
            raise ValueError("resolution %dx%d gives no pyramid level" % (self.cols, self.rows))
        if _lib.lib.ss_padded_candidates(self.candidate_count) < 0:
            raise ValueError("candidate_count %d unsupported (1..%d)" % (candidate_count, _lib.SS_MAX_CANDIDATES))
        nbytes = _lib.lib.ss_isb_workspace_bytes(self.rows, self.cols, self.candidate_count, _lib.SS_VOLUME_F32, 1)
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.device)

    def apply(self, guide, cost, sigma_i, sigma_s):
        """Filter ``cost`` ([N, rows, cols] float32) guided by ``guide``
        ([rows, cols, 3] uint8), both in place (isb_filter.py:72-136)."""
        if guide.dtype != torch.uint8 or tuple(guide.shape) != (self.rows, self.cols, 3):
            raise ValueError("guide must be uint8 [%d, %d, 3]" % (self.rows, self.cols))
        shape = (self.candidate_count, self.rows, self.cols)
        if cost.dtype != torch.float32 or cost.numel() != self.candidate_count * self.rows * self.cols:
            raise ValueError("cost must be float32 %s" % (shape,))
        if not (guide.is_contiguous() and cost.is_contiguous()):
            raise ValueError("guide and cost must be contiguous (they are filtered in place)")
        if guide.device != self.workspace.device or cost.device != self.workspace.device:
            raise ValueError("guide and cost must be on %s" % self.workspace.device)
        _lib.check(_lib.lib.ss_isb_filter_planar(
            guide.data_ptr(), cost.data_ptr(), self.rows, self.cols, self.candidate_count, float(sigma_i),
            float(sigma_s), self.workspace.data_ptr(), self.workspace.numel(), _lib.stream_of(cost)))
        return cost, guide
