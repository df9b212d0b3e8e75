"""sphere_stereo: sphere-sweeping stereo per-frame path for AMD MI355X (gfx950).

Drop-in counterparts of KAIST-VCLAB/sphere-stereo's RGBD_Estimator,
ISB_Filter and Stitcher, backed by libsphere_stereo.so (HIP kernels behind a
C ABI, include/sphere_stereo.h).
"""
from . import _lib  # noqa: F401  (fails loudly if the library is not built)
from .depth_estimation import FrameBatch, RGBD_Estimator, load_sweep_tuning, save_sweep_tuning
from .differentiable import DifferentiableDepth
from .isb_filter import ISB_Filter
from .stitcher import Stitcher, vectorize_calibration
from .streaming import StreamingDriver
from .utils import Calibration, parse_json_calib, project, rgb2yCbCr, unproject

__all__ = ["RGBD_Estimator", "FrameBatch", "save_sweep_tuning", "load_sweep_tuning", "DifferentiableDepth", "ISB_Filter", "Stitcher", "StreamingDriver", "vectorize_calibration", "Calibration", "parse_json_calib",
           "project", "unproject", "rgb2yCbCr"]
