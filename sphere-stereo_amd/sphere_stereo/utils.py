"""Host-side helpers with the reference's interface (python/utils.py).

``Calibration``, ``parse_json_calib``, ``unproject``, ``project`` and
``rgb2yCbCr`` keep the reference's names, argument meaning and tensor
semantics (utils.py:40-155): they are small torch expressions used at set-up
time and by callers.  The per-frame path never calls them; it runs the same
mathematics inside the HIP kernels.

The ``*_record`` functions turn a ``Calibration`` into the plain structures of
the C ABI, computing every fp32 product on the CPU exactly as the reference's
tensor code forms it.
"""
import math

import numpy as np
import torch

from . import _lib


class Calibration:
    def __init__(self, original_resolution, principal, fl, xi, alpha, rt, matching_scale):
        """Double-sphere intrinsics, pose and matching scale (utils.py:40-55).

        Args:
            original_resolution, principal, fl, xi, alpha: double-sphere intrinsics
            rt: [4, 4] camera pose (camera to rig)
            matching_scale: [2] scale from the original to the matching resolution
        """
        self.original_resolution = original_resolution
        self.principal = principal
        self.fl = fl
        self.xi = xi
        self.alpha = alpha
        self.rt = rt
        self.matching_scale = matching_scale


def unproject(uv, calib):
    """Pixels -> unit sphere, double-sphere model (utils.py:57-73)."""
    m_xy = (uv - calib.principal * calib.matching_scale) / (calib.fl * calib.matching_scale)
    r2 = torch.sum(m_xy ** 2, dim=-1, keepdim=True)
    m_z = ((1 - calib.alpha ** 2 * r2)
           / (calib.alpha * torch.sqrt(torch.clamp(1 - (2 * calib.alpha - 1) * r2, min=0)) + 1 - calib.alpha))
    point = torch.cat([m_xy, m_z], dim=-1)
    point = ((m_z * calib.xi + torch.sqrt(m_z ** 2 + (1 - calib.xi ** 2) * r2)) / (m_z ** 2 + r2)) * point
    point[..., 2] -= calib.xi
    valid = 1 - (2 * calib.alpha - 1) * r2 >= 0
    return point, valid[..., 0]


def project(point, calib):
    """3-D points -> pixels and validity, double-sphere model (utils.py:75-94)."""
    d1 = torch.norm(point, dim=-1, keepdim=True)
    c = calib.xi * d1 + point[..., 2:3]
    d2 = torch.norm(torch.cat([point[..., :2], c], dim=-1), dim=-1, keepdim=True)
    norm = calib.alpha * d2 + (1 - calib.alpha) * c
    w1 = (1 - calib.alpha) / calib.alpha if calib.alpha > 0.5 else calib.alpha / (1 - calib.alpha)
    w2 = (w1 + calib.xi) / math.sqrt(2 * w1 * calib.xi + calib.xi ** 2 + 1)
    valid = point[..., 2:3] > -w2 * d1
    uv = (calib.fl * calib.matching_scale * point[..., :2]) / norm + calib.principal * calib.matching_scale
    return uv, valid[..., 0]


def _quat_to_matrix(qx, qy, qz, qw):
    try:
        from scipy.spatial.transform import Rotation

        return Rotation.from_quat([qx, qy, qz, qw]).as_matrix()
    except ImportError:  # textbook unit-quaternion rotation matrix
        n = math.sqrt(qx * qx + qy * qy + qz * qz + qw * qw)
        x, y, z, w = qx / n, qy / n, qz / n, qw / n
        return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def parse_json_calib(raw_calibration, matching_resolution, device):
    """basalt calibration JSON -> [Calibration] (utils.py:96-142)."""
    calibrations = []
    for extrinsics, intrinsics, original_resolution in zip(raw_calibration["T_imu_cam"],
                                                           raw_calibration["intrinsics"],
                                                           raw_calibration["resolution"]):
        if intrinsics["camera_type"] != "ds":
            raise Exception("Unexpected camera model. The current implementation only support double sphere.")
        ci = intrinsics["intrinsics"]
        rmat = _quat_to_matrix(extrinsics["qx"], extrinsics["qy"], extrinsics["qz"], extrinsics["qw"])
        t = torch.tensor([extrinsics["px"], extrinsics["py"], extrinsics["pz"]], device=device)
        rt = torch.eye(4, device=device)
        rt[:3, :3] = torch.tensor(rmat, device=device)
        rt[:3, 3] = t
        calibrations.append(Calibration(
            original_resolution,
            torch.tensor([ci["cx"], ci["cy"]], device=device),
            torch.tensor([ci["fx"], ci["fy"]], device=device),
            ci["xi"], ci["alpha"], rt,
            torch.tensor([matching_resolution[0] / original_resolution[0],
                          matching_resolution[1] / original_resolution[1]], device=device)))
    return calibrations


def rgb2yCbCr(rgb):
    """Fixed-coefficient colour transform of the guide (utils.py:144-155)."""
    rgb = rgb.float()
    yuv = torch.zeros_like(rgb)
    yuv[:, :, 0] = torch.clamp(16 + 0.1826 * rgb[:, :, 0] + 0.6142 * rgb[:, :, 1] + 0.062 * rgb[:, :, 2],
                               min=16, max=235)
    yuv[:, :, 1] = torch.clamp(128 - 0.1006 * rgb[:, :, 0] - 0.3386 * rgb[:, :, 1] + 0.4392 * rgb[:, :, 2],
                               min=16, max=240)
    yuv[:, :, 2] = torch.clamp(128 + 0.4392 * rgb[:, :, 0] - 0.3989 * rgb[:, :, 1] - 0.0403 * rgb[:, :, 2],
                               min=16, max=240)
    return yuv


# ---------------------------------------------------------------------------
# C-ABI records
# ---------------------------------------------------------------------------
def _cpu32(t):
    return torch.as_tensor(t).detach().to("cpu", torch.float32)


def camera_record(calib):
    """ss_camera: fl*matching_scale and principal*matching_scale as the fp32
    products of utils.py:62, 94; xi and alpha as Python doubles."""
    fs = (_cpu32(calib.fl) * _cpu32(calib.matching_scale)).tolist()
    cs = (_cpu32(calib.principal) * _cpu32(calib.matching_scale)).tolist()
    rec = _lib.ss_camera()
    rec.fs[0], rec.fs[1] = fs
    rec.cs[0], rec.cs[1] = cs
    rec.xi = float(calib.xi)
    rec.alpha = float(calib.alpha)
    return rec


def camera_array(calibs):
    arr = (_lib.ss_camera * len(calibs))()
    for i, c in enumerate(calibs):
        arr[i] = camera_record(c)
    return arr


def relative_transforms(calibs, reference_calibration):
    """C x 12 floats: rows 0-2 of inv(rt_k) @ rt_ref, the fp32 matrix of
    depth_estimation.py:101, 156 (computed with torch on the CPU)."""
    ref_rt = _cpu32(reference_calibration.rt)
    rows = []
    for c in calibs:
        rt = torch.matmul(torch.inverse(_cpu32(c.rt)), ref_rt)
        rows.append(rt[:3].reshape(-1))
    return torch.cat(rows).tolist()


def distance_candidates(min_dist, max_dist, candidate_count):
    """1 / linspace(1/min, 1/max, N) in fp32 (depth_estimation.py:145-146)."""
    return (1 / torch.linspace(1 / min_dist, 1 / max_dist, candidate_count)).tolist()


_INPUT_TYPES = {torch.uint8: _lib.SS_INPUT_U8, torch.uint16: _lib.SS_INPUT_U16, torch.float32: _lib.SS_INPUT_F32}


def resize_linear(image, cols, rows, stream=None, out=None):
    """One image of read_input_images (utils.py:174-177, 198-201) on the GPU:
    a decoded HWC (or HW) uint8 / uint16 / float32 device tensor converted to
    float32 in [0, 255] (uint16 / 255; float32 taken as already in [0, 255])
    and resized to rows x cols as cv2.resize(..., cv2.INTER_AREA) does there
    (INTER_LINEAR, that flag being in the dst slot).  Returns a float32 tensor
    (rows, cols[, C]): `out` when given (contiguous, that shape), else new."""
    if image.dtype not in _INPUT_TYPES:
        raise ValueError("input images must be uint8, uint16 or float32")
    if image.dim() not in (2, 3) or (image.dim() == 3 and not 1 <= image.shape[2] <= 4):
        raise ValueError("input images must be HW or HWC with 1 to 4 channels")
    src = image.contiguous()
    C = 1 if src.dim() == 2 else src.shape[2]
    shape = (rows, cols) if src.dim() == 2 else (rows, cols, C)
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=src.device)
    elif tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous() or out.device != src.device:
        raise ValueError("out must be a contiguous float32 tensor of shape %s on %s" % (shape, src.device))
    st = _lib.C.c_void_p(stream.cuda_stream) if stream is not None else _lib.stream_of(out)
    _lib.check(_lib.lib.ss_resize_linear(src.data_ptr(), _INPUT_TYPES[src.dtype], src.shape[0], src.shape[1], C,
                                         out.data_ptr(), rows, cols, st))
    return out


def prepare_inputs(images, matching_resolution, rgb_to_stitch_resolution, references_indices, stream=None,
                   out=None):
    """read_input_images (utils.py:157-204) after decoding, on the GPU: the C
    decoded fisheyes (device tensors, as cv2.imread gives them: HWC uint8 /
    uint16 / float32) -> (images_to_match, images_to_stitch), float32 HWC in
    [0, 255] at the matching and stitch resolutions ((cols, rows) pairs).
    `out`, a previous result, is written in place (stable buffers let a
    captured frame graph be replayed)."""
    mc, mr = matching_resolution
    sc, sr = rgb_to_stitch_resolution
    om, os_ = out if out is not None else ([None] * len(images), [None] * len(references_indices))
    to_match = [resize_linear(img, mc, mr, stream, o) for img, o in zip(images, om)]
    to_stitch = [resize_linear(images[r], sc, sr, stream, o) for r, o in zip(references_indices, os_)]
    return to_match, to_stitch


# utils.py:253-260 of the reference lives in io.py (host file writers)
from .io import save_rgbd_panorama  # noqa: E402,F401
