"""ctypes binding of libsphere_stereo.so (declarations: include/sphere_stereo.h).

The library is mandatory: importing the package on a machine without the built
library raises, and there is no CPU fallback for any stage.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphere_stereo.so")

SS_MAX_CAMERAS = 8
SS_MAX_CANDIDATES = 128
SS_MAX_REFERENCES = 8
SS_MAX_BATCH_FRAMES = 64
SS_BATCH_PAIRS = 16  # frame-reference pairs per launch of a native batch (chunks of ss_frame_batch_run)
SS_RAGGED_PERSISTENT, SS_RAGGED_DISPATCHED, SS_RAGGED_AUTO = 0, 1, 2
SS_INPUT_U8, SS_INPUT_U16, SS_INPUT_F32 = 0, 1, 2

# ss_volume_type: element type of the stored cost volume
SS_VOLUME_F32, SS_VOLUME_F16, SS_VOLUME_BF16 = 0, 1, 2
SS_ISB_NOALLOC_SCRATCH_BYTES = 48 << 20  # static scratch of ss_isb_filter_noalloc, per GPU
VOLUME_TYPES = {"float32": SS_VOLUME_F32, "float16": SS_VOLUME_F16, "bfloat16": SS_VOLUME_BF16}


class ss_camera(C.Structure):
    _fields_ = [("fs", C.c_float * 2), ("cs", C.c_float * 2), ("xi", C.c_double), ("alpha", C.c_double)]


class ss_intrinsics(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float), ("xi", C.c_float),
                ("alpha", C.c_float)]


class ss_frame_desc(C.Structure):
    _fields_ = [("C", C.c_int), ("R", C.c_int), ("N", C.c_int), ("H", C.c_int), ("W", C.c_int),
                ("Hs", C.c_int), ("Ws", C.c_int), ("PH", C.c_int), ("PW", C.c_int), ("inpaint_iterations", C.c_int),
                ("volume_type", C.c_int), ("refs", C.c_int * SS_MAX_REFERENCES), ("cams", ss_camera * SS_MAX_CAMERAS),
                ("rts", (C.c_float * (SS_MAX_CAMERAS * 12)) * SS_MAX_REFERENCES),
                ("cands", C.c_float * SS_MAX_CANDIDATES), ("sigma_i", C.c_double), ("sigma_s", C.c_double),
                ("thr_empty", C.c_double), ("thr_src", C.c_double),
                ("d_raylut", C.c_void_p * SS_MAX_REFERENCES), ("d_workspace", C.c_void_p * SS_MAX_REFERENCES),
                ("workspace_bytes", C.c_size_t), ("d_dist", C.c_void_p * SS_MAX_REFERENCES),
                ("intr", ss_intrinsics * SS_MAX_REFERENCES), ("trans", (C.c_float * 3) * SS_MAX_REFERENCES),
                ("d_inpaint_lut", C.c_void_p * SS_MAX_REFERENCES), ("d_reproj", C.c_void_p * SS_MAX_REFERENCES),
                ("d_scratch", C.c_void_p * SS_MAX_REFERENCES), ("d_stitch_u8", C.c_void_p * SS_MAX_REFERENCES),
                ("d_sampling", C.c_void_p), ("d_weights", C.c_void_p), ("d_dist_pano", C.c_void_p),
                ("d_rgb_pano", C.c_void_p)]


_i, _d, _p, _z, _fl = C.c_int, C.c_double, C.c_void_p, C.c_size_t, C.c_float
_CAM = C.POINTER(ss_camera)
_INTR = C.POINTER(ss_intrinsics)

# name -> (restype, argtypes)
SIGNATURES = {
    "ss_abi_version": (_i, []),
    "ss_last_error": (C.c_char_p, []),
    "ss_padded_candidates": (_i, [_i]),
    "ss_isb_levels": (_i, [_i, _i]),
    "ss_select_camera": (_i, [_CAM, _p, _i, _i, _p, _i, _i, _d, _d, _p, _p]),
    "ss_ray_lut": (_i, [_CAM, _p, _i, _i, _p, _p]),
    "ss_inpaint_lut": (_i, [_INTR, _p, _i, _i, _d, _d, _p, _p]),
    "ss_box_filter": (_i, [_p, _p, _p, _i, _i, _i, _i, _fl, _p]),
    "ss_blend_lut": (_i, [_INTR, _p, _p, _i, _p, _i, _i, _i, _i, _d, _d, _p, _p, _p]),
    "ss_normalize_weights": (_i, [_p, _i, _i, _p]),
    "ss_guide": (_i, [_p, _i, _i, _p, _p]),
    "ss_sweep": (_i, [_p, _p, _i, _CAM, _p, _i, _p, _i, _i, _i, _i, _p, _p, _p]),
    "ss_sweep_variant_count": (_i, []),
    "ss_sweep_variant_info": (_i, [_i, _p, _p]),
    "ss_sweep_set_variant": (_i, [_i, _i, _i, _i, _i]),
    "ss_sweep_get_variant": (_i, [_i, _i, _i, _i]),
    "ss_sweep_autotune": (_i, [_p, _p, _i, _CAM, _p, _i, _p, _i, _i, _i, _i, _p, _i, _p, _p, _p]),
    "ss_isb_workspace_bytes": (_z, [_i, _i, _i, _i, _i]),
    "ss_isb_filter": (_i, [_p, _p, _i, _i, _i, _i, _d, _d, _p, _z, _p]),
    "ss_isb_filter_planar": (_i, [_p, _p, _i, _i, _i, _d, _d, _p, _z, _p]),
    "ss_isb_filter_noalloc": (_i, [_p, _p, _i, _i, _i, _i, _d, _d, _p]),
    "ss_isb_variant_count": (_i, []),
    "ss_inpaint_variant_count": (_i, []),
    "ss_inpaint_variant_info": (_i, [_i, _p]),
    "ss_inpaint_set_variant": (_i, [_i, _i, _i]),
    "ss_inpaint_get_variant": (_i, [_i, _i]),
    "ss_inpaint_autotune": (_i, [_p, _i, _p, _p, _p, _i, _i, _i, _d, _d, _i, _p, _p, _p]),
    "ss_isb_variant_info": (_i, [_i, _p, _p]),
    "ss_isb_set_variant": (_i, [_i, _i, _i, _i, _i]),
    "ss_isb_get_variant": (_i, [_i, _i, _i, _i]),
    "ss_estimate_filter_autotune": (_i, [_p, _z, _i, _i, _i, _i, _p, _d, _d, _p, _i, _p, _p, _p]),
    "ss_argmin_refine": (_i, [_p, _i, _i, _i, _i, _fl, _fl, _fl, _p, _p]),
    "ss_estimate_workspace_bytes": (_z, [_i, _i, _i, _i]),
    "ss_estimate_distance": (_i, [_p, _p, _i, _CAM, _p, _i, _p, _i, _i, _i, _d, _d, _i, _p, _p, _p, _z, _p]),
    "ss_reproject": (_i, [_p, _p, _i, _i, _INTR, _p, _p]),
    "ss_inpaint": (_i, [_p, _p, _p, _i, _i, _i, _d, _d, _p]),
    "ss_merge": (_i, [_p, _p, _p, _p, _p, _i, _i, _p, _INTR, _i, _i, _i, _i, _i, _p, _p, _p]),
    "ss_to_u8": (_i, [_p, _p, _z, _p]),
    "ss_inverse_distance": (_i, [_p, _p, _z, _p]),
    "ss_error_stats_workspace_bytes": (_z, [_z]),
    "ss_error_stats": (_i, [_p, _p, _z, _p, _i, _i, _p, _p, _z, _p]),
    "ss_resize_linear": (_i, [_p, _i, _i, _i, _i, _p, _i, _i, _p]),
    "ss_frame_create": (_i, [C.POINTER(ss_frame_desc), C.POINTER(C.c_void_p)]),
    "ss_frame_run": (_i, [_p, _p, _p, _i, _p]),
    "ss_frame_batch_create": (_i, [C.POINTER(ss_frame_desc), _i, C.POINTER(C.c_void_p)]),
    "ss_frame_batch_run": (_i, [_p, _p, _p, _i, _p]),
    "ss_frame_ragged_table_bytes": (_z, [C.POINTER(ss_frame_desc), _i]),
    "ss_frame_ragged_create": (_i, [C.POINTER(ss_frame_desc), _i, _i, _p, _z, _p, C.POINTER(C.c_void_p)]),
    "ss_frame_destroy": (_i, [_p]),
    "ss_frame_set_lanes": (_i, [_p, _i]),
    "ss_frame_get_lanes": (_i, [_p]),
    "ss_soft_argmin": (_i, [_p, _i, _i, _i, _fl, _fl, _fl, _p, _p]),
    "ss_soft_argmin_backward": (_i, [_p, _p, _i, _i, _i, _fl, _fl, _fl, _p, _p]),
    "ss_isb_backward_workspace_bytes": (_z, [_i, _i, _i]),
    "ss_isb_filter_backward": (_i, [_p, _p, _p, _i, _i, _i, _d, _d, _p, _z, _p]),
    "ss_sweep_backward_workspace_bytes": (_z, [_i, _i, _i]),
    "ss_sweep_backward": (_i, [_p, _p, _i, _CAM, _p, _i, _p, _i, _i, _i, _p, _p, _p, _p, _z, _p]),
    "ss_photometric_blocks": (_i, [_i, _i]),
    "ss_photometric_record_floats": (_i, []),
    "ss_photometric_normal_equations": (_i, [_p, _p, _p, _p, _p, _CAM, _p, _i, _i, _i, _i, _fl, _p, _p]),
    "ss_gauss_newton_solve": (_i, [_p, _i, _d, _p, _p]),
    "ss_profile_enable": (_i, [_i, C.c_uint]),
    "ss_profile_collect": (_i, [_p, _p]),
    "ss_stage_name": (C.c_char_p, [_i]),
}


class SphereStereoError(RuntimeError):
    pass


def _load():
    # torch first: it brings its own HIP runtime (libamdhip64.so.7), and the
    # library must bind to that same runtime by soname.  Loading the library
    # first would map /opt/rocm's runtime as well, and two HIP runtimes in one
    # process do not share devices, streams or allocations.
    import torch  # noqa: F401

    if not os.path.exists(LIB_PATH):
        raise ImportError("libsphere_stereo.so is not built (%s); run __graft_entry__.build() or "
                          "python sphere-stereo_amd/sphere_stereo/_build.py" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    if rc != 0:
        raise SphereStereoError("sphere_stereo error %d: %s" % (rc, lib.ss_last_error().decode()))


def ptr_array(tensors):
    arr = (C.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = t.data_ptr()
    return arr


def float_array(values):
    vals = [float(v) for v in values]
    return (C.c_float * len(vals))(*vals)


def stream_of(tensor):
    """hipStream_t of torch's current stream on the tensor's device."""
    import torch

    return C.c_void_p(torch.cuda.current_stream(tensor.device).cuda_stream)
