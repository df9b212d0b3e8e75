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


class ss_camera(C.Structure):
    _fields_ = [("fs", C.c_float * 2), ("cs", C.c_float * 2), ("xi", C.c_double), ("alpha", C.c_double)]


class ss_intrinsics(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float), ("xi", C.c_float),
                ("alpha", C.c_float)]


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
    "ss_sweep": (_i, [_p, _p, _i, _CAM, _p, _i, _p, _i, _i, _i, _p, _p]),
    "ss_isb_workspace_bytes": (_z, [_i, _i, _i, _i]),
    "ss_isb_filter": (_i, [_p, _p, _i, _i, _i, _d, _d, _p, _z, _p]),
    "ss_isb_filter_planar": (_i, [_p, _p, _i, _i, _i, _d, _d, _p, _z, _p]),
    "ss_argmin_refine": (_i, [_p, _i, _i, _i, _fl, _fl, _fl, _p, _p]),
    "ss_estimate_workspace_bytes": (_z, [_i, _i, _i]),
    "ss_estimate_distance": (_i, [_p, _p, _i, _CAM, _p, _i, _p, _i, _i, _i, _d, _d, _p, _p, _p, _z, _p]),
    "ss_reproject": (_i, [_p, _p, _i, _i, _INTR, _p, _p]),
    "ss_inpaint": (_i, [_p, _p, _p, _i, _i, _i, _d, _d, _p]),
    "ss_merge": (_i, [_p, _p, _p, _p, _p, _i, _i, _p, _INTR, _i, _i, _i, _i, _i, _p, _p, _p]),
    "ss_to_u8": (_i, [_p, _p, _z, _p]),
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
                          "python -m sphere_stereo._build" % LIB_PATH)
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
