"""The C ABI without a GPU: every declared symbol is exported, the pure queries
answer, and argument errors are reported before any device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from sphere_stereo import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "sphere_stereo.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ss_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    names = declared()
    assert len(names) >= 25
    lib = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, "no ctypes signature for %s" % n


def test_queries():
    lib = _lib.lib
    assert lib.ss_abi_version() == 2
    assert [lib.ss_padded_candidates(n) for n in (1, 2, 8, 9, 32, 33, 48, 64, 100, 128, 129)] == \
        [1, 8, 8, 16, 32, 48, 48, 64, 128, 128, -1]
    assert lib.ss_isb_levels(1024, 1024) == 9 and lib.ss_isb_levels(768, 768) == 8
    assert lib.ss_isb_levels(3, 3) == 0 and lib.ss_isb_levels(4, 1000) == 1
    assert lib.ss_isb_workspace_bytes(3, 3, 4, 0, 0) == 0
    f32 = lib.ss_estimate_workspace_bytes(1024, 1024, 32, _lib.SS_VOLUME_F32)
    f16 = lib.ss_estimate_workspace_bytes(1024, 1024, 32, _lib.SS_VOLUME_F16)
    # the level-0 volume in 16 bits, the coarser levels (1/3 of it) in fp32
    l0 = 1024 * 1024 * 32
    assert f32 > l0 * 4 and l0 * 2 < f16 < l0 * 4 and f32 - f16 >= l0 * 2
    assert lib.ss_estimate_workspace_bytes(1024, 1024, 32, _lib.SS_VOLUME_BF16) == f16
    assert lib.ss_estimate_workspace_bytes(1024, 1024, 32, 7) == 0
    # the planar drop-in is fp32 only
    assert lib.ss_isb_workspace_bytes(64, 64, 8, _lib.SS_VOLUME_F16, 1) == 0
    # the library-layout workspace holds only the (fp32) coarser levels
    assert lib.ss_isb_workspace_bytes(64, 64, 8, _lib.SS_VOLUME_F16, 0) == lib.ss_isb_workspace_bytes(64, 64, 8, 0, 0)
    assert lib.ss_stage_name(1) == b"sweep" and lib.ss_stage_name(99) == b""


def test_argument_errors_before_device_work():
    lib = _lib.lib
    cams = (_lib.ss_camera * 4)()
    rts = _lib.float_array([0.0] * 48)
    cands = _lib.float_array([1.0] * 8)
    imgs = (C.c_void_p * 4)(1, 2, 3, 4)
    one = C.c_void_p(1)
    rc = lib.ss_sweep(one, imgs, 0, cams, rts, 0, cands, 8, 16, 16, 0, one, None, None)
    assert rc == 1 and b"camera count" in lib.ss_last_error()
    rc = lib.ss_sweep(one, imgs, 4, cams, rts, 5, cands, 8, 16, 16, 0, one, None, None)
    assert rc == 1 and b"reference index" in lib.ss_last_error()
    rc = lib.ss_sweep(one, imgs, 4, cams, rts, 0, cands, 200, 16, 16, 0, one, None, None)
    assert rc == 1 and b"candidate count" in lib.ss_last_error()
    rc = lib.ss_sweep(one, imgs, 4, cams, rts, 0, cands, 8, 16, 16, 5, one, None, None)
    assert rc == 1 and b"volume type" in lib.ss_last_error()
    rc = lib.ss_sweep(one, imgs, 4, cams, rts, 0, cands, 1, 16, 16, 0, one, None, None)
    assert rc == 1 and b"at least 2" in lib.ss_last_error()
    rc = lib.ss_estimate_distance(one, imgs, 4, cams, rts, 0, cands, 8, 16, 16, 10.0, 25.0, 0, None, one, one, 16,
                                  None)
    assert rc == 4 and b"workspace" in lib.ss_last_error()
    rc = lib.ss_estimate_distance(one, imgs, 4, cams, rts, 0, cands, 8, 3, 3, 10.0, 25.0, 0, None, one, one,
                                  1 << 20, None)
    assert rc == 1 and b"pyramid" in lib.ss_last_error()
    rc = lib.ss_isb_filter(one, one, 16, 16, 1, _lib.SS_VOLUME_F16, 10.0, 25.0, one, 1 << 20, None)
    assert rc == 1 and b"fp32 only" in lib.ss_last_error()
    assert lib.ss_box_filter(C.c_void_p(8), C.c_void_p(8), C.c_void_p(16), 1, 4, 4, 1, 0.0, None) == 1
    assert lib.ss_isb_filter_planar(C.c_void_p(8), C.c_void_p(16), 16, 16, 4, 10.0, 25.0, None, 0, None) == 4
    assert lib.ss_merge(None, None, None, None, None, 8, 8, None, None, 1, 8, 8, 4, 4, None, None, None) == 1
    assert lib.ss_profile_collect(None, None) == 1


def test_inpaint_thresholds_follow_constant_substitution():
    from sphere_stereo.stitcher import inpaint_thresholds

    te, ts = inpaint_thresholds(100)  # "100 + 0.1f" is an fp32 sum
    assert te == 100.1 and ts == float(np.float32(100) + np.float32(0.1))
    te, ts = inpaint_thresholds(100.0)  # "100.0 + 0.1f" is a double sum
    assert te == 100.1 and ts == 100.0 + float(np.float32(0.1))
    assert ts != inpaint_thresholds(100)[1]


def test_candidates_and_transforms():
    import torch

    from sphere_stereo.utils import Calibration, camera_record, distance_candidates, relative_transforms

    c = distance_candidates(0.55, 100, 32)
    want = (1 / torch.linspace(1 / 0.55, 1 / 100, 32)).tolist()
    assert c == want and len(c) == 32 and abs(c[0] - 0.55) < 1e-6 and abs(c[-1] - 100) < 1e-3
    rt = torch.eye(4)
    rt[:3, 3] = torch.tensor([0.1, 0.0, -0.2])
    cal = Calibration([1216, 1216], torch.tensor([608.0, 600.0]), torch.tensor([364.8, 360.0]), -0.18, 0.59, rt,
                      torch.tensor([1024 / 1216, 1024 / 1216]))
    rec = camera_record(cal)
    assert rec.fs[0] == float(np.float32(364.8) * np.float32(1024 / 1216))
    assert rec.xi == -0.18 and rec.alpha == 0.59
    rows = relative_transforms([cal, cal], cal)
    assert len(rows) == 24 and rows[:12] == [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0]


def test_isb_filter_rejects_tiny_resolution_on_cpu():
    from sphere_stereo import ISB_Filter

    with pytest.raises(ValueError):
        ISB_Filter(4, (3, 3), "cpu")


def test_frame_desc_layout_matches_header(tmp_path):
    import subprocess

    src = tmp_path / "layout.c"
    fields = [f[0] for f in _lib.ss_frame_desc._fields_]
    body = "\n".join('printf("%s %%zu\\n", offsetof(ss_frame_desc, %s));' % (f, f) for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "sphere_stereo.h"\nint main(void){\n'
                   'printf("size %%zu\\n", sizeof(ss_frame_desc));\nprintf("camera %%zu\\n", sizeof(ss_camera));\n'
                   '%s\nreturn 0;}\n' % body)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    assert int(out["size"]) == C.sizeof(_lib.ss_frame_desc)
    assert int(out["camera"]) == C.sizeof(_lib.ss_camera)
    for f in fields:
        assert int(out[f]) == getattr(_lib.ss_frame_desc, f).offset, f
