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
    assert lib.ss_frame_set_lanes(None, 2) == 1 and b"null frame" in lib.ss_last_error()
    assert lib.ss_frame_get_lanes(None) == -1


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


def _fake_desc(H, W, N, R, vt, iters=32):
    """A frame descriptor with placeholder device addresses (the table sizing
    is host arithmetic: it never reads them)."""
    lib = _lib.lib
    d = _lib.ss_frame_desc()
    d.C, d.R, d.N, d.H, d.W = 4, R, N, H, W
    d.Hs, d.Ws, d.PH, d.PW = 96, 96, 64, 128
    d.inpaint_iterations = iters
    d.volume_type = vt
    d.workspace_bytes = lib.ss_estimate_workspace_bytes(H, W, N, vt)
    addr = 0x100000
    for r in range(R):
        d.refs[r] = r
        for arr in (d.d_raylut, d.d_workspace, d.d_dist, d.d_inpaint_lut, d.d_reproj, d.d_scratch, d.d_stitch_u8):
            arr[r] = addr
            addr += 0x100000
    for n in range(N):
        d.cands[n] = 1.0 / (1 / 0.55 + (1 / 100 - 1 / 0.55) * n / max(N - 1, 1))
    d.d_sampling = d.d_weights = d.d_dist_pano = d.d_rgb_pano = addr
    return d


def test_ragged_table_sizing_and_rejections():
    """ss_frame_ragged_table_bytes: the plan of a ragged batch (host logic,
    no device memory touched) grows with the batch and is refused for mixed
    volume types, one candidate and resolutions without two pyramid levels."""
    lib = _lib.lib
    F32, F16 = _lib.SS_VOLUME_F32, _lib.SS_VOLUME_F16

    def size(*descs):
        arr = (_lib.ss_frame_desc * len(descs))(*descs)
        return lib.ss_frame_ragged_table_bytes(arr, len(descs))

    a = _fake_desc(512, 512, 32, 2, F16)
    b = _fake_desc(768, 768, 48, 4, F16, iters=16)
    c = _fake_desc(96, 80, 16, 3, F16, iters=0)
    one = size(a)
    three = size(a, b, c)
    assert one > 0 and three > one
    assert size(a, b, c, a, b, c) > three
    assert size(a, _fake_desc(512, 512, 32, 2, F32)) == 0      # volume types differ
    assert size(_fake_desc(512, 512, 1, 2, F16)) == 0          # one candidate
    assert size(_fake_desc(6, 6, 32, 2, F16)) == 0             # a single pyramid level
    assert size(_fake_desc(8, 8, 32, 2, F16)) > 0
    assert lib.ss_frame_ragged_table_bytes(None, 1) == 0


def test_sweep_tuning_json_roundtrip(tmp_path):
    """save_sweep_tuning / load_sweep_tuning: the sweep and cost-filter
    variants tuned in a process go to JSON and back into the library's
    per-configuration registries (ss_sweep_set_variant / get_variant,
    ss_isb_set_variant / get_variant, host state only)."""
    import json

    from sphere_stereo import depth_estimation as de

    lib = _lib.lib
    saved, saved_f = dict(de._TUNED), dict(de._TUNED_FILTER)
    try:
        de._TUNED.clear()
        de._TUNED_FILTER.clear()
        de._TUNED[(96, 80, 48, _lib.SS_VOLUME_F16)] = 3
        de._TUNED[(64, 64, 16, _lib.SS_VOLUME_F32)] = 1
        de._TUNED_FILTER[(96, 80, 48, _lib.SS_VOLUME_F16)] = 2
        saved_i = dict(de._TUNED_INPAINT)
        de._TUNED_INPAINT.clear()
        de._TUNED_INPAINT[(96, 80)] = 1
        path = str(tmp_path / "tuning.json")
        assert de.save_sweep_tuning(path) == 4
        data = json.load(open(path))
        assert {e["variant"] for e in data["sweep_variants"]} == {1, 3}
        assert [e["variant"] for e in data["filter_variants"]] == [2]
        assert lib.ss_sweep_get_variant(96, 80, 48, _lib.SS_VOLUME_F16) == 0
        assert lib.ss_isb_get_variant(96, 80, 48, _lib.SS_VOLUME_F16) == 0
        assert lib.ss_inpaint_get_variant(96, 80) == 0
        assert de.load_sweep_tuning(path) == 4
        assert lib.ss_inpaint_get_variant(96, 80) == 1
        assert lib.ss_sweep_get_variant(96, 80, 48, _lib.SS_VOLUME_F16) == 3
        assert lib.ss_sweep_get_variant(64, 64, 16, _lib.SS_VOLUME_F32) == 1
        assert lib.ss_isb_get_variant(96, 80, 48, _lib.SS_VOLUME_F16) == 2
        data["device"] = "another GPU"
        json.dump(data, open(path, "w"))
        assert de.load_sweep_tuning(path) == 0  # tuned elsewhere: not applied
    finally:
        for H, W, N, vt in [(96, 80, 48, _lib.SS_VOLUME_F16), (64, 64, 16, _lib.SS_VOLUME_F32)]:
            lib.ss_sweep_set_variant(H, W, N, vt, -1)
            lib.ss_isb_set_variant(H, W, N, vt, -1)
        de._TUNED.clear()
        de._TUNED.update(saved)
        de._TUNED_FILTER.clear()
        de._TUNED_FILTER.update(saved_f)
        lib.ss_inpaint_set_variant(96, 80, -1)
        de._TUNED_INPAINT.clear()
        de._TUNED_INPAINT.update(saved_i)


def test_isb_variants():
    """The cost filter's scheduling variants: (lanes, tile order) per variant,
    groups of 4 lanes only for padded candidate counts up to 64."""
    lib = _lib.lib
    assert lib.ss_isb_variant_count() == 4
    lanes, xcd = C.c_int(), C.c_int()
    got = []
    for v in range(4):
        assert lib.ss_isb_variant_info(v, C.byref(lanes), C.byref(xcd)) == 0
        got.append((lanes.value, xcd.value))
    assert got == [(8, 1), (8, 0), (4, 1), (4, 0)]
    assert lib.ss_isb_variant_info(4, C.byref(lanes), C.byref(xcd)) == 1
    assert lib.ss_isb_set_variant(64, 64, 128, 0, 2) == 1 and b"does not exist" in lib.ss_last_error()
    assert lib.ss_isb_set_variant(64, 64, 64, 0, 3) == 0
    assert lib.ss_isb_get_variant(64, 64, 64, 0) == 3
    assert lib.ss_isb_set_variant(64, 64, 64, 0, -1) == 0
    assert lib.ss_isb_get_variant(64, 64, 64, 0) == 0
    tile = C.c_int()
    assert lib.ss_inpaint_variant_count() == 3
    assert [lib.ss_inpaint_variant_info(v, C.byref(tile)) == 0 and tile.value for v in range(3)] == [48, 32, 16]
    assert lib.ss_inpaint_set_variant(100, 90, 3) == 1


def test_isb_noalloc_limit_and_refusal():
    """ss_isb_filter_noalloc's static scratch: the header's limit is the one
    the bindings state, the C1-sized filter fits it, and a call that needs
    more is refused before any device work."""
    text = open(HEADER).read()
    m = re.search(r"#define SS_ISB_NOALLOC_SCRATCH_BYTES \((\d+)u << (\d+)\)", text)
    assert m and int(m.group(1)) << int(m.group(2)) == _lib.SS_ISB_NOALLOC_SCRATCH_BYTES
    lib = _lib.lib
    assert lib.ss_isb_workspace_bytes(1024, 1024, 32, _lib.SS_VOLUME_F32, 0) <= _lib.SS_ISB_NOALLOC_SCRATCH_BYTES
    need = lib.ss_isb_workspace_bytes(1024, 1024, 64, _lib.SS_VOLUME_F32, 0)
    assert need > _lib.SS_ISB_NOALLOC_SCRATCH_BYTES
    one = C.c_void_p(256)  # never dereferenced: the size check comes first
    rc = lib.ss_isb_filter_noalloc(one, one, 1024, 1024, 64, _lib.SS_VOLUME_F32, 10.0, 25.0, None)
    assert rc == 4 and b"static" in lib.ss_last_error()
    rc = lib.ss_isb_filter_noalloc(one, one, 3, 3, 8, _lib.SS_VOLUME_F32, 10.0, 25.0, None)
    assert rc == 1 and b"pyramid level" in lib.ss_last_error()
