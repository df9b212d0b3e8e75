"""ss_isb_filter_noalloc, the ISB filter that allocates nothing (its scratch
is a static device array per GPU), against ss_isb_filter with a workspace,
through the C ABI: two calls in a row, the second smaller (so its scratch
still holds the first call's levels), each bit for bit the workspace entry
point's result.  (The refusal of a call beyond the static scratch is checked
without a GPU in test_abi.py.)"""
import ctypes as C

import pytest
import torch

from sphere_stereo import _lib

pytestmark = pytest.mark.gpu

VT = {torch.float32: _lib.SS_VOLUME_F32, torch.float16: _lib.SS_VOLUME_F16, torch.bfloat16: _lib.SS_VOLUME_BF16}
BITS = {torch.float32: torch.int32, torch.float16: torch.int16, torch.bfloat16: torch.int16}


def _inputs(H, W, N, dtype, seed, device):
    NP = _lib.lib.ss_padded_candidates(N)
    g = torch.Generator().manual_seed(seed)
    guide = torch.randint(0, 1 << 24, [H, W], generator=g, dtype=torch.int32)  # packed bytes y, cb, cr, 0
    guide[: H // 2] = (guide[: H // 2] & 0x0F0F0F) + 0x646464  # a smoother half: the edge weights vary
    cost = torch.rand([H, W, NP], generator=g) * 500
    cost[..., N:] = 0
    return guide.to(device), cost.to(dtype).to(device)


def _both(gpu, H, W, N, dtype, seed):
    lib = _lib.lib
    print("no-alloc ISB filter: %dx%d, %d candidates, %s" % (W, H, N, dtype), flush=True)
    stream = C.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    vt = VT[dtype]
    ga, ca = _inputs(H, W, N, dtype, seed, gpu)
    gb, cb = ga.clone(), ca.clone()
    nbytes = lib.ss_isb_workspace_bytes(H, W, N, vt, 0)
    assert 0 < nbytes <= _lib.SS_ISB_NOALLOC_SCRATCH_BYTES
    ws = torch.empty(nbytes, dtype=torch.uint8, device=gpu)
    _lib.check(lib.ss_isb_filter(ga.data_ptr(), ca.data_ptr(), H, W, N, vt, 10.0, 25.0, ws.data_ptr(), nbytes,
                                 stream))
    _lib.check(lib.ss_isb_filter_noalloc(gb.data_ptr(), cb.data_ptr(), H, W, N, vt, 10.0, 25.0, stream))
    torch.cuda.synchronize()
    assert not torch.equal(ca.view(BITS[dtype]), _inputs(H, W, N, dtype, seed, gpu)[1].view(BITS[dtype]))
    assert torch.equal(gb, ga), "guide"
    assert torch.equal(cb.view(BITS[dtype]), ca.view(BITS[dtype])), "filtered volume"


def test_noalloc_matches_workspace_entry_point_twice(gpu):
    _both(gpu, 1024, 1024, 32, torch.float32, 1)  # 44 MiB of scratch
    _both(gpu, 200, 300, 48, torch.float16, 2)  # smaller, over the first call's leftovers
    _both(gpu, 37, 61, 1, torch.float32, 3)  # the distance filter's one-candidate form

