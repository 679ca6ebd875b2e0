"""The allocation-free entry point (ss_extrema_u8_noalloc), through the C ABI:
two calls in a row on one GPU (the second on a smaller image, so the library's
scratch array holds the first call's state when the second starts), each
bit-identical to ss_extrema_u8 on a caller workspace; images whose workspace
exceeds the scratch array are refused."""
import ctypes

import pytest
import torch

from ssx import sparse_seam as ss
from ssx import ssgen

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _noalloc():
    lib = ss.lib()
    fn = lib.ss_extrema_u8_noalloc
    vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    fn.argtypes = [vp, sz, i, i, vp, vp, sz, ctypes.POINTER(ss.Config), vp, vp]
    fn.restype = i
    return fn


def run_noalloc(img, tile, threads):
    h, w = img.shape
    mx = torch.full((h, w), 7, dtype=torch.uint8, device=img.device)
    mn = torch.full((h, w), 7, dtype=torch.uint8, device=img.device)
    inv = torch.full((1,), 0x5A5A, dtype=torch.int32, device=img.device)
    cfg = ss.Config(tile, threads)
    st = _noalloc()(img.data_ptr(), img.stride(0), w, h, mx.data_ptr(), mn.data_ptr(), mx.stride(0),
                    ctypes.byref(cfg), inv.data_ptr(), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, mx, mn, int(inv.item())


@pytest.mark.parametrize("tile,threads", [(64, 256), (32, 256), (16, 128)])
def test_noalloc_equals_workspace_entry_point(tile, threads):
    first = ssgen.smooth(4000, 3900, 4.0, 8, 5, device="cuda")       # 53 MiB of workspace at T = 16
    second = ssgen.smooth(700, 1100, 1.0, 4, 6, device="cuda")       # smaller: runs on the first call's leftovers
    assert ss.workspace_size(3900, 4000, tile, threads) <= SCRATCH_BYTES
    for img in (first, second):
        print("noalloc tile=%d threads=%d %s" % (tile, threads, tuple(img.shape)), flush=True)
        st, mx, mn, inv = run_noalloc(img, tile, threads)
        assert st == ss.SS_OK and inv == 0
        rmx, rmn, ws = ss.extrema(img, tile, threads)
        torch.cuda.synchronize()
        assert ss.stats(ws)["invariant"] == 0
        assert torch.equal(mx, rmx) and torch.equal(mn, rmn)


def test_noalloc_refuses_oversized_image():
    side = 6400
    assert ss.workspace_size(side, side, 64, 256) > SCRATCH_BYTES
    img = torch.zeros((side, side), dtype=torch.uint8, device="cuda")
    st, mx, mn, inv = run_noalloc(img, 64, 256)
    assert st == ss.SS_ERR_CAPACITY
    assert inv == 0x5A5A and (mx == 7).all()     # nothing was launched
    # and the scratch array still serves a small call afterwards
    small = ssgen.checker(300, 200, 3, device="cuda")
    st, mx, mn, inv = run_noalloc(small, 64, 256)
    rmx, rmn, _ = ss.extrema(small, 64, 256)
    torch.cuda.synchronize()
    assert st == ss.SS_OK and inv == 0 and torch.equal(mx, rmx) and torch.equal(mn, rmn)
