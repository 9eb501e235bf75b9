"""tdvh_fast_wta_ncc_noalloc: the fused NCC + WTA without device allocation
(box sums in the library's static device array) against the context path
tdvh_fast_wta_ncc, bit-exact, on two calls in a row (the second smaller, so
it runs over the first call's leftovers in the shared array), plus the size
limit."""
import ctypes as ct

import numpy as np
import pytest
import torch

from thundervision_amd import synth

pytestmark = pytest.mark.gpu

LIMIT_PIXELS = (64 << 20) // 16


def _noalloc(tv, L, R, D, radius, zero_mean, disp):
    H, W = L.shape
    return tv.lib().tdvh_fast_wta_ncc_noalloc(
        ct.c_void_p(L.data_ptr()), ct.c_void_p(R.data_ptr()), W, H, L.stride(0), D, radius,
        int(zero_mean), ct.c_void_p(disp.data_ptr()), disp.stride(0),
        ct.c_void_p(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("first,second", [
    ((1920, 1080, 256, 4, True), (333, 97, 64, 2, False)),
    ((640, 480, 128, 1, False), (61, 17, 20, 4, True)),
])
def test_noalloc_equals_context_path_twice(tv, dev, first, second):
    for W, H, D, r, zm in (first, second):
        L, R, _ = synth.pair(W, H, D - 1, 8, W + H + D)
        Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
        want = tv.fast_wta_ncc(Ld, Rd, D, r, zm)
        got = torch.full((H, W), -1, dtype=torch.int16, device=dev)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc = _noalloc(tv, Ld, Rd, D, r, zm, got)
        assert rc == 0, tv.lib().tdvh_last_error().decode()
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0  # no device memory taken
        assert torch.equal(got, want), (W, H, D, r, zm)


def test_noalloc_rejects_images_above_the_scratch_array(tv, dev):
    L = torch.zeros((4, 4), dtype=torch.uint8, device=dev)
    disp = torch.zeros((4, 4), dtype=torch.int16, device=dev)
    W, H = 2049, 2048  # one row past 4,194,304 pixels
    assert W * H > LIMIT_PIXELS
    rc = tv.lib().tdvh_fast_wta_ncc_noalloc(
        ct.c_void_p(L.data_ptr()), ct.c_void_p(L.data_ptr()), W, H, W, 16, 4, 1,
        ct.c_void_p(disp.data_ptr()), W, ct.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -2  # TDVH_ERR_UNSUPPORTED
    assert "scratch array" in tv.lib().tdvh_last_error().decode()
    # the largest accepted image is exactly the array's size
    assert 2048 * 2048 == LIMIT_PIXELS
