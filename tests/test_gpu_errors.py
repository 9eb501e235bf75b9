"""Argument validation at the C ABI.

The reference turns a failing CUDA call into a thrown CUException
(cuerr.hpp:73-78) and checks almost no arguments. Every tdvh entry point
checks shapes, pitches, depths, dtypes, flags and pointers before it launches
anything. A bad call returns a negative status, and tdvh_last_error() names
the problem. The context keeps working afterwards: each case is followed by a
valid call on the same context."""
import ctypes as ct
import re

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

W, H, D = 64, 16, 16


@pytest.fixture(scope="module")
def imgs(dev):
    rng = np.random.default_rng(5)
    L = torch.from_numpy(rng.integers(0, 256, (H, W), dtype=np.uint8)).to(dev)
    R = torch.from_numpy(rng.integers(0, 256, (H, W), dtype=np.uint8)).to(dev)
    return L, R


def expect_error(tv, rc, pattern):
    assert rc < 0, "call was accepted"
    msg = tv.lib().tdvh_last_error().decode()
    assert re.search(pattern, msg), msg


def dsi(tv, t, width, height, depth, dtype, layout=0):
    d = tv.DSI()
    rc = tv.lib().tdvh_dsi_init(ct.byref(d), ct.c_void_p(t.data_ptr()), width, height, depth,
                                dtype, layout)
    assert rc == 0
    return d


def stream():
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def still_works(tv, imgs):
    L, R = imgs
    got = tv.stereo_bt_sgm(L, R, D).cpu().numpy()
    assert got.shape == (H, W) and got.min() >= 0 and got.max() < D


def ctx(tv):
    return tv._ctx(None)


@pytest.mark.parametrize("width,height,pitch,pattern", [
    (0, H, W, "bad image"), (W, 0, W, "bad image"), (W, H, W - 1, "bad image"),
    (-4, H, W, "bad image")])
def test_cost_bt_bad_images(tv, dev, imgs, width, height, pitch, pattern):
    L, R = imgs
    out = torch.empty((H, W, D), dtype=torch.int16, device=dev)
    d = dsi(tv, out, W, H, D, tv.U16)
    rc = tv.lib().tdvh_cost_bt(ctx(tv), ct.c_void_p(L.data_ptr()), ct.c_void_p(R.data_ptr()),
                               width, height, pitch, ct.byref(d), stream())
    expect_error(tv, rc, pattern)
    still_works(tv, imgs)


def test_cost_bt_bad_output(tv, dev, imgs):
    L, R = imgs
    out = torch.empty((H, W, D), dtype=torch.int32, device=dev)
    lp, rp = ct.c_void_p(L.data_ptr()), ct.c_void_p(R.data_ptr())
    d = dsi(tv, out, W, H, D, tv.U32)  # BT writes uint16
    expect_error(tv, tv.lib().tdvh_cost_bt(ctx(tv), lp, rp, W, H, W, ct.byref(d), stream()),
                 "uint16")
    d = dsi(tv, out, W + 1, H, D, tv.U16)  # shape differs from the images
    expect_error(tv, tv.lib().tdvh_cost_bt(ctx(tv), lp, rp, W, H, W, ct.byref(d), stream()),
                 "does not match")
    expect_error(tv, tv.lib().tdvh_cost_bt(ctx(tv), None, rp, W, H, W, ct.byref(d), stream()),
                 "null")
    still_works(tv, imgs)


@pytest.mark.parametrize("depth,radius,pattern", [(0, 3, "depth"), (513, 3, "depth"),
                                                  (16, 0, "radius"), (16, 6, "radius")])
def test_window_costs_bad_depth_radius(tv, dev, imgs, depth, radius, pattern):
    L, R = imgs
    lp, rp = ct.c_void_p(L.data_ptr()), ct.c_void_p(R.data_ptr())
    dd = max(depth, 1)
    out = torch.empty((H, W, dd), dtype=torch.float32, device=dev)
    d = dsi(tv, out, W, H, dd, tv.F32)
    d.depth = depth
    rc = tv.lib().tdvh_cost_ncc(ctx(tv), lp, rp, W, H, W, radius, 1, ct.byref(d), stream())
    expect_error(tv, rc, pattern)
    if radius == 3:  # the SSD entry point shares the image and depth checks
        out32 = torch.empty((H, W, dd), dtype=torch.int32, device=dev)
        d = dsi(tv, out32, W, H, dd, tv.U32)
        d.depth = depth
        rc = tv.lib().tdvh_cost_ssd(ctx(tv), lp, rp, W, H, W, radius, ct.byref(d), stream())
        expect_error(tv, rc, pattern)
    still_works(tv, imgs)


def test_wta_bad_arguments(tv, dev, imgs):
    S = torch.zeros((H, W, 12), dtype=torch.int16, device=dev)
    disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    dp_ = ct.c_void_p(disp.data_ptr())
    d = dsi(tv, S, W, H, 12, tv.U16)
    expect_error(tv, tv.lib().tdvh_wta(ctx(tv), ct.byref(d), 7, dp_, W, stream()), "mode")
    expect_error(tv, tv.lib().tdvh_wta(ctx(tv), ct.byref(d), 0, dp_, W - 1, stream()),
                 "disp_pitch")
    # the tree rule needs a power-of-two depth
    expect_error(tv, tv.lib().tdvh_wta(ctx(tv), ct.byref(d), tv.WTA_TREE, dp_, W, stream()),
                 "power")
    expect_error(tv, tv.lib().tdvh_wta(ctx(tv), ct.byref(d), 0, None, W, stream()), "null")
    still_works(tv, imgs)


def test_dp_bad_arguments(tv, dev, imgs):
    C = torch.zeros((H, W, 8), dtype=torch.int32, device=dev)
    disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    dp_ = ct.c_void_p(disp.data_ptr())
    d = dsi(tv, C, W, H, 8, tv.U32)
    expect_error(tv, tv.lib().tdvh_dp(ctx(tv), ct.byref(d), 20, 8, dp_, W, stream()), "flags")
    expect_error(tv, tv.lib().tdvh_dp(ctx(tv), ct.byref(d), 20, 0, dp_, W - 1, stream()),
                 "disp_pitch")
    di = dsi(tv, C, W, H, 8, tv.I32)
    expect_error(tv, tv.lib().tdvh_dp(ctx(tv), ct.byref(di), 20, 0, dp_, W, stream()),
                 "uint16, uint32 or float")
    big = torch.zeros((1, 4, 513), dtype=torch.int32, device=dev)
    db = dsi(tv, big, 4, 1, 513, tv.U32)
    expect_error(tv, tv.lib().tdvh_dp(ctx(tv), ct.byref(db), 20, 0, dp_, 4, stream()), "D <= 512")
    expect_error(tv, tv.lib().tdvh_dp_f32(ctx(tv), ct.byref(d), 0.5, 0, dp_, W, stream()),
                 "float cost volume")
    df = dsi(tv, C, W, H, 8, tv.F32)
    expect_error(tv, tv.lib().tdvh_dp_f32(ctx(tv), ct.byref(df), float("inf"), 0, dp_, W,
                                          stream()), "penalty")
    still_works(tv, imgs)


def test_sgm_wide_bad_arguments(tv, dev, imgs):
    C = torch.zeros((H, W, 8), dtype=torch.int32, device=dev)
    disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    dp_ = ct.c_void_p(disp.data_ptr())
    d32 = dsi(tv, C, W, H, 8, tv.U32)
    p = tv.sgm_params(flags=tv.SGM_REF_GEOMETRY)
    expect_error(tv, tv.lib().tdvh_sgm(ctx(tv), ct.byref(d32), ct.byref(p), None, dp_, W,
                                       stream()), "reference geometry")
    p = tv.sgm_params(fp1=-1.0)
    df = dsi(tv, C, W, H, 8, tv.F32)
    expect_error(tv, tv.lib().tdvh_sgm(ctx(tv), ct.byref(df), ct.byref(p), None, dp_, W,
                                       stream()), "float penalties")
    p = tv.sgm_params(p1=(1 << 30) + 1)
    expect_error(tv, tv.lib().tdvh_sgm(ctx(tv), ct.byref(d32), ct.byref(p), None, dp_, W,
                                       stream()), "penalties")
    S16 = torch.zeros((H, W, 8), dtype=torch.int16, device=dev)
    s16 = dsi(tv, S16, W, H, 8, tv.U16)
    p = tv.sgm_params()
    expect_error(tv, tv.lib().tdvh_sgm(ctx(tv), ct.byref(d32), ct.byref(p), ct.byref(s16), dp_,
                                       W, stream()), "S_out")
    still_works(tv, imgs)


def test_batch_rejects_large_p1(tv, dev, imgs):
    L, R = imgs
    disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    with pytest.raises(tv.TdvhError, match="P1"):
        tv.Batch([(L, R, D, disp)], p1=70000)
    still_works(tv, imgs)


@pytest.mark.parametrize("p1,p2,mask,pattern", [(-1, 32, 0xFF, "penalties"),
                                                (8, 2000, 0xFF, "penalties"),
                                                (8, 32, 0, "dir_mask"),
                                                (8, 32, 0x1FF, "dir_mask")])
def test_sgm_bad_params(tv, dev, imgs, p1, p2, mask, pattern):
    L, R = imgs
    disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    params = tv.sgm_params(p1, p2, mask)
    rc = tv.lib().tdvh_stereo_bt_sgm(ctx(tv), ct.c_void_p(L.data_ptr()),
                                     ct.c_void_p(R.data_ptr()), W, H, W, D, ct.byref(params),
                                     tv.FUSE_COST, ct.c_void_p(disp.data_ptr()), W, stream())
    expect_error(tv, rc, pattern)
    still_works(tv, imgs)


@pytest.mark.parametrize("depth,fusion,flags,pattern", [
    (0, 1, 0, "depth"), (257, 1, 0, "depth"), (16, 5, 0, "fusion"),
    (16, 1, 1, "reference geometry")])
def test_fused_sgm_bad_call(tv, dev, imgs, depth, fusion, flags, pattern):
    L, R = imgs
    disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    params = tv.sgm_params(flags=flags)
    rc = tv.lib().tdvh_stereo_bt_sgm(ctx(tv), ct.c_void_p(L.data_ptr()),
                                     ct.c_void_p(R.data_ptr()), W, H, W, depth,
                                     ct.byref(params), fusion,
                                     ct.c_void_p(disp.data_ptr()), W, stream())
    expect_error(tv, rc, pattern)
    still_works(tv, imgs)


def test_fast_wta_ssd_bad_output(tv, dev, imgs):
    L, R = imgs
    disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    rc = tv.lib().tdvh_fast_wta_ssd(ctx(tv), ct.c_void_p(L.data_ptr()), ct.c_void_p(R.data_ptr()),
                                    W, H, W, D, 3, ct.c_void_p(disp.data_ptr()), W - 1, stream())
    expect_error(tv, rc, "disparity output")
    # the fused matchers keep one pass of at most 256 disparities
    rc = tv.lib().tdvh_fast_wta_ssd(ctx(tv), ct.c_void_p(L.data_ptr()), ct.c_void_p(R.data_ptr()),
                                    W, H, W, 257, 3, ct.c_void_p(disp.data_ptr()), W, stream())
    expect_error(tv, rc, "depth")
    still_works(tv, imgs)
