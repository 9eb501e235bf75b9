"""GPU parity of the steps around the path (prefilter, postfilter, colour
conversion, remap, reprojection, transfers) against the oracle, bit-exact.
Calls go through the C ABI (libtdvh.so via ctypes)."""
import ctypes as ct

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (1, 7), (9, 1), (37, 23), (64, 32), (65, 33), (300, 41), (1280, 720)]


def rng(seed):
    return np.random.default_rng(seed)


def dev_t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("W,H", SIZES)
def test_prefilter_gauss5(tv, dev, W, H):
    img = rng(W * 7 + H).integers(0, 256, (H, W), dtype=np.uint8)
    got = tv.prefilter_gauss5(dev_t(img, dev)).cpu().numpy()
    assert (got == oracle.gauss5(img)).all()


def test_prefilter_pitched_and_rejects_in_place(tv, dev):
    img = rng(1).integers(0, 256, (30, 50), dtype=np.uint8)
    big = torch.zeros((30, 64), dtype=torch.uint8, device=dev)
    big[:, :50] = dev_t(img, dev)
    out = torch.zeros((30, 70), dtype=torch.uint8, device=dev)
    tv.prefilter_gauss5(big[:, :50], out=out[:, :50])
    assert (out[:, :50].cpu().numpy() == oracle.gauss5(img)).all()
    assert (out[:, 50:] == 0).all()
    with pytest.raises(RuntimeError, match="in place"):
        tv.prefilter_gauss5(big[:, :50], out=big[:, :50])


@pytest.mark.parametrize("W,H", SIZES)
@pytest.mark.parametrize("it", [0, 1, 3])
def test_postfilter_open(tv, dev, W, H, it):
    d = rng(W + 3 * H + it).integers(-3, 256, (H, W)).astype(np.int16)
    got = tv.postfilter_open(dev_t(d, dev), iterations=it).cpu().numpy()
    assert (got == oracle.open_disp(d, it)).all()


def test_postfilter_in_place_and_large_radius(tv, dev):
    d = rng(2).integers(0, 100, (50, 70)).astype(np.int16)
    t = dev_t(d, dev)
    tv.postfilter_open(t, iterations=16, out=t)
    assert (t.cpu().numpy() == oracle.open_disp(d, 16)).all()


@pytest.mark.parametrize("ch,order", [(3, 0), (3, 1), (4, 0), (4, 1)])
def test_rgb_to_gray(tv, dev, ch, order):
    img = rng(ch * 10 + order).integers(0, 256, (41, 67, ch), dtype=np.uint8)
    got = tv.rgb_to_gray(dev_t(img, dev), order=order).cpu().numpy()
    assert (got == oracle.rgb_to_gray(img, order)).all()


@pytest.mark.parametrize("W,H,Wo,Ho", [(37, 23, 37, 23), (64, 48, 50, 70), (1280, 720, 1280, 720)])
def test_remap(tv, dev, W, H, Wo, Ho):
    img = rng(W + H).integers(0, 256, (H, W), dtype=np.uint8)
    yy, xx = np.mgrid[0:Ho, 0:Wo].astype(np.float64)
    # a rectification-like warp: small rotation, scale and shift, plus noise
    mx = (0.99 * xx + 0.02 * yy + 1.3 + rng(1).normal(0, 0.3, xx.shape)).astype(np.float32)
    my = (-0.015 * xx + 1.01 * yy - 2.1 + rng(2).normal(0, 0.3, xx.shape)).astype(np.float32)
    for border in (0, 200):
        got = tv.remap(dev_t(img, dev), dev_t(mx, dev), dev_t(my, dev), border=border)
        assert (got.cpu().numpy() == oracle.remap(img, mx, my, border)).all()


def test_remap_half_pixel_ties(tv, dev):
    img = rng(3).integers(0, 256, (16, 16), dtype=np.uint8)
    yy, xx = np.mgrid[0:16, 0:16].astype(np.float32)
    for off in (1 / 64, -1 / 64, 0.5, -0.5, 3 / 64):  # exact halves of 1/32 steps
        got = tv.remap(dev_t(img, dev), dev_t(xx + off, dev), dev_t(yy - off, dev))
        assert (got.cpu().numpy() == oracle.remap(img, xx + off, yy - off)).all()


@pytest.mark.parametrize("W,H", [(7, 5), (640, 480)])
def test_reproject(tv, dev, W, H):
    d = rng(W).integers(0, 128, (H, W)).astype(np.int16)
    f, cx, cy, Tx = 700.0, W / 2, H / 2, -0.12
    Q = np.array([[1, 0, 0, -cx], [0, 1, 0, -cy], [0, 0, 0, f], [0, 0, -1 / Tx, 0.3]],
                 np.float32)
    for ds, zs in ((1.0, 1.0), (1 / 128, 24.0)):
        got = tv.reproject(dev_t(d, dev), Q, disp_scale=ds, z_scale=zs).cpu().numpy()
        want = oracle.reproject(d, Q, disp_scale=ds, z_scale=zs)
        assert (got.view(np.uint32) == want.view(np.uint32)).all()


def test_pinned_round_trip_and_dsi_download(tv, dev):
    H, W, D = 20, 30, 16
    buf = tv.PinnedBuffer((H, W), np.uint8)
    try:
        buf.array[:] = rng(4).integers(0, 256, (H, W), dtype=np.uint8)
        t = torch.zeros((H, W + 5), dtype=torch.uint8, device=dev)
        tv.copy_2d(t.data_ptr(), t.stride(0), buf.array.ctypes.data, W, W, H)
        torch.cuda.synchronize()
        assert (t[:, :W].cpu().numpy() == buf.array).all() and (t[:, W:] == 0).all()
    finally:
        buf.free()
    vol = torch.from_numpy(rng(5).integers(0, 60000, (W, H, D)).astype(np.uint16)).to(dev)
    dsi = tv.DSI()
    lib = tv.lib()
    assert lib.tdvh_dsi_init(ct.byref(dsi), ct.c_void_p(vol.data_ptr()), W, H, D, 1, 0) == 0
    dsi.stride_x, dsi.stride_y = H * D, D  # the reference's [x][y][d] order
    host = np.empty((W, H, D), np.uint16)
    tv.dsi_download(dsi, host.ctypes.data)
    torch.cuda.synchronize()
    assert (host == vol.cpu().numpy()).all()
