"""Pins the oracle of the steps around the path (tdvref_io.c): closed forms,
known answers and independent restatements (scipy / numpy) that a wrong
kernel, rounding rule, border or channel order would fail."""
import numpy as np
import pytest
import scipy.ndimage as nd

import oracle

K5 = np.array([1, 4, 6, 4, 1])


def rng(seed):
    return np.random.default_rng(seed)


# ---- Gaussian prefilter (medianfiltercpu.cpp:7-19) ---------------------------

def test_gauss5_constant_and_impulse():
    img = np.full((9, 11), 77, np.uint8)
    assert (oracle.gauss5(img) == 77).all()
    img = np.zeros((9, 11), np.uint8)
    img[4, 5] = 255
    out = oracle.gauss5(img)
    want = (255 * np.outer(K5, K5) + 128) >> 8
    assert (out[2:7, 3:8] == want).all()
    assert out.sum() == want.sum()


def test_gauss5_matches_scipy_convolution_with_replicated_border():
    img = rng(1).integers(0, 256, (37, 53), dtype=np.uint8)
    s = nd.convolve(img.astype(np.float64), np.outer(K5, K5) / 256.0, mode="nearest")
    want = np.floor(s + 0.5).astype(np.uint8)  # s * 256 is an integer: exact
    assert (oracle.gauss5(img) == want).all()


def test_gauss5_mirror_symmetry():
    img = rng(2).integers(0, 256, (20, 31), dtype=np.uint8)
    assert (oracle.gauss5(img[:, ::-1]) == oracle.gauss5(img)[:, ::-1]).all()
    assert (oracle.gauss5(img[::-1]) == oracle.gauss5(img)[::-1]).all()


# ---- erode/dilate postfilter (dilate.cpp:12-22) ------------------------------

@pytest.mark.parametrize("it", [1, 2, 3])
def test_open_equals_grey_opening_with_square(it):
    d = rng(3 + it).integers(-5, 200, (23, 41)).astype(np.int16)
    size = 2 * it + 1
    want = nd.grey_dilation(nd.grey_erosion(d, size=(size, size), mode="nearest"),
                            size=(size, size), mode="nearest")
    assert (oracle.open_disp(d, it) == want).all()


def test_open_removes_small_speckles_keeps_large_regions():
    d = np.zeros((40, 40), np.int16)
    d[5:8, 5:8] = 50          # 3x3 speckle: removed by 3 iterations (7x7 opening)
    d[15:35, 15:35] = 30      # 20x20 block: kept exactly
    out = oracle.open_disp(d, 3)
    assert (out[5:8, 5:8] == 0).all()
    assert (out[15:35, 15:35] == 30).all() and out.sum() == 30 * 400


# ---- colour to gray (misc.cpp:10-16) -----------------------------------------

def test_gray_known_answers_and_order():
    px = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255]]], np.uint8)
    assert oracle.rgb_to_gray(px).tolist() == [[76, 150, 29, 255]]
    assert oracle.rgb_to_gray(px[:, :, ::-1].copy(), order=1).tolist() == [[76, 150, 29, 255]]
    v = np.arange(256, dtype=np.uint8)
    g = np.stack([v, v, v], -1)[None]
    assert (oracle.rgb_to_gray(g)[0] == v).all()  # the weights sum to 2^14


def test_gray_close_to_float_luma_and_alpha_ignored():
    img = rng(5).integers(0, 256, (13, 17, 4), dtype=np.uint8)
    f = 0.299 * img[..., 0] + 0.587 * img[..., 1] + 0.114 * img[..., 2]
    g = oracle.rgb_to_gray(img)
    assert np.abs(g - f).max() <= 0.51
    img2 = img.copy()
    img2[..., 3] = 255 - img2[..., 3]
    assert (oracle.rgb_to_gray(img2) == g).all()


# ---- rectification remap (rectificationcv.cpp:292-293) -----------------------

def test_remap_identity_shift_and_border():
    img = rng(6).integers(0, 256, (21, 30), dtype=np.uint8)
    y, x = np.mgrid[0:21, 0:30].astype(np.float32)
    assert (oracle.remap(img, x, y) == img).all()
    out = oracle.remap(img, x + 3, y - 2, border=9)
    assert (out[2:, :27] == img[:-2, 3:]).all()
    assert (out[:2] == 9).all() and (out[:, 27:] == 9).all()


def test_remap_half_pixel_is_rounded_mean_and_ramps_stay_linear():
    img = rng(7).integers(0, 256, (10, 12), dtype=np.uint8)
    y, x = np.mgrid[0:10, 0:11].astype(np.float32)
    out = oracle.remap(img, x + 0.5, y)
    want = (img[:, :11].astype(int) + img[:, 1:12] + 1) >> 1
    assert (out == want).all()
    ramp = np.tile(np.arange(0, 240, 8, dtype=np.uint8), (4, 1))  # 30 columns, step 8
    yy, xx = np.mgrid[0:4, 0:29].astype(np.float32)
    out = oracle.remap(ramp, xx + 0.25, yy)
    assert (out == (8 * np.arange(29) + 2)[None]).all()


def test_remap_matches_vectorised_restatement():
    img = rng(8).integers(0, 256, (17, 19), dtype=np.uint8)
    mx = rng(9).uniform(-2, 21, (15, 16)).astype(np.float32)
    my = rng(10).uniform(-2, 19, (15, 16)).astype(np.float32)
    fx = np.floor(mx.astype(np.float64) * 32 + 0.5).astype(np.int64)
    fy = np.floor(my.astype(np.float64) * 32 + 0.5).astype(np.int64)
    x0, y0, ax, ay = fx >> 5, fy >> 5, fx & 31, fy & 31
    pad = np.full((17 + 8, 19 + 8), 3, np.int64)
    pad[4:21, 4:23] = img

    def at(xx, yy):
        ok = (xx >= -4) & (xx < 23) & (yy >= -4) & (yy < 21)
        return np.where(ok, pad[np.clip(yy + 4, 0, 24), np.clip(xx + 4, 0, 26)], 3)

    s = ((32 - ax) * (32 - ay) * at(x0, y0) + ax * (32 - ay) * at(x0 + 1, y0) +
         (32 - ax) * ay * at(x0, y0 + 1) + ax * ay * at(x0 + 1, y0 + 1))
    # ties of the 1/32 rounding (exact halves) do not occur with these maps
    assert not np.any(np.abs(mx.astype(np.float64) * 32 % 1 - 0.5) < 1e-9)
    assert (oracle.remap(img, mx, my, border=3) == (s + 512) >> 10).all()


# ---- reprojection (cvreprojector.cpp:11-26) ----------------------------------

def test_reproject_identity_q():
    d = rng(11).integers(0, 64, (5, 7)).astype(np.int16)
    out = oracle.reproject(d, np.eye(4), disp_scale=0.5, z_scale=24.0)
    y, x = np.mgrid[0:5, 0:7]
    assert (out[..., 0] == x).all() and (out[..., 1] == 5 - y).all()
    assert (out[..., 2] == (d * np.float32(0.5)).astype(np.float32) * np.float32(24)).all()


def test_reproject_stereo_q_gives_depth_f_b_over_d():
    f, cx, cy, Tx = 500.0, 320.0, 240.0, -0.1   # Q of a rectified pair
    Q = np.array([[1, 0, 0, -cx], [0, 1, 0, -cy], [0, 0, 0, f], [0, 0, -1 / Tx, 0]])
    d = np.full((3, 4), 25, np.int16)
    d[0, 0] = 0   # w = 0: no point
    out = oracle.reproject(d, Q, disp_scale=1.0, z_scale=1.0)
    z = f / (25 / 0.1)
    assert np.allclose(out[1:, :, 2], z, rtol=1e-6)
    assert (out[0, 0] == 0).all()
    assert np.allclose(out[2, 3, 0], (3 - cx) / (25 / 0.1), rtol=1e-6)
