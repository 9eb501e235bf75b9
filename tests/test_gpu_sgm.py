"""GPU parity of the BT cost, SGM and WTA kernels against the CPU oracle
(bit-exact). Calls go through the C ABI (libtdvh.so via ctypes)."""
import numpy as np
import pytest
import torch

import oracle
from thundervision_amd import synth

pytestmark = pytest.mark.gpu


def u16(t):
    return t.cpu().numpy().view(np.uint16)


def pair(W, H, dmax, seed, nrect=6):
    L, R, _ = synth.pair(W, H, dmax, nrect, seed)
    return L, R


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


SHAPES = [(37, 23, 16), (64, 48, 32), (50, 61, 48), (97, 40, 80), (130, 70, 128),
          (300, 33, 160), (271, 19, 256), (16, 16, 5), (7, 5, 1)]


@pytest.mark.parametrize("W,H,D", SHAPES)
def test_bt_cost_strided_and_split(tv, dev, W, H, D):
    L, R = pair(W, H, max(D - 1, 0), W * 7 + H)
    want = oracle.cost_bt2(L, R, D)
    Ld, Rd = to_dev(L, dev), to_dev(R, dev)
    got = tv.cost_bt(Ld, Rd, D)
    assert (u16(got) == want).all()
    if D <= tv.MAX_DEPTH:
        sp = tv.cost_bt(Ld, Rd, D, layout=tv.LAYOUT_SPLIT16)
        back = torch.empty((H, W, D), dtype=torch.int16, device=dev)
        src = tv.dsi_for(sp, D, tv.U16, tv.LAYOUT_SPLIT16)
        dst = tv.dsi_for(back, D, tv.U16)
        import ctypes as ct
        tv._check(tv.lib().tdvh_dsi_convert(tv._ctx(None), ct.byref(src), ct.byref(dst),
                                            tv._stream()))
        assert (u16(back) == want).all()


def test_bt_cost_reference_layout(tv, dev):
    # the reference's [x][y][d] layout through explicit strides
    W, H, D = 45, 27, 20
    L, R = pair(W, H, 19, 5)
    want = oracle.cost_bt2(L, R, D)
    out = torch.empty((W, H, D), dtype=torch.int16, device=dev)
    import ctypes as ct
    d = tv.DSI()
    tv._check(tv.lib().tdvh_dsi_init(ct.byref(d), tv._ptr(out), W, H, D, tv.U16, 0))
    d.stride_x, d.stride_y = H * D, D
    Ld, Rd = to_dev(L, dev), to_dev(R, dev)
    tv._check(tv.lib().tdvh_cost_bt(tv._ctx(None), tv._ptr(Ld), tv._ptr(Rd), W, H, W,
                                    ct.byref(d), tv._stream()))
    assert (u16(out).transpose(1, 0, 2) == want).all()


@pytest.mark.parametrize("W,H,D", [s for s in SHAPES if s[2] >= 1])
def test_sgm_std_volume_and_disparity(tv, dev, W, H, D):
    L, R = pair(W, H, max(D - 1, 0), W + 3 * H)
    C = oracle.cost_bt2(L, R, D)
    S_want = oracle.sgm_std(C, 8, 32)
    d_want = oracle.wta(S_want)
    Cd = to_dev(C.view(np.int16), dev)
    disp, S = tv.sgm(Cd, p1=8, p2=32, return_S=True)
    assert (u16(S) == S_want).all()
    assert (disp.cpu().numpy() == d_want).all()


@pytest.mark.parametrize("mask", [0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x0F, 0xF0])
def test_sgm_direction_masks(tv, dev, mask):
    W, H, D = 41, 29, 24
    L, R = pair(W, H, 20, 77)
    C = oracle.cost_bt2(L, R, D)
    S_want = oracle.sgm_std(C, 8, 32, mask)
    _, S = tv.sgm(to_dev(C.view(np.int16), dev), dir_mask=mask, return_S=True)
    assert (u16(S) == S_want).all()


@pytest.mark.parametrize("p1,p2", [(0, 0), (8, 32), (3, 100), (30, 150), (8, 255), (8, 300), (40, 1000)])
def test_sgm_penalties(tv, dev, p1, p2):
    W, H, D = 60, 31, 40
    L, R = pair(W, H, 39, 11)
    C = oracle.cost_bt2(L, R, D)
    _, S = tv.sgm(to_dev(C.view(np.int16), dev), p1=p1, p2=p2, return_S=True)
    assert (u16(S) == oracle.sgm_std(C, p1, p2)).all()


@pytest.mark.parametrize("fusion", [0, 1])
def test_stereo_plain_384(tv, dev, fusion):
    L, R, _ = synth.workload_pair(synth.PLAIN_384)
    want = oracle.stereo_bt_sgm(L, R, 16)
    got = tv.stereo_bt_sgm(to_dev(L, dev), to_dev(R, dev), 16, fusion=fusion)
    assert (got.cpu().numpy() == want).all()


def test_stereo_plain_720p_fused_staged_and_oracle(tv, dev):
    L, R, gt = synth.workload_pair(synth.PLAIN_720P)
    want = oracle.stereo_bt_sgm(L, R, 128)
    Ld, Rd = to_dev(L, dev), to_dev(R, dev)
    a = tv.stereo_bt_sgm(Ld, Rd, 128, fusion=tv.FUSE_COST).cpu().numpy()
    b = tv.stereo_bt_sgm(Ld, Rd, 128, fusion=tv.FUSE_STAGED).cpu().numpy()
    assert (a == want).all()
    assert (b == want).all()
    assert (a == gt).mean() > 0.8


def test_repeat_runs_are_identical(tv, dev):
    L, R = pair(200, 120, 63, 9)
    Ld, Rd = to_dev(L, dev), to_dev(R, dev)
    a = tv.stereo_bt_sgm(Ld, Rd, 64).cpu().numpy()
    for _ in range(3):
        assert (tv.stereo_bt_sgm(Ld, Rd, 64).cpu().numpy() == a).all()


@pytest.mark.parametrize("W,H,D", [(32, 32, 8), (64, 48, 16), (48, 64, 12), (70, 30, 33)])
@pytest.mark.parametrize("flags", [1, 1 | 2, 1 | 2 | 4, 1 | 4])
def test_sgm_reference_geometry(tv, dev, W, H, D, flags):
    L, R = pair(W, H, D - 1, 5 * W + H)
    C = oracle.cost_bt2(L, R, D)
    oflags = (oracle.SGM_START_ASSIGN if flags & 2 else 0) | \
             (oracle.SGM_MIN_SKIP_LAST if flags & 4 else 0)
    S_want = oracle.sgm_ref(C, 8, 32, oflags)
    disp, S = tv.sgm(to_dev(C.view(np.int16), dev), flags=flags, return_S=True)
    assert (S.cpu().numpy() == S_want).all()
    assert (disp.cpu().numpy() == oracle.wta(S_want)).all()


@pytest.mark.parametrize("dtype", ["u16", "u32", "i32", "f32"])
@pytest.mark.parametrize("D", [37, 64, 200])  # D % 4 == 0: 4-cell vector loads per lane
def test_wta_lowest(tv, dev, dtype, D):
    rng = np.random.default_rng(3 + D)
    H, W = 13, 17
    if dtype == "u16":
        S = rng.integers(0, 20, (H, W, D)).astype(np.uint16)
        S[rng.random(S.shape) < 0.1] = 0xFFFF
        t = to_dev(S.view(np.int16), dev)
        got = tv.wta(t)
    elif dtype == "u32":
        S = rng.integers(0, 20, (H, W, D)).astype(np.uint32)
        S[rng.random(S.shape) < 0.1] = 0xFFFFFFFF
        got = tv.wta(to_dev(S.view(np.int32), dev))
    elif dtype == "i32":
        S = rng.integers(-20, 20, (H, W, D)).astype(np.int32)
        S[rng.random(S.shape) < 0.1] = oracle.INF_I32
        t = to_dev(S, dev)
        got = tv.wta_dsi(t, tv.dsi_for(t, D, tv.I32))
    else:
        S = rng.integers(0, 20, (H, W, D)).astype(np.float32) / 4
        S[rng.random(S.shape) < 0.1] = np.inf
        got = tv.wta(to_dev(S, dev))
    assert (got.cpu().numpy() == oracle.wta(S)).all()


def test_wta_tree_mode(tv, dev):
    rng = np.random.default_rng(5)
    S = rng.integers(0, 6, (9, 11, 32)).astype(np.uint16)
    got = tv.wta(to_dev(S.view(np.int16), dev), mode=tv.WTA_TREE)
    assert (got.cpu().numpy() == oracle.wta_tree(S)).all()


def test_unit_float_views(tv, dev):
    disp = torch.arange(0, 40, dtype=torch.int16, device=dev).reshape(4, 10)
    f = tv.disp_to_unit_float(disp, 64).cpu().numpy()
    assert np.allclose(f, np.arange(40).reshape(4, 10) / 64.0)
    img = torch.rand(5, 7, device=dev)
    u = tv.unit_float_to_u8(img).cpu().numpy()
    assert (u == np.clip(np.rint(img.cpu().numpy() * 255), 0, 255).astype(np.uint8)).all()


def test_errors_are_reported(tv, dev):
    L = torch.zeros((8, 8), dtype=torch.uint8, device=dev)
    with pytest.raises(tv.TdvhError, match="depth"):
        tv.stereo_bt_sgm(L, L, 300)


@pytest.mark.parametrize("fusion", [0, 1])
@pytest.mark.parametrize("W,H,D,p2", [(97, 40, 80, 32), (65, 33, 33, 300), (300, 20, 256, 32),
                                      (40, 70, 160, 7), (31, 17, 1, 32),
                                      # long sides scanned 64 lanes per line (m = 4, 8;
                                      # horizontal and vertical; with padding cells)
                                      (200, 64, 100, 32), (50, 200, 128, 32),
                                      (33, 300, 250, 20), (260, 90, 256, 32),
                                      # odd word counts (3, 5, 7) scanned one word wider
                                      (96, 40, 96, 32), (170, 45, 150, 32),
                                      (120, 50, 200, 32), (230, 30, 224, 32)])
def test_stereo_shapes_and_paths(tv, dev, fusion, W, H, D, p2):
    L, R = pair(W, H, max(D - 1, 0), 3 * W + 7 * H)
    want = oracle.wta(oracle.sgm_std(oracle.cost_bt2(L, R, D), 8, p2))
    got = tv.stereo_bt_sgm(to_dev(L, dev), to_dev(R, dev), D, p2=p2, fusion=fusion)
    assert (got.cpu().numpy() == want).all()


def test_stereo_pitched_images(tv, dev):
    W, H, D = 77, 41, 48
    L, R = pair(W, H, D - 1, 99)
    big = torch.zeros((2, H, W + 13), dtype=torch.uint8, device=dev)
    big[0, :, :W] = to_dev(L, dev)
    big[1, :, :W] = to_dev(R, dev)
    want = oracle.stereo_bt_sgm(L, R, D)
    for fusion in (0, 1):
        got = tv.stereo_bt_sgm(big[0, :, :W], big[1, :, :W], D, fusion=fusion)
        assert (got.cpu().numpy() == want).all()


@pytest.mark.parametrize("D", [64, 40])
def test_stereo_1080p_two_word_lanes(tv, dev, D):
    """1920x1080 at 2 words per lane: the line scans run 8-lane lines (8 per
    wave, 4 words per lane); D = 40 adds padding cells."""
    L, R, _ = synth.pair(1920, 1080, D - 1, 12, 7 + D)
    want = oracle.stereo_bt_sgm(L, R, D)
    got = tv.stereo_bt_sgm(to_dev(L, dev), to_dev(R, dev), D)
    assert (got.cpu().numpy() == want).all()
