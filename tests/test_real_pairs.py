"""Real image pairs the reference ships (res/tsukuba_*.png and the Middlebury
Teddy quarter-size pair with ground truth, decoded into tests/golden/
real_pairs.npz by tests/golden/make_fixtures.py). CPU: the oracle pipeline is
sane on them; GPU: the device path equals the oracle bit-exactly."""
import os

import numpy as np
import pytest

import oracle

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "real_pairs.npz")


def pairs():
    z = np.load(FIX, allow_pickle=False)
    return {"tsukuba": (z["tsukuba_left"], z["tsukuba_right"], 16, None),
            "teddy": (z["teddy_left"], z["teddy_right"], 64, z["teddy_true"])}


def teddy_rate(d, gt):
    # the 512x512 images stretch the 450-pixel-wide originals; ground truth
    # is stored as 4x disparity
    valid = gt > 0
    return (np.abs(d - gt.astype(float) * 512 / 450 / 4)[valid] <= 2).mean()


def test_oracle_is_sane_on_real_pairs():
    P = pairs()
    L, R, D, gt = P["teddy"]
    Lg, Rg = oracle.gauss5(oracle.rgb_to_gray(L)), oracle.gauss5(oracle.rgb_to_gray(R))
    assert teddy_rate(oracle.stereo_bt_sgm(Lg, Rg, D), gt) > 0.6
    # swapping the images (wrong matching direction) is far worse
    assert teddy_rate(oracle.stereo_bt_sgm(Rg, Lg, D), gt) < 0.3


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tsukuba", "teddy"])
def test_device_path_equals_oracle_on_real_pairs(tv, dev, name):
    import torch
    L, R, D, gt = pairs()[name]
    Lg, Rg = oracle.gauss5(oracle.rgb_to_gray(L)), oracle.gauss5(oracle.rgb_to_gray(R))
    want = oracle.stereo_bt_sgm(Lg, Rg, D)
    Ld = tv.prefilter_gauss5(tv.rgb_to_gray(torch.from_numpy(L).to(dev)))
    Rd = tv.prefilter_gauss5(tv.rgb_to_gray(torch.from_numpy(R).to(dev)))
    for fusion in (tv.FUSE_COST, tv.FUSE_STAGED):
        got = tv.stereo_bt_sgm(Ld, Rd, D, fusion=fusion).cpu().numpy()
        assert (got == want).all()
    H, W = Lg.shape
    p = tv.Pipeline(W, H, D, flags=tv.PIPE_PREFILTER | tv.PIPE_POSTFILTER)
    try:
        got = p.run(tv.rgb_to_gray(torch.from_numpy(L).to(dev)),
                    tv.rgb_to_gray(torch.from_numpy(R).to(dev))).cpu().numpy()
    finally:
        p.close()
    assert (got == oracle.open_disp(want, 3)).all()
    if gt is not None:
        assert teddy_rate(got, gt) > 0.6
