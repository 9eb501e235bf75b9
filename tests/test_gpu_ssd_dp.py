"""GPU parity of the fused SSD window + scanline DP (tdvh_stereo_ssd_dp):
the disparities of the oracle's SSD volume followed by its DP, bit-exact,
for every register shape (D up to 256), window radius 1..3, both reference
flags, several occlusion penalties, images narrower than the disparity
range and pitched images; the full 1920x1080 D = 256 workload against the
oracle; the widest row the LDS holds; argument validation."""
import numpy as np
import pytest
import torch

import oracle
from thundervision_amd import synth

pytestmark = pytest.mark.gpu


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def pair(W, H, dmax, seed):
    L, R, _ = synth.pair(W, H, dmax, 8, seed)
    return L, R


@pytest.mark.parametrize("W,H,D,r,flags,p_occ", [
    (40, 12, 16, 3, 0, 20), (97, 20, 64, 3, 0, 20), (64, 9, 200, 3, 0, 20),
    (70, 15, 33, 2, 1, 20), (70, 15, 33, 1, 2, 20), (70, 15, 33, 3, 3, 20),
    (1, 5, 8, 3, 0, 20), (2, 3, 8, 3, 0, 20), (3, 4, 5, 3, 1, 20), (6, 30, 4, 3, 0, 0),
    (130, 7, 256, 3, 0, 20), (300, 6, 100, 3, 0, 1000), (700, 4, 256, 3, 3, 20),
    (129, 11, 65, 3, 0, 0), (257, 9, 192, 2, 0, 7), (61, 8, 128, 1, 2, (1 << 20) - 1),
])
def test_fused_ssd_dp_matches_oracle(tv, dev, W, H, D, r, flags, p_occ):
    L, R = pair(W, H, max(D - 1, 1), 13 * W + D + r)
    want = oracle.dp(oracle.cost_ssd(L, R, D, r), p_occ, flags)
    got = tv.stereo_ssd_dp(to_dev(L, dev), to_dev(R, dev), D, radius=r, p_occ=p_occ,
                           flags=flags).cpu().numpy()
    assert (got == want).all(), int((got != want).sum())


def test_fused_ssd_dp_pitched_images(tv, dev):
    """Images with a row pitch above the width, read in place."""
    W, H, D, pitch = 75, 19, 48, 96
    L, R = pair(W, H, D - 1, 77)
    want = oracle.dp(oracle.cost_ssd(L, R, D, 3), 20)
    Lp = np.zeros((H, pitch), np.uint8)
    Rp = np.zeros((H, pitch), np.uint8)
    Lp[:, :W], Rp[:, :W] = L, R
    Ld, Rd = to_dev(Lp, dev)[:, :W], to_dev(Rp, dev)[:, :W]
    got = tv.stereo_ssd_dp(Ld, Rd, D).cpu().numpy()
    assert (got == want).all()


def test_fused_ssd_dp_window_1080p(tv, dev):
    """1920x1080, D = 256 (the window workload): the fused path equals the
    oracle's SSD volume -> DP and the device's volume path."""
    L, R, _ = synth.workload_pair(synth.WINDOW_1080P)
    want = oracle.dp(oracle.cost_ssd(L, R, 256, 3), 20)
    Ld, Rd = to_dev(L, dev), to_dev(R, dev)
    got = tv.stereo_ssd_dp(Ld, Rd, 256).cpu().numpy()
    assert (got == want).all(), int((got != want).sum())
    vol = tv.dp(tv.cost_ssd(Ld, Rd, 256, 3), 20, tv.DP_COST_BOUNDED).cpu().numpy()
    assert (vol == got).all()


def test_fused_ssd_dp_repeat_and_reuse(tv, dev):
    """Two shapes on one context in turn (the decision scratch is shared and
    grows), each call equal to the first of its shape."""
    ctx = tv.Context(0)
    a = pair(200, 30, 63, 5)
    b = pair(90, 50, 127, 6)
    da = [tv.stereo_ssd_dp(to_dev(a[0], dev), to_dev(a[1], dev), 64, ctx=ctx).cpu().numpy()
          for _ in range(2)]
    db = tv.stereo_ssd_dp(to_dev(b[0], dev), to_dev(b[1], dev), 128, ctx=ctx).cpu().numpy()
    da.append(tv.stereo_ssd_dp(to_dev(a[0], dev), to_dev(a[1], dev), 64, ctx=ctx).cpu().numpy())
    assert (da[0] == da[1]).all() and (da[0] == da[2]).all()
    assert (db == oracle.dp(oracle.cost_ssd(b[0], b[1], 128, 3), 20)).all()


@pytest.mark.parametrize("kw,msg", [
    (dict(depth=257), "depth"), (dict(depth=0), "depth"), (dict(radius=4), "radius"),
    (dict(radius=0), "radius"), (dict(p_occ=1 << 20), "p_occ"), (dict(flags=4), "flags"),
])
def test_fused_ssd_dp_rejects_bad_arguments(tv, dev, kw, msg):
    L, R = pair(40, 8, 15, 3)
    args = dict(depth=16, radius=3, p_occ=20, flags=0)
    args.update(kw)
    with pytest.raises(tv.TdvhError, match=msg):
        tv.stereo_ssd_dp(to_dev(L, dev), to_dev(R, dev), args["depth"], radius=args["radius"],
                         p_occ=args["p_occ"], flags=args["flags"])
    # the context still works
    got = tv.stereo_ssd_dp(to_dev(L, dev), to_dev(R, dev), 16).cpu().numpy()
    assert (got == oracle.dp(oracle.cost_ssd(L, R, 16, 3), 20)).all()


def test_fused_ssd_dp_widest_image(tv, dev):
    """The widest accepted row (10236 columns: 160 KiB of staged columns
    with radius 3, a workgroup's LDS) against the oracle, and the next width
    rejected."""
    L, R = pair(10236, 3, 31, 9)
    want = oracle.dp(oracle.cost_ssd(L, R, 32, 3), 20)
    got = tv.stereo_ssd_dp(to_dev(L, dev), to_dev(R, dev), 32).cpu().numpy()
    assert (got == want).all()
    W = torch.zeros((2, 10237), dtype=torch.uint8, device=dev)
    with pytest.raises(tv.TdvhError, match="width"):
        tv.stereo_ssd_dp(W, W, 16)
