"""Seeded random sweeps against the oracle, bit-exact: the per-pair BT +
SGM-8 + WTA call (image sizes 1..300, disparity ranges 1..256, so both the
scans that recompute BT and the ones over relative cost bytes run at every
register shape, penalties 0 <= P1 <= P2 <= 255, fused and staged), and SSD
-> DP (widths 1..400, D 1..512, radius 1..3, both reference flags, several
occlusion penalties; the volume DP with 32- and 64-bit sums, and the fused
SSD + DP for D <= 256), and SGM over the device's SSD and NCC volumes
(penalties on both sides of the relative-cost-byte limit, direction
masks)."""
import numpy as np
import pytest
import torch

import oracle
from thundervision_amd import synth

pytestmark = pytest.mark.gpu


def cases(n=16, seed=20261204):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W = int(rng.integers(1, 301))
        H = int(rng.integers(1, 301))
        D = int(rng.integers(1, 257))
        p1 = int(rng.integers(0, 41))
        p2 = int(rng.integers(p1, 256))
        out.append((W, H, D, p1, p2, 500 + i))
    return out


@pytest.mark.parametrize("W,H,D,p1,p2,seed", cases())
def test_random_shapes_match_oracle(tv, dev, W, H, D, p1, p2, seed):
    L, R, _ = synth.pair(W, H, max(D - 1, 0), 6, seed)
    want = oracle.stereo_bt_sgm(L, R, D, p1, p2)
    Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
    for fusion in (tv.FUSE_COST, tv.FUSE_STAGED):
        got = tv.stereo_bt_sgm(Ld, Rd, D, p1=p1, p2=p2, fusion=fusion).cpu().numpy()
        assert (got == want).all(), (fusion, int((got != want).sum()))


def dp_cases(n=12, seed=4242):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W = int(rng.integers(1, 401))
        H = int(rng.integers(1, 41))
        D = int(rng.integers(1, 513))
        r = int(rng.integers(1, 4))
        flags = int(rng.integers(0, 4))
        p_occ = int(rng.choice([0, 1, 20, 300, 5000]))
        out.append((W, H, D, r, flags, p_occ, 700 + i))
    return out


@pytest.mark.parametrize("W,H,D,r,flags,p_occ,seed", dp_cases())
def test_random_ssd_dp_match_oracle(tv, dev, W, H, D, r, flags, p_occ, seed):
    """SSD -> DP over the volume (32-bit bounded sums and 64-bit sums) and,
    for D <= 256, the fused SSD + DP, against the oracle's SSD volume and DP."""
    L, R, _ = synth.pair(W, H, max(D - 1, 0), 5, seed)
    want = oracle.dp(oracle.cost_ssd(L, R, D, r), p_occ, flags)
    Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
    vol = tv.cost_ssd(Ld, Rd, D, r)
    for f in (flags | tv.DP_COST_BOUNDED, flags):
        got = tv.dp(vol, p_occ, f).cpu().numpy()
        assert (got == want).all(), (f, int((got != want).sum()))
    if D <= 256:
        got = tv.stereo_ssd_dp(Ld, Rd, D, radius=r, p_occ=p_occ, flags=flags).cpu().numpy()
        assert (got == want).all(), ("fused", int((got != want).sum()))


def sgm_volume_cases(n=10, seed=777):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W = int(rng.integers(2, 121))
        H = int(rng.integers(2, 61))
        D = int(rng.integers(1, 257))
        p1 = int(rng.integers(0, 30))
        p2 = int(rng.integers(p1, 400))  # <= 127: relative cost bytes; above: uint32 scans
        mask = int(rng.choice([0xFF, 0xFF, 0x0F, 0xF0, 0x5A, 0x81]))
        out.append((W, H, D, p1, p2, mask, 900 + i))
    return out


@pytest.mark.parametrize("W,H,D,p1,p2,mask,seed", sgm_volume_cases())
def test_random_sgm_over_ssd_and_ncc_volumes(tv, dev, W, H, D, p1, p2, mask, seed):
    """SGM over the device's SSD (uint32) and NCC (float) volumes: S and the
    disparities (with S, and disparities only) against the oracle."""
    L, R, _ = synth.pair(W, H, max(D - 1, 0), 5, seed)
    Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
    C = tv.cost_ssd(Ld, Rd, D, 2)
    want = oracle.sgm_u32(C.cpu().numpy().view(np.uint32), p1, p2, mask)
    disp, S = tv.sgm(C, p1=p1, p2=p2, dir_mask=mask, return_S=True)
    assert (S.cpu().numpy().view(np.uint32) == want).all()
    assert (disp.cpu().numpy() == oracle.wta(want)).all()
    assert (tv.sgm(C, p1=p1, p2=p2, dir_mask=mask).cpu().numpy() == oracle.wta(want)).all()
    Cf = tv.cost_ncc(Ld, Rd, D, 2, True)
    fp1, fp2 = np.float32(p1 / 255), np.float32(p2 / 255)
    wantf = oracle.sgm_f32(Cf.cpu().numpy(), fp1, fp2, mask)
    disp, S = tv.sgm(Cf, dir_mask=mask, fp1=float(fp1), fp2=float(fp2), return_S=True)
    assert np.array_equal(S.cpu().numpy().view(np.uint32), wantf.view(np.uint32))
    assert (disp.cpu().numpy() == oracle.wta(wantf)).all()
    assert (tv.sgm(Cf, dir_mask=mask, fp1=float(fp1), fp2=float(fp2)).cpu().numpy() ==
            oracle.wta(wantf)).all()
