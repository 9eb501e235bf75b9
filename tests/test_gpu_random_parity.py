"""Seeded random sweeps against the oracle, bit-exact: the per-pair BT +
SGM-8 + WTA call (image sizes 1..300, disparity ranges 1..256, so both the
scans that recompute BT and the ones over relative cost bytes run at every
register shape, penalties 0 <= P1 <= P2 <= 255, fused and staged), and SSD
-> DP (widths 1..400, D 1..512, radius 1..3, both reference flags, several
occlusion penalties; the volume DP with 32- and 64-bit sums, and the fused
SSD + DP for D <= 256)."""
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
