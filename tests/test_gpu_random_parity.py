"""Seeded random sweep of the per-pair BT + SGM-8 + WTA call against the
oracle: image sizes 1..300, disparity ranges 1..256 (so both the scans that
recompute BT and the ones over relative cost bytes run, at every register
shape), penalties 0 <= P1 <= P2 <= 255, fused and staged; bit-exact."""
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
