"""GPU parity of the scanline DP's two decision word layouts: rows
interleaved by word block (the default) and row-major (taken where the
interleaved byte offsets would pass 32 bits). The volume DP (32- and 64-bit
sums, and the 64-bit recomputation of a row that breaks the bounded
contract) and the fused SSD + DP run on the smallest images past that bound
and just below it; sampled rows are checked against the oracle (a row's DP
depends on its own costs only, and its SSD costs on the 7 image rows around
it)."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def images(W, H, seed):
    rng = np.random.default_rng(seed)
    L = rng.integers(0, 256, (H, W), dtype=np.uint8)
    R = np.roll(L, -3, axis=1) ^ rng.integers(0, 8, (H, W), dtype=np.uint8)
    return L, R


def interleaved(W, H):
    """The library's choice for D <= 256 (4 columns per word): interleaved
    when the last row's words plus the backtrack's 32-block over-read stay
    below 2^32 bytes."""
    return ((W + 3) // 4 + 32) * H * 256 < 1 << 32


def sample_rows(H):
    return [0, 1, 2, 3, H // 3, H // 2, H - 4, H - 3, H - 2, H - 1]


@pytest.mark.parametrize("H", [16200, 16300])
def test_volume_dp_word_layouts(tv, dev, H):
    W, D = 4000, 4
    assert interleaved(W, H) == (H == 16200)
    L, R = images(W, H, H)
    vol = tv.cost_ssd(torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev), D, 3)
    # a cost above the bounded contract in one row: that row is recomputed
    # with 64-bit sums by the second launch
    bad_row = H // 2
    vol[bad_row, 100, 2] = 1 << 30
    got32 = tv.dp(vol, 20, tv.DP_COST_BOUNDED).cpu().numpy()
    got64 = tv.dp(vol, 20, 0).cpu().numpy()
    for y in sample_rows(H):
        want = oracle.dp(vol[y:y + 1].cpu().numpy().view(np.uint32), 20)[0]
        assert (got32[y] == want).all(), y
        assert (got64[y] == want).all(), y
    del vol


@pytest.mark.parametrize("H", [6470, 6480])
def test_fused_ssd_dp_word_layouts(tv, dev, H):
    W, D = 10236, 16
    assert interleaved(W, H) == (H == 6470)
    L, R = images(W, H, H + 1)
    got = tv.stereo_ssd_dp(torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev),
                           D).cpu().numpy()
    for y in sample_rows(H):
        y0, y1 = max(0, y - 3), min(H, y + 4)
        want = oracle.dp(oracle.cost_ssd(L[y0:y1], R[y0:y1], D, 3), 20)[y - y0]
        assert (got[y] == want).all(), y
